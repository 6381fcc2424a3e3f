// Decoder driver of the C ABI: Decoder::decode_frame of reitero_decode
// (reitero_decode/src/decoder.rs). Entropy decoding and MV reconstruction run on the host;
// motion compensation, dequantisation, IDCT, reconstruction, deblock and YUV -> RGB on the GPU.
#include <atomic>
#include <chrono>
#include <hip/hip_runtime.h>
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>
#include "host_common.hpp"
#include "riv_kernels.hpp"
#include "entropy.hpp"

using namespace rivhost;

// One frame's host stage: the parsed record, its vectors and masks (inter), and its quantised
// coefficients in a pinned buffer, ready for upload.
struct HostFrame {
    int status = RIV_OK;
    std::string err;
    uint64_t ts = 0;
    uint8_t ft = 0, q = 0;
    std::vector<Mv4> mv;
    std::vector<uint8_t> skip, coded, mask, payload;
    uint16_t yt[256], ct[64];
    HostBuf coeffs;   // nb * 384 int16
};

// Look-ahead of an in-memory stream: one thread parses the records in order and decodes their
// vectors (the MV contexts and the previous frame's vectors carry from frame to frame); worker
// threads decode each frame's residuals (fresh contexts per frame) into its slot. decode_frame
// takes the slots in order. Slot states: 0 free, 1 being filled, 2 ready.
struct DecodePipe {
    static constexpr int kSlots = 8;
    std::thread producer;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<int> jobs;
    bool stop = false;
    HostFrame slots[kSlots];
    int state[kSlots] = {};
};

struct riv_decoder {
    riv_reader_vtbl rv{};
    void* ruser = nullptr;
    bool mem = false;
    std::vector<uint8_t> membuf;
    size_t mempos = 0;
    riv_video_header hdr{};
    uint64_t current = 0;
    bool skip_residuals = false;
    int dev = 0;
    int W = 0, H = 0, bw = 0, bh = 0, nb = 0;
    hipStream_t st = nullptr;
    DevBuf ref[2], pred, coeffs, mvs, skip, mask, var, lnv, steps, rgb, yt, ct;
    HostBuf h_rgb;
    int ref_idx = 0;
    bool have_ref = false, have_prev_mvs = false;
    std::vector<Mv4> prev_mvs;
    rivh::MvCoder mvc;
    uint8_t last_yt_q = 0;
    bool parsed_intra = false;        // the host stage has seen the first (intra) frame
    DecodePipe* pipe = nullptr;       // in-memory streams, once decoding starts
    HostFrame inline_frame;           // reader streams: the one frame in flight
    // phase timings (ns since the last drain); host phases may run on the look-ahead threads
    std::atomic<uint64_t> t_read{0}, t_parse{0}, t_mv{0}, t_rans{0}, t_pred{0}, t_apply{0}, t_rgb{0};
    bool profiling = false;
    hipEvent_t ev[4] = {};            // device phases: before MC, after MC, after the reconstruction, after RGB out

    // read_exact (decoder.rs:502-511): a short read is EndOfStream
    bool read_exact(uint8_t* dst, size_t n) {
        const auto t0 = std::chrono::steady_clock::now();
        const bool ok = read_exact_(dst, n);
        t_read += (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
        return ok;
    }
    bool read_exact_(uint8_t* dst, size_t n) {
        if (mem) {
            if (mempos + n > membuf.size()) return false;
            std::memcpy(dst, membuf.data() + mempos, n);
            mempos += n;
            return true;
        }
        while (n) {
            size_t got = rv.read(ruser, dst, n);
            if (got == 0) return false;
            dst += got; n -= got;
        }
        return true;
    }
};

extern "C" {

static int decoder_init(riv_decoder* d) {
    uint8_t h[36];
    if (!d->read_exact(h, 36)) { set_error("truncated header"); return RIV_E_END_OF_STREAM; }
    if (!(h[0] == 'R' && h[1] == 'I' && h[2] == 'V' && h[3] == 0)) { set_error("Invalid magic bytes"); return RIV_E_INVALID_HEADER; }
    if (get_u32(h + 4) != 5) { set_error("Unsupported version"); return RIV_E_INVALID_HEADER; }
    d->hdr.display_width = get_u32(h + 8); d->hdr.display_height = get_u32(h + 12);
    d->hdr.storage_width = get_u32(h + 16); d->hdr.storage_height = get_u32(h + 20);
    d->hdr.fps = get_u32(h + 24); d->hdr.frame_count = get_u64(h + 28);
    if (d->hdr.storage_width % 16 || d->hdr.storage_height % 16 || d->hdr.storage_width == 0 || d->hdr.storage_height == 0 ||
        d->hdr.display_width > d->hdr.storage_width || d->hdr.display_height > d->hdr.storage_height) {
        set_error("storage dims must be non-zero multiples of 16 covering the display");
        return RIV_E_INVALID_HEADER;
    }
    int r = ensure_device(d->dev);
    if (r) return r;
    d->W = (int)d->hdr.storage_width; d->H = (int)d->hdr.storage_height;
    d->bw = d->W / 16; d->bh = d->H / 16; d->nb = d->bw * d->bh;
    const size_t ysz = (size_t)d->W * d->H, fsz = ysz + ysz / 2;
    RIV_HIP(hipStreamCreateWithFlags(&d->st, hipStreamNonBlocking));
    RIV_HIP(d->ref[0].alloc(fsz)); RIV_HIP(d->ref[1].alloc(fsz)); RIV_HIP(d->pred.alloc(fsz));
    RIV_HIP(d->coeffs.alloc(768 * (size_t)d->nb)); RIV_HIP(d->mvs.alloc(4 * (size_t)d->nb));
    RIV_HIP(d->skip.alloc(d->nb)); RIV_HIP(d->mask.alloc(d->nb));
    RIV_HIP(d->var.alloc(8 * (size_t)d->nb)); RIV_HIP(d->lnv.alloc(4 * ((size_t)d->nb + 1))); RIV_HIP(d->steps.alloc(4 * (size_t)d->nb));
    size_t rgbsz = (size_t)d->hdr.display_width * d->hdr.display_height * 3;
    RIV_HIP(d->rgb.alloc(rgbsz)); RIV_HIP(d->h_rgb.alloc(rgbsz));
    RIV_HIP(d->yt.alloc(512)); RIV_HIP(d->ct.alloc(128));
    return RIV_OK;
}

int riv_decoder_create(const riv_reader_vtbl* reader, void* user, int32_t device_id, riv_decoder** out) {
    if (!reader || !reader->read || !out) return RIV_E_ARG;
    riv_decoder* d = new riv_decoder();
    d->rv = *reader; d->ruser = user; d->dev = device_id;
    int r = decoder_init(d);
    if (r) { delete d; return r; }
    *out = d;
    return RIV_OK;
}
int riv_decoder_create_mem(const uint8_t* data, size_t len, int32_t device_id, riv_decoder** out) {
    if ((!data && len) || !out) return RIV_E_ARG;
    riv_decoder* d = new riv_decoder();
    d->mem = true; d->membuf.assign(data, data + len); d->dev = device_id;
    int r = decoder_init(d);
    if (r) { delete d; return r; }
    *out = d;
    return RIV_OK;
}

// decode_next (decoder.rs:195-459), in two stages.
// Host stage, part 1 (frame order): reads the next record, decodes and reconstructs its vectors
// (decoder.rs:195-382). The residual payload is kept for part 2.
static uint64_t ns_since(std::chrono::steady_clock::time_point t0) {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count();
}

static int parse_frame_(riv_decoder* d, HostFrame& f);
// The record and its vectors; parse_frame time excludes the reads and the MV rANS.
static int parse_frame(riv_decoder* d, HostFrame& f) {
    const auto t0 = std::chrono::steady_clock::now();
    const uint64_t r0 = d->t_read.load(), m0 = d->t_mv.load();
    const int r = parse_frame_(d, f);
    const uint64_t all = ns_since(t0), inner = (d->t_read.load() - r0) + (d->t_mv.load() - m0);
    d->t_parse += all > inner ? all - inner : 0;
    return r;
}

static int parse_frame_(riv_decoder* d, HostFrame& f) {
    uint8_t head[9];
    if (!d->read_exact(head, 9)) { set_error("end of stream"); return RIV_E_END_OF_STREAM; }
    f.ts = get_u64(head);
    f.ft = head[8];
    if (f.ft != 1 && f.ft != 2) { set_error("Bad frame_type"); return RIV_E_INVALID_FRAME; }
    const int nb = d->nb, bw = d->bw, bh = d->bh;
    try {
        if (f.ft == 1) {
            uint8_t qb[5];
            if (!d->read_exact(qb, 5)) return RIV_E_END_OF_STREAM;
            f.q = qb[0];
            f.payload.resize(get_u32(qb + 1));
            if (!d->read_exact(f.payload.data(), f.payload.size())) return RIV_E_END_OF_STREAM;
            f.coded.assign(nb, 1);
            intra_tables(quant_step(f.q), f.yt, f.ct);
            d->have_prev_mvs = false;
            d->mvc.reset();
            d->parsed_intra = true;
        } else {
            uint8_t hb[8];
            if (!d->read_exact(hb, 8)) return RIV_E_END_OF_STREAM;
            f.q = hb[0];
            const int bias_dx = (int8_t)hb[1], bias_dy = (int8_t)hb[2];
            std::vector<uint8_t> mvp(get_u32(hb + 4));
            if (!d->read_exact(mvp.data(), mvp.size())) return RIV_E_END_OF_STREAM;
            uint8_t rs[4];
            if (!d->read_exact(rs, 4)) return RIV_E_END_OF_STREAM;
            f.payload.resize(get_u32(rs));
            if (!d->read_exact(f.payload.data(), f.payload.size())) return RIV_E_END_OF_STREAM;
            if (!d->parsed_intra) { set_error("Inter frame before first intra frame"); return RIV_E_INVALID_FRAME; }
            const auto tm0 = std::chrono::steady_clock::now();
            std::vector<rivh::MvBlock> blocks = d->mvc.decode(mvp.data(), mvp.size(), bw, bh);
            d->t_mv += ns_since(tm0);
            // MV reconstruction (decoder.rs:316-382): base from the predictors of the growing list
            f.mv.assign(nb, Mv4{});
            f.skip.assign(nb, 0); f.coded.assign(nb, 0); f.mask.assign(nb, 0);
            const Mv4* prev = d->have_prev_mvs ? d->prev_mvs.data() : nullptr;
            for (int i = 0; i < nb; i++) {
                const rivh::MvBlock& b = blocks[i];
                Neighbours n = neighbours(f.mv.data(), i, prev, bw, bh, i % bw, i / bw);
                Mv4 base{};
                switch (b.mode) {
                    case riv::M_ZERO: break;
                    case riv::M_NEAREST: base = n.nearest; break;
                    case riv::M_NEAR: base = n.nearp; break;
                    case riv::M_TOPRIGHT: base = n.has_tr ? n.tr : n.nearest; break;
                    case riv::M_TOPLEFT: base = n.has_tl ? n.tl : n.nearest; break;
                    case riv::M_TEMPORAL: base = n.temporal; break;
                    default:
                        base = b.new_base == 1 ? n.nearp : b.new_base == 2 ? (n.has_tr ? n.tr : n.nearest)
                               : b.new_base == 3 ? (n.has_tl ? n.tl : n.nearest) : b.new_base == 4 ? n.temporal : n.nearest;
                }
                int ddx = 0, ddy = 0;
                if (b.mode == riv::M_NEW) {
                    ddx = std::max(-128, std::min(127, (int)b.dx + bias_dx));
                    ddy = std::max(-128, std::min(127, (int)b.dy + bias_dy));
                }
                Mv4 m;
                m.v[0] = (int8_t)std::max(-128, std::min(127, (int)base.v[0] + ddx));
                m.v[1] = (int8_t)std::max(-128, std::min(127, (int)base.v[1] + ddy));
                m.v[2] = b.mode == riv::M_NEW ? b.sx : base.v[2];
                m.v[3] = b.mode == riv::M_NEW ? b.sy : base.v[3];
                f.mv[i] = m;
                f.skip[i] = b.skip;
                f.coded[i] = !b.skip;
                f.mask[i] = !b.skip || m.v[0] || m.v[1] || m.v[2] || m.v[3];   // decoder.rs:412-423
            }
            d->prev_mvs = f.mv;
            d->have_prev_mvs = true;
        }
    } catch (const rivh::CorruptStream& ex) {
        set_error(std::string("corrupt entropy data: ") + ex.what());
        return RIV_E_DECODING_FAILED;
    } catch (const std::exception& ex) {   // the stage may run on a look-ahead thread: no exception leaves it
        set_error(std::string("decoding failed: ") + ex.what());
        return RIV_E_DECODING_FAILED;
    }
    return RIV_OK;
}

// Host stage, part 2 (any order, any thread): the residuals (ResidualDecoder::decode_intra /
// decode_inter, residual.rs:815-966) into the frame's coefficient buffer.
static int decode_frame_residuals(riv_decoder* d, HostFrame& f) {
    int16_t* hco = f.coeffs.as<int16_t>();
    const auto t0 = std::chrono::steady_clock::now();
    std::fill(hco, hco + (size_t)d->nb * 384, 0);
    struct Timed { riv_decoder* d; std::chrono::steady_clock::time_point t0; ~Timed() { d->t_rans += ns_since(t0); } } timed{d, t0};
    try {
        rivh::decode_residuals(f.payload.data(), f.payload.size(), f.coded.data(), d->nb, f.ft == 1, hco);
    } catch (const rivh::CorruptStream& ex) {
        set_error(std::string("corrupt entropy data: ") + ex.what());
        return RIV_E_DECODING_FAILED;
    } catch (const std::exception& ex) {
        set_error(std::string("decoding failed: ") + ex.what());
        return RIV_E_DECODING_FAILED;
    }
    return RIV_OK;
}

// Device stage (frame order): reconstruction, deblock and RGB out.
static int reconstruct(riv_decoder* d, HostFrame& f, uint8_t* rgb_out, riv_decoded_info* info) {
    RIV_HIP(hipSetDevice(d->dev));
    const int W = d->W, H = d->H, nb = d->nb;
    const size_t ysz = (size_t)W * H, csz = ysz / 4;
    uint8_t* nY = d->ref[1 - d->ref_idx].as<uint8_t>(); uint8_t* nU = nY + ysz; uint8_t* nV = nU + csz;
    uint8_t* rY = d->ref[d->ref_idx].as<uint8_t>(); uint8_t* rU = rY + ysz; uint8_t* rV = rU + csz;
    uint8_t* pY = d->pred.as<uint8_t>(); uint8_t* pU = pY + ysz; uint8_t* pV = pU + csz;
    hipStream_t st = d->st;
    const float qs = quant_step(f.q);
    const bool prof = d->profiling;
    if (prof && !d->ev[0])
        for (auto& e : d->ev) RIV_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    if (prof) RIV_HIP(hipEventRecord(d->ev[0], st));
    if (f.ft == 1) {
        RIV_HIP(hipMemcpyAsync(d->yt.p, f.yt, sizeof(f.yt), hipMemcpyHostToDevice, st));
        RIV_HIP(hipMemcpyAsync(d->ct.p, f.ct, sizeof(f.ct), hipMemcpyHostToDevice, st));
        RIV_HIP(hipMemcpyAsync(d->coeffs.p, f.coeffs.p, 768 * (size_t)nb, hipMemcpyHostToDevice, st));
        if (prof) RIV_HIP(hipEventRecord(d->ev[1], st));   // an intra frame has no prediction
        rivk::intra_recon(st, W, H, d->yt.as<uint16_t>(), d->ct.as<uint16_t>(), d->coeffs.as<int16_t>(), nY, nU, nV);
        if (!d->skip_residuals) rivk::deblock(st, nY, nU, nV, W, H, deblock_level(qs), true, nullptr);
    } else {
        if (!d->have_ref) { set_error("Inter frame before first intra frame"); return RIV_E_INVALID_FRAME; }
        RIV_HIP(hipMemcpyAsync(d->mvs.p, f.mv.data(), 4 * (size_t)nb, hipMemcpyHostToDevice, st));
        rivk::mc(st, rY, rU, rV, W, H, d->mvs.as<int8_t>(), 4, pY, pU, pV, d->var.as<uint32_t>(), d->lnv.as<float>());
        if (prof) RIV_HIP(hipEventRecord(d->ev[1], st));
        if (d->skip_residuals) {
            RIV_HIP(hipMemcpyAsync(nY, pY, ysz + 2 * csz, hipMemcpyDeviceToDevice, st));
        } else {
            RIV_HIP(hipMemcpyAsync(d->coeffs.p, f.coeffs.p, 768 * (size_t)nb, hipMemcpyHostToDevice, st));
            RIV_HIP(hipMemcpyAsync(d->skip.p, f.skip.data(), nb, hipMemcpyHostToDevice, st));
            RIV_HIP(hipMemcpyAsync(d->mask.p, f.mask.data(), nb, hipMemcpyHostToDevice, st));
            rivk::aq(st, nb, qs, d->lnv.as<float>(), d->steps.as<float>());
            rivk::inter_recon(st, pY, pU, pV, W, H, d->skip.as<uint8_t>(), d->steps.as<float>(), d->coeffs.as<int16_t>(), nY, nU, nV);
            rivk::deblock(st, nY, nU, nV, W, H, deblock_level(qs), false, d->mask.as<uint8_t>());
        }
    }
    d->ref_idx = 1 - d->ref_idx;
    d->have_ref = true;
    if (prof) RIV_HIP(hipEventRecord(d->ev[2], st));
    const int dw = (int)d->hdr.display_width, dh = (int)d->hdr.display_height;
    if (rgb_out) {
        rivk::yuv420_to_rgb(st, nY, nU, nV, W, dw, dh, d->rgb.as<uint8_t>());
        RIV_HIP(hipMemcpyAsync(rgb_out, d->rgb.p, (size_t)dw * dh * 3, hipMemcpyDeviceToHost, st));
    }
    if (prof) RIV_HIP(hipEventRecord(d->ev[3], st));
    RIV_HIP(hipGetLastError());
    RIV_HIP(hipStreamSynchronize(st));   // the frame's host buffers are free again after this
    if (prof) {
        float a = 0, b = 0, c = 0;
        RIV_HIP(hipEventElapsedTime(&a, d->ev[0], d->ev[1]));
        RIV_HIP(hipEventElapsedTime(&b, d->ev[1], d->ev[2]));
        RIV_HIP(hipEventElapsedTime(&c, d->ev[2], d->ev[3]));
        d->t_pred += (uint64_t)(a * 1e6);
        d->t_apply += (uint64_t)(b * 1e6);
        d->t_rgb += (uint64_t)(c * 1e6);
    }
    if (info) { info->width = (uint32_t)dw; info->height = (uint32_t)dh; info->timestamp_ms = f.ts; info->frame_index = d->current; info->frame_type = f.ft; }
    return RIV_OK;
}

static void pipe_producer(riv_decoder* d, uint64_t first) {
    DecodePipe& P = *d->pipe;
    for (uint64_t i = first; i < d->hdr.frame_count; i++) {
        const int k = (int)(i % DecodePipe::kSlots);
        {
            std::unique_lock<std::mutex> lk(P.mu);
            P.cv.wait(lk, [&] { return P.stop || P.state[k] == 0; });
            if (P.stop) return;
            P.state[k] = 1;
        }
        HostFrame& f = P.slots[k];
        f.status = parse_frame(d, f);
        if (f.status != RIV_OK) {   // the frame carries the error; nothing after it is read
            f.err = last_error();
            std::lock_guard<std::mutex> lk(P.mu);
            P.state[k] = 2;
            P.cv.notify_all();
            return;
        }
        std::lock_guard<std::mutex> lk(P.mu);
        P.jobs.push_back(k);
        P.cv.notify_all();
    }
}

static void pipe_worker(riv_decoder* d) {
    DecodePipe& P = *d->pipe;
    for (;;) {
        int k;
        {
            std::unique_lock<std::mutex> lk(P.mu);
            P.cv.wait(lk, [&] { return P.stop || !P.jobs.empty(); });
            if (P.stop) return;
            k = P.jobs.front();
            P.jobs.pop_front();
        }
        HostFrame& f = P.slots[k];
        f.status = decode_frame_residuals(d, f);
        if (f.status != RIV_OK) f.err = last_error();
        std::lock_guard<std::mutex> lk(P.mu);
        P.state[k] = 2;
        P.cv.notify_all();
    }
}

static void stop_pipe(riv_decoder* d) {
    if (!d->pipe) return;
    {
        std::lock_guard<std::mutex> lk(d->pipe->mu);
        d->pipe->stop = true;
    }
    d->pipe->cv.notify_all();
    if (d->pipe->producer.joinable()) d->pipe->producer.join();
    for (auto& t : d->pipe->workers) t.join();
    delete d->pipe;
    d->pipe = nullptr;
}

static int start_pipe(riv_decoder* d) {
    DecodePipe* P = new DecodePipe();
    for (HostFrame& f : P->slots) {
        if (f.coeffs.alloc(768 * (size_t)d->nb) != hipSuccess) { delete P; set_error("decoder: pinned buffers"); return RIV_E_GPU; }
    }
    d->pipe = P;
    const int cores = (int)std::max(2u, std::thread::hardware_concurrency());
    const int nw = std::max(1, std::min(DecodePipe::kSlots - 1, cores - 2));
    P->producer = std::thread(pipe_producer, d, d->current);
    for (int i = 0; i < nw; i++) P->workers.emplace_back(pipe_worker, d);
    return RIV_OK;
}

// The next frame: from the look-ahead (in-memory streams) or decoded in place (reader streams).
static int decode_next(riv_decoder* d, uint8_t* rgb_out, riv_decoded_info* info) {
    if (!d->mem) {
        HostFrame& f = d->inline_frame;
        if (!f.coeffs.p) RIV_HIP(f.coeffs.alloc(768 * (size_t)d->nb));
        int r = parse_frame(d, f);
        if (!r) r = decode_frame_residuals(d, f);
        if (r) return r;
        return reconstruct(d, f, rgb_out, info);
    }
    if (!d->pipe) {
        int r = start_pipe(d);
        if (r) return r;
    }
    DecodePipe& P = *d->pipe;
    const int k = (int)(d->current % DecodePipe::kSlots);
    {
        std::unique_lock<std::mutex> lk(P.mu);
        P.cv.wait(lk, [&] { return P.state[k] == 2; });
    }
    HostFrame& f = P.slots[k];
    if (f.status != RIV_OK) { set_error(f.err); return f.status; }   // kept: the frame stays failed
    const int r = reconstruct(d, f, rgb_out, info);
    if (r) return r;
    std::lock_guard<std::mutex> lk(P.mu);
    P.state[k] = 0;
    P.cv.notify_all();
    return RIV_OK;
}

int riv_decode_frame(riv_decoder* d, uint8_t* rgb_out, size_t cap, riv_decoded_info* info) {
    if (!d) return RIV_E_ARG;
    if (d->current >= d->hdr.frame_count) { set_error("end of stream"); return RIV_E_END_OF_STREAM; }   // decoder.rs:175-177
    if (rgb_out && cap < (size_t)d->hdr.display_width * d->hdr.display_height * 3) return RIV_E_ARG;
    int r = decode_next(d, rgb_out, info);
    if (r) return r;
    d->current++;
    return RIV_OK;
}
int riv_decoder_decode_frame_null(riv_decoder* d) {
    if (!d) return RIV_E_ARG;
    if (d->current >= d->hdr.frame_count) return RIV_E_END_OF_STREAM;
    std::vector<uint8_t> sink((size_t)d->hdr.display_width * d->hdr.display_height * 3);
    int r = decode_next(d, sink.data(), nullptr);
    if (r) return r;
    d->current++;
    return RIV_OK;
}
int riv_decoder_has_more_frames(const riv_decoder* d) { return d && d->current < d->hdr.frame_count; }
int riv_decoder_skip_residuals(const riv_decoder* d) { return d && d->skip_residuals ? 1 : 0; }
uint64_t riv_decoder_current_frame(const riv_decoder* d) { return d ? d->current : 0; }
int riv_decoder_drain_timings(riv_decoder* d, riv_decode_timings* out) {
    if (!d || !out) return RIV_E_ARG;
    riv_decode_timings t{};
    t.read_bits_ns = d->t_read.exchange(0);
    t.parse_frame_ns = d->t_parse.exchange(0);
    t.mv_decode_ns = d->t_mv.exchange(0);
    t.residual_rans_ns = d->t_rans.exchange(0);
    t.build_pred_ns = d->t_pred.exchange(0);
    t.residual_apply_ns = d->t_apply.exchange(0);
    t.yuv_to_rgb_ns = d->t_rgb.exchange(0);
    t.residual_ns = t.residual_rans_ns + t.residual_deinterleave_ns + t.residual_dct_y_ns + t.residual_dct_uv_ns + t.residual_apply_ns;
    *out = t;
    return RIV_OK;
}
int riv_decoder_set_profiling(riv_decoder* d, int on) {
    if (!d) return RIV_E_ARG;
    d->profiling = on != 0;
    return RIV_OK;
}
int riv_decoder_header(const riv_decoder* d, riv_video_header* out) {
    if (!d || !out) return RIV_E_ARG;
    *out = d->hdr;
    return RIV_OK;
}
int riv_decoder_set_skip_residuals(riv_decoder* d, int skip) {
    if (!d) return RIV_E_ARG;
    d->skip_residuals = skip != 0;
    return RIV_OK;
}
int riv_decoder_debug_yuv(riv_decoder* d, uint8_t* y, uint8_t* u, uint8_t* v) {
    if (!d || !d->have_ref) return RIV_E_ARG;
    RIV_HIP(hipSetDevice(d->dev));
    const size_t ysz = (size_t)d->W * d->H, csz = ysz / 4;
    const uint8_t* src = d->ref[d->ref_idx].as<uint8_t>();
    RIV_HIP(hipMemcpy(y, src, ysz, hipMemcpyDeviceToHost));
    RIV_HIP(hipMemcpy(u, src + ysz, csz, hipMemcpyDeviceToHost));
    RIV_HIP(hipMemcpy(v, src + ysz + csz, csz, hipMemcpyDeviceToHost));
    return RIV_OK;
}
void riv_decoder_destroy(riv_decoder* d) {
    if (!d) return;
    stop_pipe(d);
    (void)hipSetDevice(d->dev);
    if (d->st) { (void)hipStreamSynchronize(d->st); (void)hipStreamDestroy(d->st); }
    for (auto& e : d->ev) if (e) (void)hipEventDestroy(e);
    delete d;
}

}  // extern "C"
