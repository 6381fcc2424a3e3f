// Encoder driver of the C ABI: Encoder / EncoderConfig / encode_frame of reitero_encode
// (reitero_encode/src/encoder.rs, config.rs). Per-frame analysis runs on the GPU; rANS entropy
// coding and packing run on the host. Frames flow through a ring of slots: the GPU writes a
// frame's block records and coefficients into its slot, host workers entropy-code slots
// concurrently (residual coding has fresh contexts per frame), and the ordered tail (MV coding,
// whose contexts carry across frames, packing, writing) runs in frame order.
#include <hip/hip_runtime.h>
#include <chrono>
#include <cmath>
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <exception>
#include <memory>
#include <deque>
#include <map>
#include <tuple>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>
#include "host_common.hpp"
#include "riv_kernels.hpp"
#include "entropy.hpp"
#include "mvnet_host.hpp"

using namespace rivhost;
using riv::BlockRec;

namespace {

struct Slot {
    DevBuf rec;                          // the frame's block records (they also seed the next frame's temporal MVs)
    HostBuf h_rec, h_pack, h_optskip;    // mapped host memory the kernels write for the host entropy stage
    HostBuf h_main;                      // inter frames: each coded MB's masks (and few values), PACK_MAIN bytes per MB
    HostBuf h_passes;                    // mapped: the motion search's pass count
    HostBuf h_sym, h_rmh;                // pinned: a P-frame's residual symbols modelled on the GPU; mapped: their count and status
    DevBuf d_sym;                        // the slot's symbols as the modelling writes them (copied to h_sym by DMA)
    hipEvent_t sym_ev = nullptr;         // that copy's completion
    bool gpu_model = false;              // this P-frame's residual symbols were modelled on the GPU
    int rm_q = 0;                        // then: its record buffer on the device
    hipEvent_t done = nullptr;           // the frame's kernels (and so its host-side outputs) finished
    int prof = 0;                        // this frame's profiling level
    hipEvent_t marks[RIV_NUM_STAGES + 1] = {};
    int mark_stage[RIV_NUM_STAGES + 1] = {};
    int n_marks = 0;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    uint64_t index = 0, ts = 0;
    bool intra = false;
    bool timed = false;      // t0 / t1 were recorded (stage profiling): each event costs the stream microseconds
    bool busy = false;       // holds a frame that is not yet written
    // host-stage products of the frame, passed from its coding job to the ordered tail
    bool gpu_ready = false;  // the GPU work finished; waiting for the ordered MV modelling
    bool coded = false;      // the residual coding finished (products below are valid unless failed)
    bool modeled = false;    // the ordered MV modelling finished
    bool packed = false;     // the frame's bytes are assembled (or it failed)
    bool failed = false;     // the GPU part failed (set before the host parts start)
    bool res_failed = false, mv_failed = false;   // each written only by its own part
    std::vector<uint8_t> mvb, bytes;                 // MV payload, whole frame record
    std::unique_ptr<rivh::RansWriter> mvw;           // the MV symbols between their contexts and their rANS coding
    riv_frame_stats st{};
    std::vector<uint8_t> res;
    std::vector<rivh::MvBlock> blocks;
    std::vector<uint16_t> mv_evs;                    // the frame's MV decisions as (context, bit) events
    size_t mv_nev = 0;
    bool events_ready = false;                       // its MV events are formed (the contexts may run)
    int bdx = 0, bdy = 0;
    double smark[RIV_NUM_STAGES] = {};
    uint32_t nmark[RIV_NUM_STAGES] = {};
    double gpu_ms = 0;
    double t_submit = 0, t_done = 0;   // host timeline (ms since the encoder was created)
};

class Pool {
public:
    // n workers, each running init once before taking jobs
    void start(int n, std::function<void()> init) {
        for (int i = 0; i < n; i++) th_.emplace_back([this, init] { if (init) init(); run(); });
    }
    void submit(std::function<void()> f) {
        { std::lock_guard<std::mutex> lk(mu_); q_.push_back(std::move(f)); }
        cv_.notify_one();
    }
    // Workers waiting for a job right now (a hint: it may change at once).
    int idle() const { return idle_.load(std::memory_order_relaxed); }
    // Runs fn(0) .. fn(n - 1) and returns when all have run: the calling thread takes tasks from
    // a shared counter, and so do helpers queued at the front of the queue for the idle workers.
    // A helper that starts after every task is taken returns without touching fn, so the call
    // never waits on a task no thread holds, from a worker or from any other thread.
    void parallel(int n, const std::function<void(int)>& fn) {
        struct Job {
            std::atomic<int> next{0}, done{0};
            std::mutex mu;
            std::condition_variable cv;
            std::exception_ptr err;   // the first task's exception, rethrown by the caller
        };
        auto job = std::make_shared<Job>();
        const std::function<void(int)>* f = &fn;
        auto work = [job, n, f] {
            for (;;) {
                const int i = job->next.fetch_add(1);
                if (i >= n) return;
                try {
                    (*f)(i);
                } catch (...) {
                    std::lock_guard<std::mutex> lk(job->mu);
                    if (!job->err) job->err = std::current_exception();
                }
                if (job->done.fetch_add(1) + 1 == n) {
                    std::lock_guard<std::mutex> lk(job->mu);
                    job->cv.notify_all();
                }
            }
        };
        const int helpers = std::min(n - 1, idle());
        if (helpers > 0) {
            {
                std::lock_guard<std::mutex> lk(mu_);
                for (int h = 0; h < helpers; h++) q_.push_front(work);
            }
            if (helpers == 1) cv_.notify_one(); else cv_.notify_all();
        }
        work();
        std::unique_lock<std::mutex> lk(job->mu);
        job->cv.wait(lk, [&] { return job->done.load() == n; });
        if (job->err) std::rethrow_exception(job->err);
    }
    ~Pool() {
        { std::lock_guard<std::mutex> lk(mu_); stop_ = true; }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
private:
    void run() {
        for (;;) {
            std::function<void()> f;
            {
                std::unique_lock<std::mutex> lk(mu_);
                idle_.fetch_add(1, std::memory_order_relaxed);
                cv_.wait(lk, [&] { return stop_ || !q_.empty(); });
                idle_.fetch_sub(1, std::memory_order_relaxed);
                if (stop_ && q_.empty()) return;
                f = std::move(q_.front());
                q_.pop_front();
            }
            f();
        }
    }
    std::vector<std::thread> th_;
    std::deque<std::function<void()>> q_;
    std::mutex mu_;
    std::condition_variable cv_;
    bool stop_ = false;
    std::atomic<int> idle_{0};
};

}  // namespace

struct riv_encoder;
namespace {
// Host stage shared by every pipelined encoder of the process: the worker pool that entropy-codes
// frames and the thread that waits for the GPU to finish them, in submission order, before handing
// them to the pool (so no worker ever blocks on the GPU). Created with the first pipelined encoder
// and joined when the last one is destroyed.
struct HostStage {
    Pool pool;
    std::thread waiter;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<std::pair<riv_encoder*, int>> q;   // (encoder, slot) in submission order
    bool stop = false;
};
std::mutex g_hs_mu;
HostStage* g_hs = nullptr;
int g_hs_refs = 0;
}  // namespace

// Persistent motion-search launches on one device run one at a time. Every workgroup of a launch
// must be resident at once (its passes end at a grid barrier), so two launches from different
// streams sharing the CUs could each hold part of the device and wait on each other until their
// barriers time out. A launch on another stream than the device's previous one first waits, on
// the device, for an event recorded on that stream; a stream that launches alone pays nothing.
struct SearchGate {
    std::mutex mu;
    hipEvent_t ev = nullptr;
    hipStream_t last = nullptr;   // stream of the device's latest search launch (null: none in flight)
};
SearchGate& search_gate(int device) {
    static SearchGate gates[64];
    return gates[(unsigned)device % 64u];
}

// Online autotuner of the batched motion search's launch shape, per (blocks, streams, search
// range): the first steps of a key each launch one candidate shape (every shape reaches the same
// fixed point, so the bitstreams do not change), timed by two events around the launch that are
// read once they have completed; after kTuneSamples samples of every candidate the fastest
// mean is kept.
struct ShapeTuner {
    static constexpr int kTuneSamples = 3;
    struct Cand { rivk::SearchShape sh{}; double ms = 0; int n = 0; int issued = 0; };
    struct Pending { int ci; hipEvent_t a, b; };
    std::vector<Cand> cands;
    std::deque<Pending> pending;
    std::vector<hipEvent_t> spare;
    int best = -1;
    int last = -1;   // the candidate of the latest step
    ~ShapeTuner() {
        for (const Pending& q : pending) { (void)hipEventDestroy(q.a); (void)hipEventDestroy(q.b); }
        for (hipEvent_t ev : spare) (void)hipEventDestroy(ev);
    }
    void collect() {   // completed timings, in launch order
        while (!pending.empty() && hipEventQuery(pending.front().b) == hipSuccess) {
            const Pending q = pending.front();
            pending.pop_front();
            float ms = 0;
            if (hipEventElapsedTime(&ms, q.a, q.b) == hipSuccess) { cands[q.ci].ms += ms; cands[q.ci].n++; }
            spare.push_back(q.a); spare.push_back(q.b);
        }
        if (best >= 0 || !pending.empty()) return;
        for (const Cand& c : cands) if (c.n < kTuneSamples) return;
        best = 0;
        for (int i = 1; i < (int)cands.size(); i++)
            if (cands[i].ms / cands[i].n < cands[best].ms / cands[best].n) best = i;
    }
    // the candidate for the next step; -1 means tuning is over (use `best`)
    int next() const {
        if (best >= 0) return -1;
        int pick = -1;
        for (int i = 0; i < (int)cands.size(); i++)
            if (cands[i].issued < kTuneSamples && (pick < 0 || cands[i].issued < cands[pick].issued)) pick = i;
        return pick;
    }
    hipEvent_t event() {
        hipEvent_t ev = nullptr;
        if (!spare.empty()) { ev = spare.back(); spare.pop_back(); return ev; }
        if (hipEventCreate(&ev) != hipSuccess) return nullptr;
        return ev;
    }
};

struct riv_encoder {
    riv_encoder_config cfg{};
    riv_writer_vtbl wv{};
    void* wuser = nullptr;
    bool mem = false;
    std::vector<uint8_t> membuf;
    uint64_t mempos = 0;
    uint64_t header_pos = 0;
    riv_video_header hdr{};
    uint64_t frame_count = 0;      // frames submitted

    int W = 0, H = 0, bw = 0, bh = 0, nb = 0;
    rivk::SearchShape shape{};                 // motion-search launch shape
    int s_maxp = 0, s_bank = 0;                // pass limit, counter bank
    hipStream_t st = nullptr;    // kernels, frame after frame
    DevBuf rgb, cur, pred, ref[2], var, lnv, steps, mask, field, arrive, bits, rdo, yt, ct;
    // GPU residual modelling of P-frames: the residual kernel's device records and the modelling's
    // scratch (sized for nb MBs), and the symbols a frame may have
    DevBuf rm_pack[2], rm_coded[2];   // the residual kernel's records, by frame parity
    DevBuf rm_cnt, rm_last, rm_ke, rm_bcnt, rm_wst, rm_sigw, rm_eobw, rm_bad, rm_off, rm_boff, rm_tt, rm_tp, rm_ev, rm_hdr;
    rivk::ResModelBufs rm{};
    // The modelling runs on a stream of its own, so that it overlaps the next frame's motion
    // search: frame k's residual kernel writes record buffer k % 2 and records rm_ready; its
    // modelling waits for that, reads the buffer and records rm_free, which frame k + 2's residual
    // kernel waits for. Made at the first GPU-modelled frame, with every slot's mapped symbol buffer.
    bool rm_host = false;
    hipStream_t ms = nullptr;
    hipStream_t dl = nullptr;      // device-to-host copies of modelled symbols (the copy engine, no CUs)
    hipEvent_t rm_ready[2] = {}, rm_free[2] = {};
    bool rm_used[2] = {};
    int rm_q = 0;
    uint32_t rm_cap_max = 0;       // the symbol buffers' capacity
    int ref_idx = 0;
    bool have_ref = false;
    int prev_inter_slot = -1;      // slot holding the previous P-frame's records, -1 after an intra frame
    double lambda = 0;
    int c_zero = 0, c_pred = 0, rdo_max = -1;
    float inter_qs = 0, intra_qs = 0;

    std::vector<Slot> slots;
    int depth = 0;                 // configured pipeline depth (0 = synchronous)
    HostStage* hs = nullptr;       // the process's host stage, while pipelined
    HostStage* par = nullptr;      // the same, held by every encoder: its idle workers help code a frame
    hipEvent_t sync_ev = nullptr;  // orders this encoder's stream against a batch step's stream
    hipEvent_t fin_ev[4] = {};     // flush: the end of each of the encoder's streams (blocking-sync events)
    // frames handed over in host memory (riv_encode_frame): each is copied on a stream of its own
    // into one of two device buffers (the first two thirds of `rgb`), so that the copy of frame k,
    // which holds the submitting thread when the source is pageable, overlaps the GPU work of
    // earlier frames; a batch step copies into the last third on the batch's stream
    hipStream_t up = nullptr;
    hipEvent_t up_ready[2] = {}, up_free[2] = {};
    bool up_used[2] = {};
    int up_idx = 0;
    // batch-step scratch (held by the first encoder of a batch): field and barrier counters over all
    // streams of the batch, the stream descriptors and their pinned staging copies
    DevBuf b_field, b_arrive, b_desc;
    size_t b_field_n = 0, b_arrive_n = 0, b_desc_n = 0;
    int b_bank = 0, b_stage = 0, b_maxp = 0;
    HostBuf b_stage_buf[2];
    hipEvent_t b_stage_ev[2] = {nullptr, nullptr};
    hipEvent_t b_copy_ev = nullptr;            // after a batch step's copies of host frames
    bool b_autotune = true;
    DevBuf b_prop, b_lstat;   // learned form in batch steps: proposals of every block, per-wave counters
    size_t b_prop_n = 0;
    std::map<std::tuple<int, int, int>, ShapeTuner> b_tuners;   // by (blocks, streams, search range)
    ShapeTuner* b_last_tuner = nullptr;
    // batch steps led by this encoder with profiling on: events around each search launch, read
    // once complete into stage 7
    std::deque<std::pair<hipEvent_t, hipEvent_t>> b_prof;
    std::vector<hipEvent_t> b_prof_spare;

    // ordered tail + completion state
    std::mutex mu;
    std::condition_variable cv;
    // ordered stages of the host tail: MV modelling (its contexts carry from frame to frame), then,
    // after the packing that any thread may do, the write
    uint64_t next_tail = 0;        // next frame for MV events (and the candidate history)
    uint64_t next_ctx = 0;         // next frame for the MV contexts
    uint64_t next_write = 0;       // next frame to write
    uint64_t completed = 0;        // frames fully written
    bool tail_active = false;      // a thread is forming ready frames' MV events in order
    bool ctx_active = false;       // a thread is running ready frames' MV contexts in order
    bool write_active = false;     // a thread is writing packed frames in order
    int err_code = 0;
    std::string err_msg;
    rivh::MvCoder mvc;
    std::deque<riv_frame_stats> done_stats;
    int last_slot = -1;            // slot of the most recently completed frame (debug accessors)

    // learned form: the proposal network, its confidence threshold (a logit), the per-block
    // proposals of the frame being searched and the running counts
    riv_mvnet* net = nullptr;
    float net_logit = 0.0f;
    int learned_passes = 0;        // learned form: relaxation passes per search (0: to the fixed point)
    DevBuf prop, lstat;

    int pass_cap = 0;              // test seam: relaxation passes allowed per search (0: the proven bound)
    std::atomic<int> split_coding{0};   // test seam: P-frame residual coding 0 GPU modelling, 1 host serial, 2 host split, 3 host by idle workers
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double now_ms() const { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
    int profiling = 0;             // 0 off, 1 motion-search stage only, 2 every stage
    uint64_t prof_every = 1;       // frames whose index is a multiple of it are profiled
    double stage_ms[RIV_NUM_STAGES] = {};
    uint32_t stage_n[RIV_NUM_STAGES] = {};

    int write(const uint8_t* d, size_t n) {
        if (mem) {
            if (mempos + n > membuf.size()) membuf.resize(mempos + n);
            std::memcpy(membuf.data() + mempos, d, n);
            mempos += n;
            return RIV_OK;
        }
        return wv.write(wuser, d, n) == 0 ? RIV_OK : RIV_E_IO;
    }
    uint64_t position() const { return mem ? mempos : wv.position(wuser); }
    int seek(uint64_t pos) {
        if (mem) { if (pos > membuf.size()) membuf.resize(pos); mempos = pos; return RIV_OK; }
        return wv.seek(wuser, pos) == 0 ? RIV_OK : RIV_E_IO;
    }
    void fail(int code, const std::string& msg) {
        std::lock_guard<std::mutex> lk(mu);
        if (!err_code) { err_code = code; err_msg = msg; }
        cv.notify_all();
    }
};

static HostStage* acquire_host_stage();
static void release_host_stage();

// Completed batch-search timings of the encoder leading batch steps, into stage 7.
static void collect_batch_search_times(riv_encoder* e) {
    while (!e->b_prof.empty() && hipEventQuery(e->b_prof.front().second) == hipSuccess) {
        const auto q = e->b_prof.front();
        e->b_prof.pop_front();
        float ms = 0;
        if (hipEventElapsedTime(&ms, q.first, q.second) == hipSuccess) {
            std::lock_guard<std::mutex> lk(e->mu);
            e->stage_ms[7] += ms;
            e->stage_n[7]++;
        }
        e->b_prof_spare.push_back(q.first);
        e->b_prof_spare.push_back(q.second);
    }
}
static hipEvent_t batch_prof_event(riv_encoder* e) {
    hipEvent_t ev = nullptr;
    if (!e->b_prof_spare.empty()) { ev = e->b_prof_spare.back(); e->b_prof_spare.pop_back(); return ev; }
    if (hipEventCreateWithFlags(&ev, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return ev;
}

// Launches the motion search on st once the device's previous search (from any stream) is done.
static int launch_search(int device, hipStream_t st, const riv::SearchParams& p, int groups, int waves, hipEvent_t ev0 = nullptr,
                         hipEvent_t ev1 = nullptr) {
    SearchGate& g = search_gate(device);
    std::lock_guard<std::mutex> lk(g.mu);
    if (!g.ev) RIV_HIP(hipEventCreateWithFlags(&g.ev, hipEventDisableTiming | hipEventDisableSystemFence));
    if (g.last && g.last != st) {   // everything queued on the other stream so far, its search included
        RIV_HIP(hipEventRecord(g.ev, g.last));
        RIV_HIP(hipStreamWaitEvent(st, g.ev, 0));
    }
    RIV_HIP(rivk::me_search(st, p, groups, waves, ev0, ev1));
    g.last = st;
    return RIV_OK;
}

// Error word of a finished inter frame's search (mapped host memory the kernel writes only on
// failure), as a message, or null.
static const char* search_error(const Slot& s) {
    if (s.intra) return nullptr;
    const int code = s.h_passes.as<int>()[1];
    if (!code) return nullptr;
    return code == 2 ? "mode decision: reconstructed MV differs from search MV"
         : code == 3 ? "motion search did not settle within its pass limit" : "motion-search barrier timed out";
}

extern "C" {

void riv_encoder_config_default(riv_encoder_config* c, uint32_t w, uint32_t h, uint32_t fps) {
    std::memset(c, 0, sizeof(*c));
    c->display_width = w; c->display_height = h;
    c->storage_width = (w + 15) / 16 * 16; c->storage_height = (h + 15) / 16 * 16;
    c->fps = fps;
    c->keyframe_mode = RIV_KEYFRAME_FIXED; c->keyframe_interval = 30;
    c->intra_quality = 90; c->inter_quality = 35;
    c->search_range = 12; c->mv_deflate_level = 6; c->skip_threshold = 3;
    c->me_zero_mv_threshold = 0; c->me_predictor_threshold = 0;
    c->rdo_lambda_mult = 0.49; c->inter_dead_zone = 0.75f;
    c->device_id = 0;
    c->pipeline_depth = 0;
}

int riv_encoder_config_validate(const riv_encoder_config* c) {
    if (!c) return RIV_E_ARG;
    if (c->display_width == 0 || c->display_height == 0) { set_error("Width and height must be greater than 0"); return RIV_E_INVALID_CONFIG; }
    if (c->storage_width < c->display_width || c->storage_height < c->display_height) { set_error("Storage dimensions must be >= display dimensions"); return RIV_E_INVALID_CONFIG; }
    if (c->fps == 0) { set_error("FPS must be greater than 0"); return RIV_E_INVALID_CONFIG; }
    if (c->search_range > 127) { set_error("search_range must be <= 127 for current MV encoding"); return RIV_E_INVALID_CONFIG; }
    if (c->mv_deflate_level > 10) { set_error("mv_deflate_level must be in 0..=10"); return RIV_E_INVALID_CONFIG; }
    if (c->keyframe_mode == RIV_KEYFRAME_FIXED && c->keyframe_interval == 0) { set_error("Keyframe interval must be greater than 0"); return RIV_E_INVALID_CONFIG; }
    if (c->pipeline_depth < 0 || c->pipeline_depth > 64) { set_error("pipeline_depth must be in 0..=64"); return RIV_E_INVALID_CONFIG; }
    return RIV_OK;
}

static int encoder_init(riv_encoder* e) {
    const riv_encoder_config& c = e->cfg;
    if (c.storage_width % 16 || c.storage_height % 16) { set_error("storage dims must be multiples of 16"); return RIV_E_INVALID_CONFIG; }
    // MV words hold whole-pixel components in 8 bits, and a NEW block's reconstructed vector equals
    // its search vector while |search - base| fits the delta (every vector within [-sr-1, sr])
    if (c.search_range > 63) { set_error("this backend supports search_range <= 63"); return RIV_E_UNSUPPORTED; }
    int r = ensure_device(c.device_id);
    if (r) return r;
    e->W = (int)c.storage_width; e->H = (int)c.storage_height;
    e->bw = e->W / 16; e->bh = e->H / 16; e->nb = e->bw * e->bh;
    // motion-search launch shape
    {
        hipDeviceProp_t prop;
        RIV_HIP(hipGetDeviceProperties(&prop, c.device_id));
        // 12 waves per workgroup (3 per SIMD) rather than 16: the search waits on its pass barriers
        // far more than it issues, and the quarter of each SIMD's registers (and the LDS) left free
        // lets the residual-modelling kernels of the previous frame run on the same CUs beside it;
        // with 16 waves they wait for the search to end (1080p: the search itself 183 -> 187 us,
        // the pipelined encode 2,740 -> 2,920 frames/s); fewer waves give each wave more blocks
        // to search in the full passes (8 and 10 waves: ~3,240 and ~3,270 against ~3,360 frames/s
        // over 20 pipelined frames)
        if (!rivk::search_shape(c.search_range, e->nb, 1, prop.multiProcessorCount, prop.sharedMemPerBlock, &e->shape, 12) &&
            !rivk::search_shape(c.search_range, e->nb, 1, prop.multiProcessorCount, prop.sharedMemPerBlock, &e->shape)) {
            set_error("frame has more macroblocks than the motion search's resident waves can hold");
            return RIV_E_UNSUPPORTED;
        }
        e->s_maxp = 2 * (e->bh - 1) + e->bw + 2;
    }
    const size_t ysz = (size_t)e->W * e->H, csz = ysz / 4, fsz = ysz + 2 * csz;
    RIV_HIP(hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking));
    RIV_HIP(e->rgb.alloc((size_t)c.display_width * c.display_height * 3 * 3));
    RIV_HIP(e->cur.alloc(fsz)); RIV_HIP(e->pred.alloc(fsz));
    RIV_HIP(e->ref[0].alloc(fsz)); RIV_HIP(e->ref[1].alloc(fsz));
    RIV_HIP(e->var.alloc(8 * (size_t)e->nb)); RIV_HIP(e->lnv.alloc(4 * ((size_t)e->nb + 1))); RIV_HIP(e->steps.alloc(4 * (size_t)e->nb));
    RIV_HIP(e->mask.alloc(e->nb));
    RIV_HIP(e->field.alloc(8 * (size_t)e->nb)); RIV_HIP(hipMemset(e->field.p, 0, 8 * (size_t)e->nb));
    {
        const size_t bb = 2 * sizeof(unsigned) * (size_t)riv::barrier_words(e->s_maxp);
        RIV_HIP(e->arrive.alloc(bb)); RIV_HIP(hipMemset(e->arrive.p, 0, bb));
    }
    RIV_HIP(e->prop.alloc(16 * (size_t)e->nb));
    {   // three learned-form counters per wave of the search launch
        const size_t lb = 3 * sizeof(unsigned long long) * (size_t)e->shape.groups * e->shape.waves;
        RIV_HIP(e->lstat.alloc(lb)); RIV_HIP(hipMemset(e->lstat.p, 0, lb));
    }
    {   // GPU residual modelling: up to 256 symbols per MB on average (a heavier frame is coded on the host)
        const size_t nb = (size_t)e->nb;
        e->rm.nb = e->nb;
        e->rm.cap = e->rm_cap_max = (uint32_t)std::min<size_t>(256 * nb, (1u << 28) - 16);
        for (int q = 0; q < 2; q++) { RIV_HIP(e->rm_pack[q].alloc(riv::PACK_STRIDE * nb)); RIV_HIP(e->rm_coded[q].alloc(nb)); }
        RIV_HIP(e->rm_cnt.alloc(3 * 2 * nb)); RIV_HIP(e->rm_last.alloc(3 * nb)); RIV_HIP(e->rm_ke.alloc(3 * nb));
        RIV_HIP(e->rm_bcnt.alloc(7 * 2 * nb)); RIV_HIP(e->rm_wst.alloc(6 * 2 * nb)); RIV_HIP(e->rm_bad.alloc(nb));
        RIV_HIP(e->rm_sigw.alloc(384 * 4 * nb)); RIV_HIP(e->rm_eobw.alloc(20 * 4 * nb));
        RIV_HIP(e->rm_off.alloc(4 * nb)); RIV_HIP(e->rm_boff.alloc(7 * 4 * nb)); RIV_HIP(e->rm_ev.alloc(384 * 4 * nb));
        const size_t ntiles = (nb + 1023) / 1024;
        RIV_HIP(e->rm_tt.alloc(9 * 4 * ntiles)); RIV_HIP(e->rm_tp.alloc(8 * 4 * ntiles));
        RIV_HIP(e->rm_hdr.alloc(64));
        e->rm.cnt = e->rm_cnt.as<uint16_t>(); e->rm.last = e->rm_last.as<uint8_t>(); e->rm.ke = e->rm_ke.as<uint8_t>();
        e->rm.bcnt = e->rm_bcnt.as<uint16_t>(); e->rm.wst = e->rm_wst.as<uint16_t>(); e->rm.bad = e->rm_bad.as<uint8_t>();
        e->rm.sigw = e->rm_sigw.as<uint32_t>(); e->rm.eobw = e->rm_eobw.as<uint32_t>();
        e->rm.off = e->rm_off.as<uint32_t>(); e->rm.boff = e->rm_boff.as<uint32_t>(); e->rm.ev = e->rm_ev.as<uint32_t>();
        e->rm.tt = e->rm_tt.as<uint32_t>(); e->rm.tp = e->rm_tp.as<uint32_t>();
        e->rm.hdr = e->rm_hdr.as<uint32_t>();   // the symbol buffer is the frame's slot's
    }
    e->depth = c.pipeline_depth;
    e->slots.resize(std::max(2, e->depth));
    for (Slot& s : e->slots) {
        RIV_HIP(s.rec.alloc(sizeof(BlockRec) * e->nb));
        RIV_HIP(s.h_rec.alloc(sizeof(BlockRec) * e->nb)); RIV_HIP(s.h_pack.alloc((size_t)riv::PACK_STRIDE * e->nb));
        RIV_HIP(s.h_optskip.alloc(e->nb)); RIV_HIP(s.h_passes.alloc(16)); RIV_HIP(s.h_main.alloc((size_t)riv::PACK_MAIN * e->nb));
        RIV_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming | hipEventBlockingSync));   // waiting workers sleep instead of spinning
        // timing-only events skip the system-scope fence (a cache write-back per record); done keeps
        // it: the host reads what the kernels wrote to mapped memory once done has fired
        RIV_HIP(hipEventCreateWithFlags(&s.t0, hipEventDisableSystemFence));
        RIV_HIP(hipEventCreateWithFlags(&s.t1, hipEventDisableSystemFence));
    }
    RIV_HIP(hipEventCreateWithFlags(&e->sync_ev, hipEventDisableTiming));
    e->par = acquire_host_stage();
    if (e->depth > 0) e->hs = e->par;
    // MV-cost bit table (motion/scalar.rs:495-525) and lambda-derived costs
    rivhost::RdoModel rdo(c.inter_quality, c.rdo_lambda_mult, c.skip_threshold);
    e->lambda = rdo.lambda;
    double bitsv[256];
    for (int a = 0; a < 256; a++) bitsv[a] = a == 0 ? 1.0 : 2.0 * std::log2((double)a) + 2.0;
    RIV_HIP(e->bits.alloc(sizeof(bitsv)));
    RIV_HIP(hipMemcpy(e->bits.p, bitsv, sizeof(bitsv), hipMemcpyHostToDevice));
    const double lam2 = e->lambda * 2.0;
    // every MV cost is lam2 * (6 + bits_x + bits_y) <= lam2 * 42; the kernels keep SAD + cost in
    // 27 bits of a search key
    if (!(e->lambda >= 0.0) || lam2 * 42.0 >= 67108864.0) {
        set_error("rdo_lambda_mult gives a lambda outside 0 <= lambda < 7.9e5, which this backend does not support");
        return RIV_E_UNSUPPORTED;
    }
    e->c_zero = (int)std::min(1e9, std::trunc(lam2 * 0.5));
    e->c_pred = (int)std::min(1e9, std::trunc(lam2 * 1.5));
    // RDO skip table over every reachable SAD (a 16x16 SAD is at most 65280)
    e->rdo_max = c.skip_threshold == 0 ? -1 : (int)std::min<int64_t>(65280, 768 * (int64_t)c.skip_threshold);
    std::vector<uint32_t> rb((65281 + 31) / 32, 0);
    for (int s = 0; s <= e->rdo_max; s++) if (rdo.skip(s, true)) rb[s >> 5] |= 1u << (s & 31);
    RIV_HIP(e->rdo.alloc(rb.size() * 4));
    RIV_HIP(hipMemcpy(e->rdo.p, rb.data(), rb.size() * 4, hipMemcpyHostToDevice));
    e->inter_qs = quant_step(c.inter_quality);
    e->intra_qs = quant_step(c.intra_quality);
    uint16_t yt[256], ct[64];
    intra_tables(e->intra_qs, yt, ct);
    RIV_HIP(e->yt.alloc(sizeof(yt))); RIV_HIP(e->ct.alloc(sizeof(ct)));
    RIV_HIP(hipMemcpy(e->yt.p, yt, sizeof(yt), hipMemcpyHostToDevice));
    RIV_HIP(hipMemcpy(e->ct.p, ct, sizeof(ct), hipMemcpyHostToDevice));
    return RIV_OK;
}

static int write_header(riv_encoder* e) {
    std::vector<uint8_t> h;
    put_header(h, e->hdr);
    return e->write(h.data(), h.size());
}

static int create_common(const riv_encoder_config* cfg, riv_encoder* e) {
    int r = riv_encoder_config_validate(cfg);
    if (r) return r;
    e->cfg = *cfg;
    r = encoder_init(e);
    if (r) return r;
    e->hdr.display_width = cfg->display_width; e->hdr.display_height = cfg->display_height;
    e->hdr.storage_width = cfg->storage_width; e->hdr.storage_height = cfg->storage_height;
    e->hdr.fps = cfg->fps; e->hdr.frame_count = 0;
    e->header_pos = e->position();
    return write_header(e);   // Encoder::new writes the header immediately (encoder.rs:181-185)
}

static void destroy_encoder(riv_encoder* e) {
    if (!e) return;
    (void)hipSetDevice(e->cfg.device_id);
    if (e->hs) {   // every submitted frame reaches its tail (errors included) before the buffers go away
        std::unique_lock<std::mutex> lk(e->mu);
        e->cv.wait(lk, [&] { return e->completed == e->frame_count; });
        lk.unlock();
        e->hs = nullptr;
    }
    if (e->par) {
        release_host_stage();
        e->par = nullptr;
    }
    if (e->st) {
        (void)hipStreamSynchronize(e->st);
        SearchGate& g = search_gate(e->cfg.device_id);
        std::lock_guard<std::mutex> lk(g.mu);
        if (g.last == e->st) g.last = nullptr;   // its searches are done; the stream goes away
    }
    if (e->dl) {
        (void)hipStreamSynchronize(e->dl);
        (void)hipStreamDestroy(e->dl);
    }
    for (Slot& s : e->slots) {
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.t0) (void)hipEventDestroy(s.t0);
        if (s.t1) (void)hipEventDestroy(s.t1);
        if (s.sym_ev) (void)hipEventDestroy(s.sym_ev);
        for (auto& m : s.marks) if (m) (void)hipEventDestroy(m);
    }
    if (e->sync_ev) (void)hipEventDestroy(e->sync_ev);
    for (hipEvent_t& f : e->fin_ev) if (f) (void)hipEventDestroy(f);
    if (e->ms) {
        (void)hipStreamSynchronize(e->ms);
        for (int k = 0; k < 2; k++) { (void)hipEventDestroy(e->rm_ready[k]); (void)hipEventDestroy(e->rm_free[k]); }
        (void)hipStreamDestroy(e->ms);
    }
    if (e->up) {
        (void)hipStreamSynchronize(e->up);
        for (int b = 0; b < 2; b++) { (void)hipEventDestroy(e->up_ready[b]); (void)hipEventDestroy(e->up_free[b]); }
        (void)hipStreamDestroy(e->up);
    }
    for (auto& q : e->b_prof) { (void)hipEventDestroy(q.first); (void)hipEventDestroy(q.second); }
    for (hipEvent_t ev : e->b_prof_spare) (void)hipEventDestroy(ev);
    for (auto& ev : e->b_stage_ev) if (ev) (void)hipEventDestroy(ev);
    if (e->b_copy_ev) (void)hipEventDestroy(e->b_copy_ev);
    if (e->st) (void)hipStreamDestroy(e->st);
    delete e;
}

int riv_encoder_create(const riv_encoder_config* cfg, const riv_writer_vtbl* writer, void* user, riv_encoder** out) {
    if (!cfg || !writer || !out || !writer->write || !writer->position || !writer->seek || !writer->flush) return RIV_E_ARG;
    riv_encoder* e = new riv_encoder();
    e->wv = *writer; e->wuser = user;
    int r = create_common(cfg, e);
    if (r) { destroy_encoder(e); return r; }
    *out = e;
    return RIV_OK;
}

int riv_encoder_create_mem(const riv_encoder_config* cfg, riv_encoder** out) {
    if (!cfg || !out) return RIV_E_ARG;
    riv_encoder* e = new riv_encoder();
    e->mem = true;
    int r = create_common(cfg, e);
    if (r) { destroy_encoder(e); return r; }
    *out = e;
    return RIV_OK;
}

// compute_delta_bias_for_new (encoder.rs:984-1048) from a histogram of the raw NEW deltas:
// per axis, the bias in [-24, 24] with the most zero deltas, ties to the smallest L1 sum.
static void compute_bias(const BlockRec* rec, int nb, int& bx_out, int& by_out) {
    uint32_t hx[256] = {0}, hy[256] = {0};
    uint32_t n = 0;
    for (int i = 0; i < nb; i++)
        if (rec[i].mode == riv::M_NEW) { hx[(uint8_t)rec[i].raw_dx]++; hy[(uint8_t)rec[i].raw_dy]++; n++; }
    bx_out = 0; by_out = 0;
    if (!n) return;
    for (int axis = 0; axis < 2; axis++) {
        const uint32_t* h = axis ? hy : hx;
        int64_t best_zero = -1, best_cost = INT64_MAX; int best = 0;
        for (int b = -24; b <= 24; b++) {
            int64_t zero = 0, cost = 0;
            for (int v = -128; v < 128; v++) {
                uint32_t cnt = h[(uint8_t)(int8_t)v];
                if (!cnt) continue;
                int adj = std::max(-128, std::min(127, v - b));
                if (adj == 0) zero += cnt;
                cost += (int64_t)cnt * std::abs(adj);
            }
            if (zero > best_zero || (zero == best_zero && cost < best_cost)) { best_zero = zero; best_cost = cost; best = b; }
        }
        (axis ? by_out : bx_out) = best;
    }
}

// Host stage of one frame, part 1 (any order, any thread): residual rANS. The frame's GPU work
// has finished. The products stay in the slot for the ordered tail.
// Device times of a finished frame, read once its events have completed (by the thread that
// waited for them, so that the host-stage workers make no runtime calls).
static void collect_gpu_times(Slot& s) {
    float gpu_ms = 0;
    s.gpu_ms = s.timed && hipEventElapsedTime(&gpu_ms, s.t0, s.t1) == hipSuccess ? gpu_ms : 0.0;
    for (int k = 0; k < RIV_NUM_STAGES; k++) { s.smark[k] = 0; s.nmark[k] = 0; }
    for (int k = 1; k < s.n_marks; k++) {
        float ms = 0;
        if (s.mark_stage[k] >= 0 && hipEventElapsedTime(&ms, s.marks[k - 1], s.marks[k]) == hipSuccess) { s.smark[s.mark_stage[k]] += ms; s.nmark[s.mark_stage[k]]++; }
    }
}

// Host stage of one frame, once its GPU work has finished: two parts that run concurrently, then
// the packing and the ordered write.
//  * the residual coding (any order, any thread; its contexts are fresh every frame);
//  * the MV modelling (frame order, one thread at a time: the MV contexts and candidate history
//    carry from frame to frame), which needs only the block records, so it starts as soon as the
//    GPU is done instead of after the residuals, then the MV rANS coding (any order);
// whichever of the two finishes second hands the frame to the packing (any thread), and the
// writes run in frame order.
static void frame_part_done(riv_encoder* e, int si, bool mv_part);

// A GPU-modelled frame's n symbols from the slot's device buffer to its pinned host buffer, by the
// copy engine (a kernel storing to host memory would hold CUs for the whole PCIe transfer, and keep
// the next motion search from starting on them).
static void fetch_symbols(riv_encoder* e, Slot& s, uint32_t n) {
    if (n == 0) return;
    hipError_t r = hipSetDevice(e->cfg.device_id);
    if (r == hipSuccess) r = hipMemcpyAsync(s.h_sym.p, s.d_sym.p, (size_t)n * 2, hipMemcpyDeviceToHost, e->dl);
    if (r == hipSuccess) r = hipEventRecord(s.sym_ev, e->dl);
    if (r == hipSuccess) r = hipEventSynchronize(s.sym_ev);
    if (r != hipSuccess) throw std::runtime_error(std::string("copying the modelled symbols: ") + hipGetErrorString(r));
}

static void code_residuals(riv_encoder* e, int si) {
    Slot& s = e->slots[si];
    const int nb = e->nb;
    auto t0 = std::chrono::steady_clock::now();
    riv_frame_stats& st = s.st;
    const uint8_t* pack = s.h_pack.as<uint8_t>();
    s.res.clear();
    try {
        if (s.failed) {
            // nothing to code: the frame is dropped at the write
        } else if (s.intra) {
            std::vector<uint8_t> coded(nb, 1);
            s.res = rivh::encode_residuals_packed(pack, riv::PACK_STRIDE, coded.data(), nb, true);
        } else if (s.gpu_model && s.h_rmh.as<uint32_t>()[1] == 0) {
            // the GPU modelled the symbols: only their rANS32 coding is left, in groups of sixteen
            // 16,386-symbol blocks side by side, across idle workers when there are any
            const uint32_t n = s.h_rmh.as<uint32_t>()[0];
            fetch_symbols(e, s, n);
            st.residual_symbols_on_gpu = 1;
            // groups of eight blocks when the idle workers (and this thread) can take them all at
            // once: twice the tasks of the sixteen-block groups, each done in about half the time
            const int idle = e->par ? e->par->pool.idle() : 0;
            const int blocks = (int)((n + 16385) / 16386);
            const int gb = idle + 1 >= (blocks + 7) / 8 ? 8 : 16;
            s.res = rivh::code_symbol_stream(s.h_sym.as<uint16_t>(), n, idle >= 2 ? rivh::ParallelFor([e](int k, const std::function<void(int)>& f) {
                e->par->pool.parallel(k, f);
            }) : rivh::ParallelFor(), gb);
        } else if (s.gpu_model) {
            // the GPU could not model this frame (too many symbols, or a magnitude beyond the class
            // table): it sent the full records instead, and the host codes them (or rejects them)
            const uint8_t* os = s.h_optskip.as<uint8_t>();
            std::vector<uint8_t> coded(nb);
            for (int i = 0; i < nb; i++) coded[i] = !os[i];
            s.res = rivh::encode_residuals_packed(pack, riv::PACK_STRIDE, coded.data(), nb, false);
        } else {
            const uint8_t* os = s.h_optskip.as<uint8_t>();
            std::vector<uint8_t> coded(nb);
            for (int i = 0; i < nb; i++) coded[i] = !os[i];
            // with idle host workers the frame's coding is split across them (the same bytes; the
            // split costs more work in all, so a pool busy with other frames codes it serially)
            const int mode = e->split_coding.load();
            const bool split = e->par && (mode == 2 || ((mode == 0 || mode == 3) && e->par->pool.idle() >= 2));
            if (split)
                s.res = rivh::encode_residuals_split_par(s.h_main.as<uint8_t>(), pack, coded.data(), nb,
                                                         [e](int n, const std::function<void(int)>& f) { e->par->pool.parallel(n, f); });
            else
                s.res = rivh::encode_residuals_split(s.h_main.as<uint8_t>(), pack, coded.data(), nb);
        }
    } catch (const std::exception& ex) {
        s.res_failed = true;
        e->fail(RIV_E_ENCODING_FAILED, std::string("entropy coding failed: ") + ex.what());
    }
    st.host_entropy_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    st.t_coded_ms = e->now_ms();
    frame_part_done(e, si, false);
}

// The block list of P-frame s from its records (NEW deltas after the frame bias,
// encoder.rs:775-818, bias by compute_delta_bias_for_new, :984-1048) and its stats: any order,
// any thread (nothing carries from frame to frame).
static int prepare_blocks(riv_encoder* e, Slot& s) {
    riv_frame_stats& st = s.st;
    const int nb = e->nb;
    try {
        if (s.intra) {
            st.frame_type = 1;
        } else {
            const BlockRec* hr = s.h_rec.as<BlockRec>();
            const uint8_t* os = s.h_optskip.as<uint8_t>();
            int bdx = 0, bdy = 0;
            compute_bias(hr, nb, bdx, bdy);
            s.blocks.resize(nb);
            for (int i = 0; i < nb; i++) {
                rivh::MvBlock& b = s.blocks[i];
                b.mode = hr[i].mode; b.new_base = hr[i].new_base;
                b.dx = b.dy = 0; b.sx = b.sy = 0;
                if (b.mode == riv::M_NEW) {
                    b.dx = (int8_t)std::max(-128, std::min(127, (int)hr[i].raw_dx - bdx));
                    b.dy = (int8_t)std::max(-128, std::min(127, (int)hr[i].raw_dy - bdy));
                    b.sx = hr[i].s[2]; b.sy = hr[i].s[3];
                    st.mv_new_blocks++;
                }
                b.skip = os[i] != 0;
                st.blocks_skipped += b.skip;
                st.mv_mode_counts[b.mode]++;
                st.rdo_blocks_selected_skip += hr[i].rdo_skip;
                if (hr[i].s[2] || hr[i].s[3]) st.rdo_forced_fractional++;
            }
            s.bdx = bdx; s.bdy = bdy;
            st.frame_type = 2;
            st.mv_bias_dx = bdx; st.mv_bias_dy = bdy;
            st.rdo_lambda = e->lambda; st.rdo_blocks_evaluated = (uint32_t)nb;
            if (e->cfg.skip_threshold == 0) st.rdo_forced_disabled = (uint32_t)nb;
            st.me_passes = (uint32_t)*s.h_passes.as<int>();
        }
    } catch (const std::exception& ex) {
        e->fail(RIV_E_ENCODING_FAILED, std::string("MV coding failed: ") + ex.what());
        return RIV_E_ENCODING_FAILED;
    }
    return RIV_OK;
}

// The MV events of frame s and the history update (mv_rans/encoder.rs; contexts and history
// restart at every I-frame, encoder.rs:335-337): frame order.
static int model_frame(riv_encoder* e, Slot& s) {
    try {
        if (s.intra) e->mvc.reset_history();
        else s.mv_nev = e->mvc.events(s.blocks, e->bw, s.mv_evs);
    } catch (const std::exception& ex) {
        e->fail(RIV_E_ENCODING_FAILED, std::string("MV coding failed: ") + ex.what());
        return RIV_E_ENCODING_FAILED;
    }
    return RIV_OK;
}

// The MV contexts (frame order; the contexts carry from frame to frame) over the frame's events,
// its symbols into the frame's MV writer. Runs beside the next frame's model_frame.
static int model_contexts(riv_encoder* e, Slot& s) {
    try {
        if (s.intra) {
            e->mvc.reset_contexts();
        } else {
            s.mvb.clear();
            s.mvw.reset(new rivh::RansWriter(s.mvb));
            e->mvc.contexts(s.mv_evs.data(), s.mv_nev, *s.mvw);
        }
    } catch (const std::exception& ex) {
        e->fail(RIV_E_ENCODING_FAILED, std::string("MV coding failed: ") + ex.what());
        return RIV_E_ENCODING_FAILED;
    }
    return RIV_OK;
}

// Part 3 (any order, any thread): the frame record (format.rs:171-205).
static void pack_frame(riv_encoder* e, Slot& s) {
    const riv_encoder_config& c = e->cfg;
    std::vector<uint8_t>& bytes = s.bytes;
    bytes.clear();
    if (s.intra) {
        bytes.reserve(14 + s.res.size());
        put_u64(bytes, s.ts); bytes.push_back(1); bytes.push_back(c.intra_quality);
        put_u32(bytes, (uint32_t)s.res.size());
        bytes.insert(bytes.end(), s.res.begin(), s.res.end());
        s.st.residual_bytes = (uint32_t)s.res.size();
    } else {
        bytes.reserve(21 + s.mvb.size() + s.res.size());
        put_u64(bytes, s.ts); bytes.push_back(2); bytes.push_back(c.inter_quality);
        bytes.push_back((uint8_t)(int8_t)s.bdx); bytes.push_back((uint8_t)(int8_t)s.bdy); bytes.push_back(0);
        put_u32(bytes, (uint32_t)s.mvb.size()); bytes.insert(bytes.end(), s.mvb.begin(), s.mvb.end());
        put_u32(bytes, (uint32_t)s.res.size()); bytes.insert(bytes.end(), s.res.begin(), s.res.end());
        s.st.mv_bytes = (uint32_t)s.mvb.size(); s.st.residual_bytes = (uint32_t)s.res.size();
    }
    s.st.total_bytes = (uint32_t)bytes.size();
}

static void run_writes(riv_encoder* e, int si);

// Packs slot si (on the pool when pipelined), then runs the ordered writes.
static void pack_then_write(riv_encoder* e, int si) {
    Slot& s = e->slots[si];
    const auto t0 = std::chrono::steady_clock::now();
    bool ok;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        ok = !s.failed && !s.res_failed && !s.mv_failed && !e->err_code;
    }
    if (ok) pack_frame(e, s);
    s.st.host_tail_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    run_writes(e, si);
}

// One of the frame's two parts finished: the second one sends the frame to the packing.
static void frame_part_done(riv_encoder* e, int si, bool mv_part) {
    Slot& s = e->slots[si];
    bool both;
    {
        std::lock_guard<std::mutex> lk(e->mu);
        (mv_part ? s.modeled : s.coded) = true;
        both = s.coded && s.modeled;
    }
    if (!both) return;
    if (e->hs) e->hs->pool.submit([e, si] { pack_then_write(e, si); });
    else pack_then_write(e, si);
}

// The frame's MV rANS coding (any order, any thread), as soon as its contexts are done, so that it
// runs beside the frame's residual coding rather than after it, its blocks across idle workers
// when there are any; then the MV part is done.
static void drain_mv(riv_encoder* e, int si) {
    Slot& s = e->slots[si];
    const auto t0 = std::chrono::steady_clock::now();
    if (s.mvw) {
        try {
            if (!s.mv_failed) {
                if (e->par && e->par->pool.idle() >= 2)
                    s.mvw->flush_parallel([e](int k, const std::function<void(int)>& f) { e->par->pool.parallel(k, f); });
                else
                    s.mvw->flush();
            }
        } catch (const std::exception& ex) {
            s.mv_failed = true;
            e->fail(RIV_E_ENCODING_FAILED, std::string("MV coding failed: ") + ex.what());
        }
        s.mvw.reset();
    }
    s.st.host_tail_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    frame_part_done(e, si, true);
}

// The MV part of a frame runs in two ordered stages on the host's workers: its events (and the
// candidate history for the next frame), then its contexts, so that frame k's contexts run beside
// frame k + 1's events. Each stage, unless another thread is already doing so, takes every ready
// frame in frame order; it works outside the lock.
static void run_ctx(riv_encoder* e, int si) {
    std::unique_lock<std::mutex> lk(e->mu);
    e->slots[si].events_ready = true;
    if (e->ctx_active) return;
    e->ctx_active = true;
    for (;;) {
        const int ti = (int)(e->next_ctx % e->slots.size());
        Slot& t = e->slots[ti];
        if (!t.busy || !t.events_ready || t.index != e->next_ctx) break;
        t.events_ready = false;
        const bool skip = t.mv_failed || t.failed || e->err_code;
        lk.unlock();
        const auto t0 = std::chrono::steady_clock::now();
        if (skip || model_contexts(e, t) != RIV_OK) t.mv_failed = true;
        t.st.host_tail_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        lk.lock();
        e->next_ctx++;
        lk.unlock();
        if (e->hs) e->hs->pool.submit([e, ti] { drain_mv(e, ti); });
        else drain_mv(e, ti);
        lk.lock();
    }
    e->ctx_active = false;
}

// Marks slot si's GPU work done, then forms the MV events of every ready frame in order; the
// frame's block list first, outside the ordered stage.
static void run_mv(riv_encoder* e, int si) {
    {
        Slot& s = e->slots[si];
        const auto t0 = std::chrono::steady_clock::now();
        if (!s.failed && prepare_blocks(e, s) != RIV_OK) s.mv_failed = true;
        s.st.host_tail_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    std::unique_lock<std::mutex> lk(e->mu);
    e->slots[si].gpu_ready = true;
    if (e->tail_active) return;
    e->tail_active = true;
    for (;;) {
        const int ti = (int)(e->next_tail % e->slots.size());
        Slot& t = e->slots[ti];
        if (!t.busy || !t.gpu_ready || t.index != e->next_tail) break;
        t.gpu_ready = false;
        const bool skip = t.failed || t.mv_failed || e->err_code;
        lk.unlock();
        const auto t0 = std::chrono::steady_clock::now();
        if (skip || model_frame(e, t) != RIV_OK) t.mv_failed = true;
        t.st.host_tail_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        lk.lock();
        e->next_tail++;
        lk.unlock();
        if (e->hs) e->hs->pool.submit([e, ti] { run_ctx(e, ti); });
        else run_ctx(e, ti);
        lk.lock();
    }
    e->tail_active = false;
}

// A frame whose GPU work has finished (or failed, when `failed`): its stats start afresh and both
// parts are started, on the pool when pipelined, else on the calling thread.
static void start_host_parts(riv_encoder* e, int si, bool failed) {
    Slot& s = e->slots[si];
    riv_frame_stats& st = s.st;
    st = riv_frame_stats{};
    st.frame_index = s.index; st.timestamp_ms = s.ts;
    st.storage_width = (uint32_t)e->W; st.storage_height = (uint32_t)e->H; st.blocks_total = (uint32_t)e->nb;
    st.gpu_ms = s.gpu_ms;
    st.t_submit_ms = s.t_submit; st.t_gpu_done_ms = s.t_done;
    if (failed) s.failed = true;
    if (e->hs) {
        e->hs->pool.submit([e, si] { run_mv(e, si); });
        e->hs->pool.submit([e, si] { code_residuals(e, si); });
    } else {
        run_mv(e, si);
        code_residuals(e, si);
    }
}

// Marks slot si packed, then, unless another thread is already doing so, writes every packed
// frame in frame order.
static void run_writes(riv_encoder* e, int si) {
    std::unique_lock<std::mutex> lk(e->mu);
    e->slots[si].packed = true;
    if (e->write_active) return;
    e->write_active = true;
    for (;;) {
        const int ti = (int)(e->next_write % e->slots.size());
        Slot& t = e->slots[ti];
        if (!t.busy || !t.packed || t.index != e->next_write) break;
        t.packed = false;
        const bool ok = !t.failed && !t.res_failed && !t.mv_failed && !e->err_code;
        lk.unlock();
        int r = RIV_OK;
        if (ok) {
            r = e->write(t.bytes.data(), t.bytes.size());
            if (r) e->fail(r, "writer failed");
        }
        lk.lock();
        if (ok && r == RIV_OK) {
            t.st.t_written_ms = e->now_ms();
            if (e->profiling)
                for (int k = 0; k < RIV_NUM_STAGES; k++) { e->stage_ms[k] += t.smark[k]; e->stage_n[k] += t.nmark[k]; }
            e->done_stats.push_back(t.st);
            if (e->done_stats.size() > 4096) e->done_stats.pop_front();
            e->hdr.frame_count = t.index + 1;
            e->last_slot = ti;
        }
        e->next_write++;
        e->completed++;
        t.busy = false;
        e->cv.notify_all();
    }
    e->write_active = false;
}

// The GPU-completion thread of the host stage.
static void waiter_loop(HostStage* hs) {
    for (;;) {
        std::pair<riv_encoder*, int> job;
        {
            std::unique_lock<std::mutex> lk(hs->mu);
            hs->cv.wait(lk, [&] { return hs->stop || !hs->q.empty(); });
            if (hs->q.empty()) return;
            job = hs->q.front();
            hs->q.pop_front();
        }
        riv_encoder* e = job.first;
        const int si = job.second;
        (void)hipSetDevice(e->cfg.device_id);
        if (hipEventSynchronize(e->slots[si].done) != hipSuccess) {
            e->fail(RIV_E_GPU, "GPU frame failed");
            start_host_parts(e, si, true);
            continue;
        }
        e->slots[si].t_done = e->now_ms();
        if (const char* msg = search_error(e->slots[si])) {   // no bytes for a frame whose search failed
            e->fail(RIV_E_GPU, msg);
            start_host_parts(e, si, true);
            continue;
        }
        collect_gpu_times(e->slots[si]);
        start_host_parts(e, si, false);
    }
}

static HostStage* acquire_host_stage() {
    std::lock_guard<std::mutex> lk(g_hs_mu);
    if (!g_hs) {
        g_hs = new HostStage();
        // workers: up to 16, leaving one core to the submitting and completion threads (both sleep
        // on the GPU and on free slots most of the time)
        const int cores = (int)std::max(2u, std::thread::hardware_concurrency());
        // each worker's rANS buffers sized for a 1080p frame's residual symbols up front
        g_hs->pool.start(std::max(1, std::min(16, cores - 1)), [] { rivh::RansWriter::prewarm(1u << 20); });
        g_hs->waiter = std::thread(waiter_loop, g_hs);
    }
    g_hs_refs++;
    return g_hs;
}

static void release_host_stage() {
    std::lock_guard<std::mutex> lk(g_hs_mu);
    if (--g_hs_refs > 0) return;
    { std::lock_guard<std::mutex> q(g_hs->mu); g_hs->stop = true; }
    g_hs->cv.notify_all();
    g_hs->waiter.join();
    delete g_hs;   // joins the workers
    g_hs = nullptr;
}

// ---- one frame, in phases: begin (up to the motion search), search, after the search, end ----
struct FrameCtx {
    int si = 0;
    bool intra = false;
    bool batch = false;            // part of a batch step (its residuals are coded on the host)
    hipStream_t st = nullptr;
    uint8_t *cY, *cU, *cV, *rY, *rU, *rV, *nY, *nU, *nV, *pY, *pU, *pV;
    const int8_t* prev_mv = nullptr;
};

static void mark(riv_encoder* e, Slot& s, hipStream_t st, int stage) {
    if (!s.prof || s.n_marks > RIV_NUM_STAGES) return;
    if (s.prof == 1 && stage != 0 && stage != 1) return;   // two events per frame: around the search
    (void)hipEventRecord(s.marks[s.n_marks], st);
    s.mark_stage[s.n_marks++] = stage;
}
// The frame's next profiling event for `stage`, taken but not recorded: a kernel launch records it
// at the kernel's start or end, which costs the stream nothing (a recorded event holds it for
// microseconds). Null when profiling is off.
static hipEvent_t mark_by_launch(riv_encoder* e, Slot& s, int stage) {
    if (!s.prof || s.n_marks > RIV_NUM_STAGES) return nullptr;
    s.mark_stage[s.n_marks] = stage;
    return s.marks[s.n_marks++];
}
// Single-frame plain P-frames under motion-search profiling: the search launch records the two
// events around it.
static bool search_marks_by_launch(const riv_encoder* e, const FrameCtx& f) {
    return e->slots[f.si].prof == 1 && !f.intra && !f.batch && !e->net;
}

static int check_frame(riv_encoder* e, size_t len, uint32_t w, uint32_t h) {
    const riv_encoder_config& c = e->cfg;
    if (w != c.display_width || h != c.display_height) { set_error("Frame dimensions do not match config"); return RIV_E_FRAME; }
    if (len != (size_t)w * h * 3) { set_error("Display RGB24 size mismatch"); return RIV_E_FRAME; }
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->err_code) { set_error(e->err_msg); return e->err_code; }
    return RIV_OK;
}

// Takes a slot, converts the frame, and for an intra frame runs all of its GPU work.
static int frame_begin(riv_encoder* e, const void* rgb, bool on_device, size_t len, uint32_t w, uint32_t h, uint64_t ts,
                       hipStream_t st, FrameCtx& f) {
    const riv_encoder_config& c = e->cfg;
    int r = check_frame(e, len, w, h);
    if (r) return r;
    RIV_HIP(hipSetDevice(c.device_id));
    // keyframe decision (encoder.rs:265-291)
    bool intra;
    if (c.keyframe_mode == RIV_KEYFRAME_ALL_INTRA) intra = true;
    else if (c.keyframe_mode == RIV_KEYFRAME_AUTOMATIC) intra = e->frame_count % 30 == 0;
    else intra = e->frame_count % c.keyframe_interval == 0;
    if (!e->have_ref) intra = true;

    const int si = (int)(e->frame_count % e->slots.size());
    Slot& s = e->slots[si];
    {   // wait for the slot's previous frame to leave the host stage
        std::unique_lock<std::mutex> lk(e->mu);
        e->cv.wait(lk, [&] { return !s.busy || e->err_code; });
        if (e->err_code) { set_error(e->err_msg); return e->err_code; }
        s.index = e->frame_count; s.ts = ts; s.intra = intra; s.busy = true; s.coded = false; s.modeled = false; s.packed = false;
        s.gpu_ready = false;
        s.failed = false; s.res_failed = false; s.mv_failed = false;
        s.gpu_model = false;
    }
    s.h_passes.as<int>()[1] = 0;   // the search's error word: the kernel writes it only on failure
    s.t_submit = e->now_ms();
    s.n_marks = 0;
    // profiling events on every prof_every-th frame (a sample: each event holds the stream)
    s.prof = e->frame_count % e->prof_every == 0 ? e->profiling : 0;
    if (s.prof && !s.marks[0])
        for (auto& m : s.marks) RIV_HIP(hipEventCreateWithFlags(&m, hipEventDisableSystemFence));

    f.si = si; f.intra = intra; f.st = st;
    const size_t ysz = (size_t)e->W * e->H, csz = ysz / 4;
    f.cY = e->cur.as<uint8_t>(); f.cU = f.cY + ysz; f.cV = f.cU + csz;
    f.rY = e->ref[e->ref_idx].as<uint8_t>(); f.rU = f.rY + ysz; f.rV = f.rU + csz;
    f.nY = e->ref[1 - e->ref_idx].as<uint8_t>(); f.nU = f.nY + ysz; f.nV = f.nU + csz;
    f.pY = e->pred.as<uint8_t>(); f.pU = f.pY + ysz; f.pV = f.pU + csz;
    f.prev_mv = e->prev_inter_slot >= 0 ? (const int8_t*)e->slots[e->prev_inter_slot].rec.p + 4 : nullptr;

    s.timed = s.prof >= 2;
    if (s.timed) RIV_HIP(hipEventRecord(s.t0, st));
    const uint8_t* d_rgb = (const uint8_t*)rgb;
    int ub = -1;   // the upload buffer this frame reads, if any
    if (!on_device && st == e->st) {
        if (!e->up) {
            RIV_HIP(hipStreamCreateWithFlags(&e->up, hipStreamNonBlocking));
            for (int b = 0; b < 2; b++) {
                RIV_HIP(hipEventCreateWithFlags(&e->up_ready[b], hipEventDisableTiming | hipEventDisableSystemFence));
                RIV_HIP(hipEventCreateWithFlags(&e->up_free[b], hipEventDisableTiming | hipEventDisableSystemFence));
            }
        }
        ub = e->up_idx;
        e->up_idx ^= 1;
        uint8_t* buf = e->rgb.as<uint8_t>() + (size_t)ub * len;
        if (e->up_used[ub]) RIV_HIP(hipStreamWaitEvent(e->up, e->up_free[ub], 0));   // read by frame k-2
        RIV_HIP(hipMemcpyAsync(buf, rgb, len, hipMemcpyHostToDevice, e->up));
        RIV_HIP(hipEventRecord(e->up_ready[ub], e->up));
        RIV_HIP(hipStreamWaitEvent(st, e->up_ready[ub], 0));
        // the caller may reuse its frame once this returns: wait for the copy (from pinned memory
        // it is asynchronous; it waits for nothing but frame k-2's conversion)
        RIV_HIP(hipEventSynchronize(e->up_ready[ub]));
        d_rgb = buf;
    } else if (!on_device) {
        uint8_t* buf = e->rgb.as<uint8_t>() + 2 * len;
        RIV_HIP(hipMemcpyAsync(buf, rgb, len, hipMemcpyHostToDevice, st));
        d_rgb = buf;
    }
    mark(e, s, st, -1);
    rivk::rgb_to_yuv420(st, d_rgb, (int)w, (int)h, e->W, e->H, f.cY, f.cU, f.cV);
    if (ub >= 0) {
        RIV_HIP(hipEventRecord(e->up_free[ub], st));
        e->up_used[ub] = true;
    }
    if (!search_marks_by_launch(e, f)) mark(e, s, st, 0);
    if (intra) {
        rivk::intra(st, f.cY, f.cU, f.cV, e->W, e->H, e->yt.as<uint16_t>(), e->ct.as<uint16_t>(), f.nY, f.nU, f.nV, s.h_pack.dev<uint8_t>());
        mark(e, s, st, 6);
        rivk::deblock(st, f.nY, f.nU, f.nV, e->W, e->H, deblock_level(e->intra_qs), true, nullptr);
        mark(e, s, st, 5);
    }
    RIV_HIP(hipGetLastError());
    return RIV_OK;
}

// The inter frame's part of a search launch.
static riv::SearchStream frame_search_stream(riv_encoder* e, const FrameCtx& f, int blk0) {
    riv::SearchStream S{};
    S.cur = f.cY; S.ref = f.rY; S.prev_mv = f.prev_mv; S.prev_stride = sizeof(BlockRec);
    Slot& s = e->slots[f.si];
    S.rec = s.rec.as<BlockRec>(); S.host_rec = s.h_rec.dev<BlockRec>(); S.passes = s.h_passes.dev<int>();
    S.rdo_bits = e->rdo.as<uint32_t>(); S.rdo_max = e->rdo_max;
    S.W = e->W; S.H = e->H; S.bw = e->bw; S.bh = e->bh; S.blk0 = blk0;
    S.lnv = e->lnv.as<float>();
    return S;
}

// Search parameters shared by every stream of a launch (from the launching encoder's config).
static riv::SearchParams search_params(riv_encoder* e) {
    const riv_encoder_config& c = e->cfg;
    riv::SearchParams p{};
    p.sr = c.search_range;
    p.bits = e->bits.as<double>(); p.lam2 = e->lambda * 2.0; p.c_zero = e->c_zero; p.c_pred = e->c_pred;
    p.lam_zero = e->lambda == 0.0;
    p.zthr = c.me_zero_mv_threshold; p.pthr = c.me_predictor_threshold;
    return p;
}

// Everything of an inter frame after its motion search.
static int frame_after_search(riv_encoder* e, const FrameCtx& f) {
    const riv_encoder_config& c = e->cfg;
    Slot& s = e->slots[f.si];
    hipStream_t st = f.st;
    BlockRec* rec = s.rec.as<BlockRec>();   // written by the search's mode-decision epilogue
    if (!search_marks_by_launch(e, f)) mark(e, s, st, 1);
    // the prediction's variances came from the search (lnv); the residual kernel forms the prediction
    rivk::aq(st, e->nb, e->inter_qs, e->lnv.as<float>(), nullptr);   // the mean; the residual kernel forms the steps
    mark(e, s, st, 3);
    // the residual symbols are modelled on the GPU for single-frame calls (the host codes them for
    // batch steps, and when the test seam asks it to)
    s.gpu_model = !f.batch && e->split_coding.load() == 0;
    int q = -1;   // the record buffer
    if (s.gpu_model) {
        if (!e->rm_host) {
            for (Slot& t : e->slots) {
                RIV_HIP(t.h_sym.alloc(2 * ((size_t)e->rm_cap_max + 8))); RIV_HIP(t.h_rmh.alloc(64));
                RIV_HIP(t.d_sym.alloc(2 * ((size_t)e->rm_cap_max + 8)));
                RIV_HIP(hipEventCreateWithFlags(&t.sym_ev, hipEventDisableTiming | hipEventBlockingSync));
            }
            RIV_HIP(hipStreamCreateWithFlags(&e->ms, hipStreamNonBlocking));
            RIV_HIP(hipStreamCreateWithFlags(&e->dl, hipStreamNonBlocking));
            for (int k = 0; k < 2; k++) {
                RIV_HIP(hipEventCreateWithFlags(&e->rm_ready[k], hipEventDisableTiming | hipEventDisableSystemFence));
                RIV_HIP(hipEventCreateWithFlags(&e->rm_free[k], hipEventDisableTiming | hipEventDisableSystemFence));
            }
            e->rm_host = true;
        }
        q = e->rm_q;
        e->rm_q ^= 1;
        // frame k - 2's modelling has read it (a wait packet holds the stream for microseconds even
        // when its event has completed, so it is queued only while the event is pending)
        if (e->rm_used[q] && hipEventQuery(e->rm_free[q]) != hipSuccess) RIV_HIP(hipStreamWaitEvent(st, e->rm_free[q], 0));
    }
    // GPU-modelled frames keep their coefficient records on the device (the host gets them only if
    // it has to code the frame itself); host-coded frames get them in mapped host memory
    s.rm_q = q;
    rivk::inter_residual(st, f.cY, f.cU, f.cV, f.rY, f.rU, f.rV, f.pY, f.pU, f.pV, e->W, e->H, rec, e->lnv.as<float>(), e->inter_qs,
                         e->steps.as<float>(), c.inter_dead_zone,
                         f.nY, f.nU, f.nV,
                         q >= 0 ? nullptr : s.h_main.dev<uint8_t>(), q >= 0 ? nullptr : s.h_pack.dev<uint8_t>(), e->mask.as<uint8_t>(),
                         s.h_optskip.dev<uint8_t>(), q >= 0 ? e->rm_pack[q].as<uint8_t>() : nullptr, q >= 0 ? e->rm_coded[q].as<uint8_t>() : nullptr);
    mark(e, s, st, 4);
    rivk::deblock(st, f.nY, f.nU, f.nV, e->W, e->H, deblock_level(e->inter_qs), false, e->mask.as<uint8_t>());
    mark(e, s, st, 5);
    if (s.gpu_model) {   // the frame ends on the modelling stream, which waits for all of its work here
        RIV_HIP(hipEventRecord(e->rm_ready[q], st));
        RIV_HIP(hipStreamWaitEvent(e->ms, e->rm_ready[q], 0));
        rivk::ResModelBufs B = e->rm;
        B.pack = e->rm_pack[q].as<uint8_t>();
        B.coded = e->rm_coded[q].as<uint8_t>();
        B.sym = s.d_sym.as<uint16_t>();   // the slot's own: it is read by a copy the host stage starts
        RIV_HIP(rivk::res_model(e->ms, B, s.h_rmh.dev<uint32_t>(), s.h_pack.dev<uint8_t>()));
        RIV_HIP(hipEventRecord(e->rm_free[q], e->ms));
        e->rm_used[q] = true;
        mark(e, s, e->ms, 8);
    }
    RIV_HIP(hipGetLastError());
    return RIV_OK;
}

// Closes the frame's GPU work and hands it to the host stage (pipelined) or codes it (synchronous).
static int frame_end(riv_encoder* e, const FrameCtx& f, riv_frame_stats* stats) {
    Slot& s = e->slots[f.si];
    // the kernels wrote the records, skip flags and packed coefficients straight to the slot's
    // mapped host memory: the frame is ready for the host stage once the stream reaches here
    const hipStream_t last = !f.intra && s.gpu_model ? e->ms : f.st;   // the frame's last work
    if (s.timed) RIV_HIP(hipEventRecord(s.t1, last));
    RIV_HIP(hipEventRecord(s.done, last));
    e->prev_inter_slot = f.intra ? -1 : f.si;
    e->ref_idx = 1 - e->ref_idx;
    e->have_ref = true;
    e->frame_count++;

    if (e->hs) {
        { std::lock_guard<std::mutex> lk(e->hs->mu); e->hs->q.push_back({e, f.si}); }
        e->hs->cv.notify_one();
        if (stats) { std::memset(stats, 0, sizeof(*stats)); stats->frame_index = s.index; stats->timestamp_ms = s.ts; stats->frame_type = f.intra ? 1 : 2; }
        return RIV_OK;
    }
    if (hipEventSynchronize(s.done) != hipSuccess) { e->fail(RIV_E_GPU, "GPU frame failed"); start_host_parts(e, f.si, true); }
    else if (const char* msg = search_error(s)) { e->fail(RIV_E_GPU, msg); start_host_parts(e, f.si, true); }
    else { s.t_done = e->now_ms(); collect_gpu_times(s); start_host_parts(e, f.si, false); }
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->err_code) { set_error(e->err_msg); return e->err_code; }
    if (stats && !e->done_stats.empty()) *stats = e->done_stats.back();
    e->done_stats.clear();
    return RIV_OK;
}

static int encode_impl(riv_encoder* e, const void* rgb, bool on_device, size_t len, uint32_t w, uint32_t h, uint64_t ts, riv_frame_stats* stats) {
    FrameCtx f;
    int r = frame_begin(e, rgb, on_device, len, w, h, ts, e->st, f);
    if (r) return r;
    if (!f.intra) {
        riv::SearchParams p = search_params(e);
        p.s0 = frame_search_stream(e, f, 0);
        p.streams = nullptr; p.nstreams = 1; p.nb = e->nb;
        p.field = e->field.as<uint32_t>(); p.arrive = e->arrive.as<unsigned>(); p.bank = e->s_bank; p.maxp = e->s_maxp;
        p.nslot = e->shape.nslot; p.nres = e->shape.nres;
        if (e->net) {   // learned form: the network's proposals for every block seed the search
            const rivnet::Source src{f.cY, f.rY, f.prev_mv, (int)sizeof(BlockRec), e->W, e->H, e->bw, e->bh};
            RIV_HIP(rivnet::infer(f.st, nullptr, src, nullptr, e->nb, rivnet::net_weights(e->net), e->net_logit, e->cfg.search_range,
                                  e->prop.as<uint4>(), nullptr));
            RIV_HIP(rivnet::note_use(e->net, f.st));
            p.prop = e->prop.as<uint4>();
            p.lstat = e->lstat.as<unsigned long long>();
        }
        e->s_bank ^= 1;
        p.plimit = e->pass_cap > 0 ? std::min(p.maxp, e->pass_cap) : p.maxp;
        if (e->net && e->learned_passes > 0) {   // learned form: a fixed number of passes
            p.plimit = std::min(p.plimit, e->learned_passes);
            p.truncate = 1;
        }
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        if (search_marks_by_launch(e, f)) {
            Slot& s = e->slots[f.si];
            ev0 = mark_by_launch(e, s, 0);
            ev1 = mark_by_launch(e, s, 1);
        }
        int lr = launch_search(e->cfg.device_id, f.st, p, e->shape.groups, e->shape.waves, ev0, ev1);
        if (lr) return lr;
        r = frame_after_search(e, f);
        if (r) return r;
    }
    return frame_end(e, f, stats);
}

// One frame for each of n encoders; their GPU work is queued on the first encoder's stream and the
// motion searches of all inter frames run as one launch.
static int batch_step(riv_encoder* const* encs, size_t n, const void* const* rgb, const size_t* len, const uint64_t* ts, bool on_device,
                      riv_frame_stats* stats) {
    riv_encoder* e0 = encs[0];
    const riv_encoder_config& c0 = e0->cfg;
    for (size_t i = 0; i < n; i++) {
        riv_encoder* e = encs[i];
        if (!e || !rgb[i]) return RIV_E_ARG;
        for (size_t j = 0; j < i; j++) if (encs[j] == e) { set_error("an encoder appears twice in a batch step"); return RIV_E_ARG; }
        const riv_encoder_config& c = e->cfg;
        if (c.device_id != c0.device_id || c.search_range != c0.search_range || e->lambda != e0->lambda ||
            c.me_zero_mv_threshold != c0.me_zero_mv_threshold || c.me_predictor_threshold != c0.me_predictor_threshold) {
            set_error("batched encoders must share device, search range, inter quality / lambda and ME thresholds");
            return RIV_E_ARG;
        }
        int r = check_frame(e, len[i], c.display_width, c.display_height);
        if (r) return r;
    }
    RIV_HIP(hipSetDevice(c0.device_id));
    hipStream_t st = e0->st;
    for (size_t i = 1; i < n; i++) {   // the batch stream follows each encoder's earlier work
        RIV_HIP(hipEventRecord(encs[i]->sync_ev, encs[i]->st));
        RIV_HIP(hipStreamWaitEvent(st, encs[i]->sync_ev, 0));
    }
    std::vector<FrameCtx> f(n);
    std::vector<riv::SearchStream> desc;
    int total = 0, maxp = 0;
    for (size_t i = 0; i < n; i++) {
        riv_encoder* e = encs[i];
        f[i].batch = true;
        int r = frame_begin(e, rgb[i], on_device, len[i], e->cfg.display_width, e->cfg.display_height, ts ? ts[i] : 0, st, f[i]);
        if (r) return r;
        if (!f[i].intra) {
            desc.push_back(frame_search_stream(e, f[i], total));
            total += e->nb;
            maxp = std::max(maxp, e->s_maxp);
        }
    }
    if (!on_device) {   // the callers' host frames may be reused once the step returns
        if (!e0->b_copy_ev) RIV_HIP(hipEventCreateWithFlags(&e0->b_copy_ev, hipEventDisableTiming));
        RIV_HIP(hipEventRecord(e0->b_copy_ev, st));
        RIV_HIP(hipEventSynchronize(e0->b_copy_ev));
    }
    if (!desc.empty()) {
        riv::SearchParams p = search_params(e0);
        p.nstreams = (int)desc.size(); p.nb = total; p.maxp = maxp;
        // launch shape for the whole batch: the autotuner's choice for this batch's key
        ShapeTuner& T = e0->b_tuners[std::make_tuple(total, p.nstreams, p.sr)];
        if (T.cands.empty()) {
            hipDeviceProp_t prop;
            RIV_HIP(hipGetDeviceProperties(&prop, c0.device_id));
            for (const rivk::SearchShape& c : rivk::search_shapes(p.sr, total, p.nstreams, prop.multiProcessorCount, prop.sharedMemPerBlock))
                T.cands.push_back(ShapeTuner::Cand{c});
            if (T.cands.empty()) {
                set_error("batch has more macroblocks or streams than one motion-search launch can hold");
                return RIV_E_UNSUPPORTED;
            }
            if (!e0->b_autotune || T.cands.size() == 1) T.best = 0;
        }
        T.collect();
        const int ci = T.next();
        T.last = ci >= 0 ? ci : T.best;
        e0->b_last_tuner = &T;
        const rivk::SearchShape sh = T.cands[T.last].sh;
        p.nslot = sh.nslot; p.nres = sh.nres;
        // scratch sized for the batch (grown when a larger batch comes; the stream is drained first)
        const size_t need_field = 2 * (size_t)total, need_arrive = 2 * (size_t)riv::barrier_words(maxp);
        if (need_field > e0->b_field_n || need_arrive > e0->b_arrive_n || desc.size() > e0->b_desc_n) {
            RIV_HIP(hipStreamSynchronize(st));
            if (need_field > e0->b_field_n) {
                RIV_HIP(e0->b_field.alloc(4 * need_field)); RIV_HIP(hipMemset(e0->b_field.p, 0, 4 * need_field));
                e0->b_field_n = need_field;
            }
            if (need_arrive > e0->b_arrive_n) {
                RIV_HIP(e0->b_arrive.alloc(4 * need_arrive)); RIV_HIP(hipMemset(e0->b_arrive.p, 0, 4 * need_arrive));
                e0->b_arrive_n = need_arrive; e0->b_bank = 0;
            }
            if (desc.size() > e0->b_desc_n) {
                const size_t cap = std::max<size_t>(desc.size(), 64);
                RIV_HIP(e0->b_desc.alloc(cap * sizeof(riv::SearchStream)));
                for (int k = 0; k < 2; k++) {
                    RIV_HIP(e0->b_stage_buf[k].alloc(cap * sizeof(riv::SearchStream)));
                    if (!e0->b_stage_ev[k]) RIV_HIP(hipEventCreateWithFlags(&e0->b_stage_ev[k], hipEventDisableTiming));
                }
                e0->b_desc_n = cap;
            }
        }
        // the counter banks are laid out for one pass limit: clear them when it changes
        if (maxp != e0->b_maxp) {
            RIV_HIP(hipMemsetAsync(e0->b_arrive.p, 0, 4 * need_arrive, st));
            e0->b_maxp = maxp; e0->b_bank = 0;
        }
        // the descriptors go up through a pinned staging buffer that no pending copy still reads
        const int k = e0->b_stage;
        e0->b_stage ^= 1;
        RIV_HIP(hipEventSynchronize(e0->b_stage_ev[k]));
        std::memcpy(e0->b_stage_buf[k].p, desc.data(), desc.size() * sizeof(riv::SearchStream));
        RIV_HIP(hipMemcpyAsync(e0->b_desc.p, e0->b_stage_buf[k].p, desc.size() * sizeof(riv::SearchStream), hipMemcpyHostToDevice, st));
        RIV_HIP(hipEventRecord(e0->b_stage_ev[k], st));
        p.streams = e0->b_desc.as<riv::SearchStream>();
        p.field = e0->b_field.as<uint32_t>(); p.arrive = e0->b_arrive.as<unsigned>(); p.bank = e0->b_bank;
        // learned form: each learned stream's proposals at its first block; the other streams' zero
        bool any_net = false;
        for (size_t i = 0; i < n; i++) any_net |= !f[i].intra && encs[i]->net;
        if (any_net) {
            if ((size_t)total > e0->b_prop_n) {
                RIV_HIP(hipStreamSynchronize(st));
                RIV_HIP(e0->b_prop.alloc(16 * (size_t)total));
                e0->b_prop_n = total;
            }
            if (!e0->b_lstat.p) {   // three counters for each wave any launch shape can have
                const size_t lb = 3 * sizeof(unsigned long long) * riv::kMaxSearchWaves;
                RIV_HIP(e0->b_lstat.alloc(lb)); RIV_HIP(hipMemset(e0->b_lstat.p, 0, lb));
            }
            uint4* bp = e0->b_prop.as<uint4>();
            std::vector<const riv_mvnet*> used;   // networks read on the batch stream, each noted once
            for (size_t i = 0, d = 0; i < n; i++) {
                if (f[i].intra) continue;
                riv_encoder* e = encs[i];
                const riv::SearchStream& S = desc[d++];
                if (e->net) {
                    const rivnet::Source src{S.cur, S.ref, f[i].prev_mv, (int)sizeof(BlockRec), e->W, e->H, e->bw, e->bh};
                    RIV_HIP(rivnet::infer(st, nullptr, src, nullptr, e->nb, rivnet::net_weights(e->net), e->net_logit, p.sr, bp + S.blk0, nullptr));
                    if (std::find(used.begin(), used.end(), e->net) == used.end()) used.push_back(e->net);
                } else {
                    RIV_HIP(hipMemsetAsync(bp + S.blk0, 0, 16 * (size_t)e->nb, st));
                }
            }
            for (const riv_mvnet* net : used) RIV_HIP(rivnet::note_use(net, st));
            p.prop = bp;
            // counters only while the launch fits the counter array (every MI355X shape does)
            p.lstat = sh.groups * sh.waves <= riv::kMaxSearchWaves ? e0->b_lstat.as<unsigned long long>() : nullptr;
        }
        e0->b_bank ^= 1;
        hipEvent_t ta = ci >= 0 ? T.event() : nullptr, tb = ci >= 0 ? T.event() : nullptr;
        if (ta && tb) RIV_HIP(hipEventRecord(ta, st));
        collect_batch_search_times(e0);
        hipEvent_t pa = nullptr, pb = nullptr;   // profiling: the whole launch, timed into stage 7
        if (e0->profiling && ci < 0) {
            pa = batch_prof_event(e0); pb = batch_prof_event(e0);
            if (pa && pb) RIV_HIP(hipEventRecord(pa, st));
        }
        p.plimit = e0->pass_cap > 0 ? std::min(p.maxp, e0->pass_cap) : p.maxp;
        {   // a fixed number of passes only when every inter stream of the launch is learned with one
            int lp = -1;
            for (size_t i = 0; i < n && lp != 0; i++) {
                if (f[i].intra) continue;
                const int k = encs[i]->net ? encs[i]->learned_passes : 0;
                lp = lp < 0 ? k : (lp == k ? k : 0);
            }
            if (lp > 0) { p.plimit = std::min(p.plimit, lp); p.truncate = 1; }
        }
        int lr = launch_search(c0.device_id, st, p, sh.groups, sh.waves);
        if (lr) return lr;
        if (pa && pb) {
            RIV_HIP(hipEventRecord(pb, st));
            e0->b_prof.push_back({pa, pb});
        } else {
            if (pa) e0->b_prof_spare.push_back(pa);
            if (pb) e0->b_prof_spare.push_back(pb);
        }
        if (ta && tb) {
            RIV_HIP(hipEventRecord(tb, st));
            T.pending.push_back({ci, ta, tb});
            T.cands[ci].issued++;
        } else {
            if (ta) T.spare.push_back(ta);
            if (tb) T.spare.push_back(tb);
        }
    }
    int status = RIV_OK;
    for (size_t i = 0; i < n; i++) {
        riv_encoder* e = encs[i];
        int r = f[i].intra ? RIV_OK : frame_after_search(e, f[i]);
        if (!r) r = frame_end(e, f[i], stats ? &stats[i] : nullptr);
        if (r && !status) status = r;
    }
    for (size_t i = 1; i < n; i++) {   // each encoder's own stream follows the batch
        RIV_HIP(hipEventRecord(encs[i]->sync_ev, st));
        RIV_HIP(hipStreamWaitEvent(encs[i]->st, encs[i]->sync_ev, 0));
    }
    return status;
}

int riv_encode_frame(riv_encoder* e, const uint8_t* rgb, size_t len, uint32_t w, uint32_t h, uint64_t ts, riv_frame_stats* stats) {
    if (!e || !rgb) return RIV_E_ARG;
    return encode_impl(e, rgb, false, len, w, h, ts, stats);
}
int riv_encode_frame_device(riv_encoder* e, const void* d_rgb, size_t len, uint32_t w, uint32_t h, uint64_t ts, riv_frame_stats* stats) {
    if (!e || !d_rgb) return RIV_E_ARG;
    return encode_impl(e, d_rgb, true, len, w, h, ts, stats);
}

int riv_encode_batch_step(riv_encoder* const* encs, size_t n, const void* const* rgb, const size_t* len, const uint64_t* timestamp_ms,
                          int on_device, riv_frame_stats* stats) {
    if (!encs || !n || !rgb || !len) return RIV_E_ARG;
    return batch_step(encs, n, rgb, len, timestamp_ms, on_device != 0, stats);
}

int riv_encoder_set_batch_autotune(riv_encoder* e, int on) {
    if (!e) return RIV_E_ARG;
    e->b_autotune = on != 0;
    return RIV_OK;
}

int riv_encoder_batch_tuning(riv_encoder* e, int32_t* out) {
    if (!e || !out) return RIV_E_ARG;
    const ShapeTuner* T = e->b_last_tuner;
    if (!T || T->last < 0) { set_error("no batch step has been launched with this encoder first"); return RIV_E_ARG; }
    const rivk::SearchShape& sh = T->cands[T->last].sh;
    out[0] = sh.waves; out[1] = sh.groups; out[2] = sh.nslot; out[3] = (int32_t)T->cands.size(); out[4] = T->best >= 0 ? 1 : 0;
    out[5] = T->best >= 0 ? T->best : -1;
    return RIV_OK;
}

int riv_encoder_flush(riv_encoder* e) {
    if (!e) return RIV_E_ARG;
    // The encoder's streams first: this thread waits for the GPU while the host stage codes the
    // frames already done, so no work of this encoder is left for a later device-wide wait (the
    // first one after a run of frames costs ~0.15 ms, which this wait takes beside the coding).
    // The waits sleep (blocking-sync events), leaving the cores to the host stage; then each
    // stream is synchronised, now quickly, which settles the runtime's own records of its work
    // (waiting on events alone leaves that to the next device-wide wait). Each frame's search
    // failures arrive through its own error word, read by the host stage.
    RIV_HIP(hipSetDevice(e->cfg.device_id));
    const hipStream_t ss[4] = {e->st, e->ms, e->up, e->dl};
    for (int k = 0; k < 4; k++) {
        if (!ss[k]) continue;
        if (!e->fin_ev[k]) RIV_HIP(hipEventCreateWithFlags(&e->fin_ev[k], hipEventDisableTiming | hipEventBlockingSync));
        RIV_HIP(hipEventRecord(e->fin_ev[k], ss[k]));
    }
    for (int k = 0; k < 4; k++)
        if (ss[k]) RIV_HIP(hipEventSynchronize(e->fin_ev[k]));
    for (int k = 0; k < 4; k++)
        if (ss[k]) RIV_HIP(hipStreamSynchronize(ss[k]));
    std::unique_lock<std::mutex> lk(e->mu);
    e->cv.wait(lk, [&] { return e->completed == e->frame_count || e->err_code; });
    if (e->err_code) { set_error(e->err_msg); return e->err_code; }
    return RIV_OK;
}

int riv_encoder_pop_stats(riv_encoder* e, riv_frame_stats* out, size_t cap, size_t* n) {
    if (!e || !n || (!out && cap)) return RIV_E_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    size_t k = 0;
    while (k < cap && !e->done_stats.empty()) { out[k++] = e->done_stats.front(); e->done_stats.pop_front(); }
    *n = k;
    return RIV_OK;
}

int riv_encoder_finish(riv_encoder* e) {   // encoder.rs:204-221, :950-961
    if (!e) return RIV_E_ARG;
    int r = riv_encoder_flush(e);
    if (r) return r;
    std::lock_guard<std::mutex> lk(e->mu);
    uint64_t end = e->position();
    r = e->seek(e->header_pos);
    if (!r) r = write_header(e);
    if (!r) r = e->seek(end);
    if (!r && !e->mem && e->wv.flush(e->wuser) != 0) r = RIV_E_IO;
    return r;
}
uint64_t riv_encoder_frame_count(const riv_encoder* e) { return e ? e->frame_count : 0; }
int riv_encoder_header(const riv_encoder* e, riv_video_header* out) {
    if (!e || !out) return RIV_E_ARG;
    *out = e->hdr;
    return RIV_OK;
}
int riv_encoder_get_config(const riv_encoder* e, riv_encoder_config* out) {
    if (!e || !out) return RIV_E_ARG;
    *out = e->cfg;
    return RIV_OK;
}
int riv_encoder_output(const riv_encoder* e, const uint8_t** data, size_t* len) {
    if (!e || !e->mem || !data || !len) return RIV_E_ARG;
    *data = e->membuf.data(); *len = e->membuf.size();
    return RIV_OK;
}
void riv_encoder_destroy(riv_encoder* e) { destroy_encoder(e); }

int riv_encoder_set_mvnet(riv_encoder* e, riv_mvnet* net, float confidence) {
    if (!e) return RIV_E_ARG;
    if (net && rivnet::net_device(net) != e->cfg.device_id) { set_error("the network lives on another device"); return RIV_E_ARG; }
    if (net && !(confidence > 0.0f && confidence <= 1.0f)) { set_error("confidence must be in (0, 1]"); return RIV_E_ARG; }
    e->net = net;
    e->net_logit = net ? rivnet::conf_logit(confidence) : 0.0f;
    return RIV_OK;
}

int riv_encoder_set_learned_passes(riv_encoder* e, int32_t passes) {
    if (!e || passes < 0) return RIV_E_ARG;
    e->learned_passes = passes;
    return RIV_OK;
}

int riv_encoder_learned_stats(riv_encoder* e, uint64_t* counts) {
    if (!e || !counts) return RIV_E_ARG;
    RIV_HIP(hipSetDevice(e->cfg.device_id));
    RIV_HIP(hipStreamSynchronize(e->st));
    for (int i = 0; i < 3; i++) counts[i] = 0;
    for (const DevBuf* b : {&e->lstat, &e->b_lstat}) {
        if (!b->p) continue;
        std::vector<unsigned long long> h(b->n / sizeof(unsigned long long));
        RIV_HIP(hipMemcpy(h.data(), b->p, b->n, hipMemcpyDeviceToHost));
        for (size_t w = 0; w + 2 < h.size(); w += 3)
            for (int i = 0; i < 3; i++) counts[i] += h[w + i];
    }
    return RIV_OK;
}

int riv_encoder_debug_pass_limit(riv_encoder* e, int32_t max_passes) {
    if (!e || max_passes < 0) return RIV_E_ARG;
    e->pass_cap = max_passes;
    return RIV_OK;
}

int riv_encoder_debug_model_cap(riv_encoder* e, uint32_t cap) {
    if (!e) return RIV_E_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    e->rm.cap = cap == 0 ? e->rm_cap_max : std::min(cap, e->rm_cap_max);
    return RIV_OK;
}

int riv_encoder_debug_split_coding(riv_encoder* e, int32_t mode) {
    if (!e || mode < 0 || mode > 3) return RIV_E_ARG;
    e->split_coding = mode;
    return RIV_OK;
}

int riv_encoder_set_profiling(riv_encoder* e, int on) {
    if (!e) return RIV_E_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    e->profiling = on < 0 ? 0 : on > 2 ? 2 : on;
    for (int i = 0; i < RIV_NUM_STAGES; i++) { e->stage_ms[i] = 0; e->stage_n[i] = 0; }
    return RIV_OK;
}
int riv_encoder_set_profiling_sample(riv_encoder* e, uint32_t every) {
    if (!e || every == 0) return RIV_E_ARG;
    std::lock_guard<std::mutex> lk(e->mu);
    e->prof_every = every;
    return RIV_OK;
}
int riv_encoder_stage_times(riv_encoder* e, double* ms, uint32_t* counts) {
    if (!e || !ms || !counts) return RIV_E_ARG;
    if (!e->b_prof.empty()) {
        RIV_HIP(hipSetDevice(e->cfg.device_id));
        collect_batch_search_times(e);
    }
    std::lock_guard<std::mutex> lk(e->mu);
    for (int i = 0; i < RIV_NUM_STAGES; i++) { ms[i] = e->stage_ms[i]; counts[i] = e->stage_n[i]; }
    return RIV_OK;
}

// ---- last completed frame introspection (meant for the synchronous mode) ----
int riv_encoder_debug_blocks(riv_encoder* e, int8_t* rec12, int64_t* sad, uint8_t* rdo_skip, uint8_t* opt_skip) {
    if (!e || e->last_slot < 0 || e->slots[e->last_slot].intra) return RIV_E_ARG;
    const Slot& s = e->slots[e->last_slot];
    const BlockRec* hr = s.h_rec.as<BlockRec>();
    for (int i = 0; i < e->nb; i++) {
        int8_t* r = rec12 + 12 * i;
        std::memcpy(r, hr[i].s, 4); std::memcpy(r + 4, hr[i].r, 4);
        r[8] = (int8_t)hr[i].mode; r[9] = (int8_t)hr[i].new_base; r[10] = hr[i].raw_dx; r[11] = hr[i].raw_dy;
        if (sad) sad[i] = hr[i].sad;
        if (rdo_skip) rdo_skip[i] = hr[i].rdo_skip;
        if (opt_skip) opt_skip[i] = s.h_optskip.as<uint8_t>()[i];
    }
    return RIV_OK;
}
int riv_encoder_debug_yuv(riv_encoder* e, int which, uint8_t* y, uint8_t* u, uint8_t* v) {
    if (!e || which < 0 || which > 2) return RIV_E_ARG;
    RIV_HIP(hipSetDevice(e->cfg.device_id));
    RIV_HIP(hipStreamSynchronize(e->st));
    const size_t ysz = (size_t)e->W * e->H, csz = ysz / 4;
    const uint8_t* src = which == 0 ? e->cur.as<uint8_t>() : which == 1 ? e->pred.as<uint8_t>() : e->ref[e->ref_idx].as<uint8_t>();
    RIV_HIP(hipMemcpy(y, src, ysz, hipMemcpyDeviceToHost));
    RIV_HIP(hipMemcpy(u, src + ysz, csz, hipMemcpyDeviceToHost));
    RIV_HIP(hipMemcpy(v, src + ysz + csz, csz, hipMemcpyDeviceToHost));
    return RIV_OK;
}
int riv_encoder_debug_coeffs(riv_encoder* e, int16_t* out) {
    if (!e || e->last_slot < 0) return RIV_E_ARG;
    const Slot& s = e->slots[e->last_slot];
    const uint8_t* os = s.h_optskip.as<uint8_t>();
    std::vector<uint8_t> dev;   // a GPU-modelled frame's records, from the device (the last frame's are intact)
    if (!s.intra && s.gpu_model) {
        RIV_HIP(hipSetDevice(e->cfg.device_id));
        RIV_HIP(hipStreamSynchronize(e->st));
        dev.resize((size_t)riv::PACK_STRIDE * e->nb);
        RIV_HIP(hipMemcpy(dev.data(), e->rm_pack[s.rm_q].p, dev.size(), hipMemcpyDeviceToHost));
    }
    for (int b = 0; b < e->nb; b++) {
        int16_t* q = out + (size_t)b * 384;
        if (!s.intra && os[b]) std::memset(q, 0, 768);
        else if (!dev.empty()) rivh::unpack_coeffs(dev.data() + (size_t)b * riv::PACK_STRIDE, q);
        else if (!s.intra) rivh::unpack_split(s.h_main.as<uint8_t>(), s.h_pack.as<uint8_t>(), b, q);
        else rivh::unpack_coeffs(s.h_pack.as<uint8_t>() + (size_t)b * riv::PACK_STRIDE, q);
    }
    return RIV_OK;
}
int riv_encoder_debug_aq(riv_encoder* e, float* steps) {
    if (!e || e->last_slot < 0 || e->slots[e->last_slot].intra) return RIV_E_ARG;
    RIV_HIP(hipSetDevice(e->cfg.device_id));
    RIV_HIP(hipStreamSynchronize(e->st));
    RIV_HIP(hipMemcpy(steps, e->steps.p, 4 * (size_t)e->nb, hipMemcpyDeviceToHost));
    return RIV_OK;
}

const char* riv_last_error(void) { return rivhost::last_error(); }

}  // extern "C"
