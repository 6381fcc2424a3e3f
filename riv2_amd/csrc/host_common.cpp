// Host-side helpers shared by the encoder and decoder drivers (see host_common.hpp).
#include "host_common.hpp"
#include "riv_kernels.hpp"
#include <cmath>
#include <mutex>
#include <algorithm>

namespace rivhost {

static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

int ensure_device(int device) {
    static std::mutex mu;
    static std::vector<bool> ready;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        set_error(std::string("no HIP device available: ") + hipGetErrorString(e));
        return RIV_E_GPU;
    }
    if (device < 0 || device >= n) { set_error("device_id out of range"); return RIV_E_ARG; }
    RIV_HIP(hipSetDevice(device));
    std::lock_guard<std::mutex> lk(mu);
    if ((int)ready.size() < n) ready.resize(n, false);
    if (!ready[device]) {
        // DCT bases round(2^18 * c_k * cos((2n+1) k pi / 2N)): entry for entry the reference's
        // FDCT_TRANSPOSE_16 / IDCT_ROWS_16 (transposed) / FDCT_TRANSPOSE_8 (reitero_dct/src/dct16.rs:7-145, dct8.rs:6-25)
        int32_t b16[16][16], b8[8][8];
        for (int i = 0; i < 16; i++)
            for (int k = 0; k < 16; k++)
                b16[i][k] = (int32_t)std::llround(262144.0 * (k == 0 ? std::sqrt(1.0 / 16) : std::sqrt(2.0 / 16)) * std::cos((2 * i + 1) * k * M_PI / 32.0));
        for (int i = 0; i < 8; i++)
            for (int k = 0; k < 8; k++)
                b8[i][k] = (int32_t)std::llround(262144.0 * (k == 0 ? std::sqrt(1.0 / 8) : std::sqrt(2.0 / 8)) * std::cos((2 * i + 1) * k * M_PI / 16.0));
        RIV_HIP(rivk::upload_bases(&b16[0][0], &b8[0][0]));
        ready[device] = true;
    }
    return RIV_OK;
}

float quant_step(uint8_t q) {
    if (q == 100) return 1.0f;
    if (q == 1) return 112.0f;
    const float range = 112.0f - 1.0f;
    return 1.0f + (100.0f - (float)q) / 99.0f * range;
}

int deblock_level(float qs) {
    float r = std::round(qs * 0.6f);
    r = std::min(63.0f, std::max(0.0f, r));
    return (int)r;
}

static const float kJpegLuma[64] = {16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55,
                                    14, 13, 16, 24, 40, 57, 69, 56, 14, 17, 22, 29, 51, 87, 80, 62,
                                    18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
                                    49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
static const float kJpegChroma[64] = {17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
                                      24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
                                      99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
                                      99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
static uint16_t to_u16(float f) { return f <= 0.0f ? 0 : (f >= 65535.0f ? 65535 : (uint16_t)f); }

void intra_tables(float qs, uint16_t* yt, uint16_t* ct) {
    // 8x8 JPEG table bilinearly upscaled to 16x16 in f32, scaled so DC == qs (residual.rs:58-85)
    for (int v = 0; v < 16; v++)
        for (int u = 0; u < 16; u++) {
            float fx = (float)u * 7.0f / 15.0f, fy = (float)v * 7.0f / 15.0f;
            int x0 = (int)fx, y0 = (int)fy, x1 = std::min(x0 + 1, 7), y1 = std::min(y0 + 1, 7);
            float dx = fx - (float)x0, dy = fy - (float)y0;
            float val = kJpegLuma[y0 * 8 + x0] * (1.0f - dx) * (1.0f - dy) + kJpegLuma[y0 * 8 + x1] * dx * (1.0f - dy) +
                        kJpegLuma[y1 * 8 + x0] * (1.0f - dx) * dy + kJpegLuma[y1 * 8 + x1] * dx * dy;
            yt[v * 16 + u] = to_u16(std::max(std::round((val / kJpegLuma[0]) * qs), 1.0f));
        }
    for (int i = 0; i < 64; i++) ct[i] = to_u16(std::max(std::round((kJpegChroma[i] / kJpegChroma[0]) * qs), 1.0f));
}

RdoModel::RdoModel(uint8_t quality, double mult, uint8_t thr) : skip_threshold(thr) {
    quality = std::min<uint8_t>(100, std::max<uint8_t>(1, quality));
    int qp = (int)std::round((double)(100u - quality) * 0.6);   // quality_to_qp (rdo.rs:75-80)
    qp = std::min(60, std::max(0, qp));
    lambda = mult * std::pow(2.0, ((double)qp - 12.0) / 3.0);   // lambda_from_qp (rdo.rs:82-86)
    double nq = (double)quality / 100.0, q = (double)quant_step(quality);
    slope = 0.5 / q;
    intercept = 5.0 + 0.56 * q;
    retention = 0.15 + ((q - 1.0) / 111.0) * 0.55;
    skip_bits = 0.55 - nq * 0.15;
    keep_bits = 1.20 + nq * 0.35;
}

bool RdoModel::skip(int64_t sad, bool aligned) const {
    if (skip_threshold == 0 || !aligned) return false;
    double satd = (double)std::max<int64_t>(sad, 0);
    if (satd > 768.0 * (double)skip_threshold) return false;   // 16*16*3 bytes per block
    double skip_cost = satd + lambda * skip_bits;
    double rate = keep_bits + std::max(intercept + slope * satd, 0.0);
    double coded = satd * retention + lambda * rate;
    return skip_cost <= coded;
}

Neighbours neighbours(const Mv4* cur, int known, const Mv4* prev, int bw, int bh, int bx, int by) {
    Neighbours n;
    const int idx = by * bw + bx;
    if (bx > 0 && idx - 1 < known) { n.has_l = true; n.l = cur[idx - 1]; }
    if (by > 0) {
        int t = idx - bw;
        if (t < known) { n.has_t = true; n.t = cur[t]; }
        if (bx + 1 < bw && t + 1 < known) { n.has_tr = true; n.tr = cur[t + 1]; }
        if (bx > 0 && t - 1 < known) { n.has_tl = true; n.tl = cur[t - 1]; }
    }
    if (prev) { n.has_tp = true; n.tp = prev[idx]; }
    (void)bh;
    Mv4 list[5]; int cnt = 0;
    auto add = [&](bool have, const Mv4& m) {
        if (!have) return;
        for (int i = 0; i < cnt; i++) if (list[i] == m) return;
        list[cnt++] = m;
    };
    add(n.has_l, n.l); add(n.has_t, n.t); add(n.has_tr, n.tr); add(n.has_tl, n.tl); add(n.has_tp, n.tp);
    Mv4 zero{};
    n.nearest = cnt > 0 ? list[0] : zero;
    n.nearp = cnt > 1 ? list[1] : zero;
    n.temporal = n.has_tp ? n.tp : zero;
    return n;
}

void put_header(std::vector<uint8_t>& b, const riv_video_header& h) {
    b.push_back('R'); b.push_back('I'); b.push_back('V'); b.push_back(0);
    put_u32(b, 5);   // RIV_VERSION (format.rs:9)
    put_u32(b, h.display_width); put_u32(b, h.display_height);
    put_u32(b, h.storage_width); put_u32(b, h.storage_height);
    put_u32(b, h.fps); put_u64(b, h.frame_count);
}

}  // namespace rivhost
