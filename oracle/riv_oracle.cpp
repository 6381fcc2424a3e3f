// Parity oracle for the RIV codec: a plain, single-file C++ restatement of the
// reference CPU encoder/decoder (reitero_video, Rust).
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or calls
// this file; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
// load liboracle_riv.so, as the checker and the CPU baseline.
//
// Every routine cites the reference file:line it restates (paths are relative to
// the reference tree, under reitero_video/ unless stated). Where the SPEC and the
// code disagree the code is followed. Build flags: -O2 -ffp-contract=off (Rust never
// contracts a*b+c into an FMA, and the RDO / MV-cost / AQ maths is float-exact).
//
// Known deliberate difference: RGB<->YUV420 in the reference goes through the
// third-party `yuv` crate 0.8.9 (BT.709, limited range, Balanced mode), which is not
// available here. This file uses the project's own BT.709 limited-range fixed-point
// conversion (rgb_to_yuv420 / yuv420_to_rgb below); parity of that step with the Rust
// reference is unpinned. Everything downstream of the YUV planes follows the reference.

#include <cstdint>
#include <cstring>
#include <cmath>
#include <cstdlib>
#include <vector>
#include <deque>
#include <map>
#include <tuple>
#include <algorithm>
#include <thread>
#include <stdexcept>

namespace {

// ---------------------------------------------------------------------------
// Rust numeric semantics helpers
// ---------------------------------------------------------------------------
template <typename T> static inline T clampv(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }
// Rust `f as i16` (saturating, NaN -> 0)
static inline int16_t f32_to_i16_sat(float f) {
    if (std::isnan(f)) return 0;
    if (f <= -32768.0f) return -32768;
    if (f >= 32767.0f) return 32767;
    return (int16_t)f;
}
static inline int32_t f32_to_i32_sat(float f) {
    if (std::isnan(f)) return 0;
    if (f <= -2147483648.0f) return INT32_MIN;
    if (f >= 2147483648.0f) return INT32_MAX;
    return (int32_t)f;
}
static inline uint16_t f32_to_u16_sat(float f) {
    if (std::isnan(f) || f <= 0.0f) return 0;
    if (f >= 65535.0f) return 65535;
    return (uint16_t)f;
}
static inline int64_t f64_to_i64_sat(double d) {
    if (std::isnan(d)) return 0;
    if (d <= -9223372036854775808.0) return INT64_MIN;
    if (d >= 9223372036854775808.0) return INT64_MAX;
    return (int64_t)d;
}
// wrapping i32 arithmetic (Rust release / `wide` lane semantics)
static inline int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static inline int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
static inline int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
static inline int32_t wshl(int32_t a, int s) { return (int32_t)((uint32_t)a << s); }

// ---------------------------------------------------------------------------
// MotionVector  (reitero_video_common/src/motion_vector.rs:3-117)
// ---------------------------------------------------------------------------
struct MV {
    int8_t dx = 0, dy = 0, sx = 0, sy = 0;
    bool skip = false;
    int dx_hp() const { return (int)dx * 2 + sx; }   // motion_vector.rs:68
    int dy_hp() const { return (int)dy * 2 + sy; }
    static int8_t clamp_sub(int v) { return (int8_t)(v < -1 ? -1 : (v > 1 ? 1 : v)); }
    static MV make(int8_t dx, int8_t dy, int8_t sx, int8_t sy, bool skip) {
        MV m; m.dx = dx; m.dy = dy; m.sx = clamp_sub(sx); m.sy = clamp_sub(sy); m.skip = skip; return m;
    }
    // motion_vector.rs:81-97
    uint8_t to_flags() const {
        uint8_t f = 0;
        if (sx == 1) f |= 0x01; else if (sx == -1) f |= 0x02;
        if (sy == 1) f |= 0x04; else if (sy == -1) f |= 0x08;
        if (skip) f |= 0x40;
        return f;
    }
    // motion_vector.rs:103-116
    MV canonicalized() const {
        MV m = *this;
        if (m.dx == -128 && m.sx == -1) m.sx = 0;
        if (m.dy == -128 && m.sy == -1) m.sy = 0;
        if (m.sx == -1) { m.sx = 1; m.dx = (int8_t)std::max(-128, (int)m.dx - 1); }
        if (m.sy == -1) { m.sy = 1; m.dy = (int8_t)std::max(-128, (int)m.dy - 1); }
        return m;
    }
};
struct Q4 { int8_t dx, dy, sx, sy; bool operator==(const Q4& o) const { return dx == o.dx && dy == o.dy && sx == o.sx && sy == o.sy; } };
static inline Q4 q4(const MV& m) { return Q4{m.dx, m.dy, m.sx, m.sy}; }

// ---------------------------------------------------------------------------
// YUV420 frame
// ---------------------------------------------------------------------------
struct Yuv {
    int w = 0, h = 0;
    std::vector<uint8_t> y, u, v;
    Yuv() {}
    Yuv(int w_, int h_) : w(w_), h(h_), y((size_t)w_ * h_), u((size_t)(w_ / 2) * (h_ / 2)), v((size_t)(w_ / 2) * (h_ / 2)) {}
};

// Project-owned BT.709 limited-range conversion, 14-bit fixed point, 2x2 box-averaged
// chroma. Stands in for yuv::rgb_to_yuv420 (reitero_video_common/src/yuv.rs:41-74);
// parity with the third-party crate is unpinned.
static void rgb_to_yuv420(const uint8_t* rgb, int w, int h, Yuv& out) {
    out = Yuv(w, h);
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            const uint8_t* p = rgb + ((size_t)y * w + x) * 3;
            int Y = (2991 * p[0] + 10064 * p[1] + 1016 * p[2] + (16 << 14) + (1 << 13)) >> 14;
            out.y[(size_t)y * w + x] = (uint8_t)clampv(Y, 0, 255);
        }
    int cw = w / 2;
    for (int cy = 0; cy < h / 2; cy++)
        for (int cx = 0; cx < cw; cx++) {
            int R = 0, G = 0, B = 0;
            for (int j = 0; j < 2; j++)
                for (int i = 0; i < 2; i++) {
                    const uint8_t* p = rgb + ((size_t)(2 * cy + j) * w + 2 * cx + i) * 3;
                    R += p[0]; G += p[1]; B += p[2];
                }
            int U = (-1649 * R - 5547 * G + 7196 * B + (128 << 16) + (1 << 15)) >> 16;
            int V = (7196 * R - 6536 * G - 660 * B + (128 << 16) + (1 << 15)) >> 16;
            out.u[(size_t)cy * cw + cx] = (uint8_t)clampv(U, 0, 255);
            out.v[(size_t)cy * cw + cx] = (uint8_t)clampv(V, 0, 255);
        }
}
// Stands in for yuv::yuv420_to_rgb (yuv.rs:94-110); nearest chroma, unpinned.
static void yuv420_to_rgb(const Yuv& f, uint8_t* rgb) {
    int cw = f.w / 2;
    for (int y = 0; y < f.h; y++)
        for (int x = 0; x < f.w; x++) {
            int Y = f.y[(size_t)y * f.w + x] - 16;
            int U = f.u[(size_t)(y >> 1) * cw + (x >> 1)] - 128;
            int V = f.v[(size_t)(y >> 1) * cw + (x >> 1)] - 128;
            int R = (19077 * Y + 29372 * V + (1 << 13)) >> 14;
            int G = (19077 * Y - 3494 * U - 8731 * V + (1 << 13)) >> 14;
            int B = (19077 * Y + 34610 * U + (1 << 13)) >> 14;
            uint8_t* p = rgb + ((size_t)y * f.w + x) * 3;
            p[0] = (uint8_t)clampv(R, 0, 255); p[1] = (uint8_t)clampv(G, 0, 255); p[2] = (uint8_t)clampv(B, 0, 255);
        }
}

// pad_rgb24  (reitero_encode/src/encoder.rs:1049-1096)
static std::vector<uint8_t> pad_rgb24(const uint8_t* d, int dw, int dh, int sw, int sh) {
    std::vector<uint8_t> out((size_t)sw * sh * 3);
    for (int y = 0; y < dh; y++) {
        memcpy(&out[(size_t)y * sw * 3], d + (size_t)y * dw * 3, (size_t)dw * 3);
        const uint8_t* last = d + ((size_t)y * dw + dw - 1) * 3;
        for (int x = dw; x < sw; x++) memcpy(&out[((size_t)y * sw + x) * 3], last, 3);
    }
    for (int y = dh; y < sh; y++) memcpy(&out[(size_t)y * sw * 3], &out[(size_t)(dh - 1) * sw * 3], (size_t)sw * 3);
    return out;
}
// crop_rgb24  (reitero_decode/src/decoder.rs:18-32)
static void crop_rgb24(const uint8_t* s, int sw, int dw, int dh, uint8_t* out) {
    for (int y = 0; y < dh; y++) memcpy(out + (size_t)y * dw * 3, s + (size_t)y * sw * 3, (size_t)dw * 3);
}

// ---------------------------------------------------------------------------
// BinProb + rANS  (reitero_video_common/src/rans.rs)
// ---------------------------------------------------------------------------
static uint8_t g_zero_prob[65536];
static bool g_zero_prob_init = false;
// rans.rs:30-44
static void init_zero_prob() {
    if (g_zero_prob_init) return;
    g_zero_prob[0] = 1;
    for (int i = 1; i < 65536; i++) {
        int z = i >> 8, o = i & 0xff;
        uint8_t p = (uint8_t)((z << 8) / (z + o));
        g_zero_prob[i] = p ? p : 1;
    }
    g_zero_prob_init = true;
}
struct BinProb {
    uint16_t packed = 0x0101;   // rans.rs:55-58
    uint8_t prob_zero() const { return g_zero_prob[packed]; }
    static uint16_t rotl8(uint16_t v) { return (uint16_t)((v << 8) | (v >> 8)); }
    // rans.rs:73-84
    void observe(bool bit) {
        uint16_t rot = bit ? rotl8(packed) : packed;
        uint32_t sum = (uint32_t)rot + 0x100u;
        uint16_t upd;
        if (sum > 0xffffu) {
            uint16_t mask = (rot == 0xff01) ? 0xff00 : 0x8100;
            upd = (uint16_t)((uint16_t)(rot + 0x101u) >> 1) | mask;
        } else {
            upd = (uint16_t)sum;
        }
        packed = bit ? rotl8(upd) : upd;
    }
};

static const uint32_t RANS_L = 1u << 16;   // rans.rs:107
static const int BATCH = 16386;            // rans.rs:201

// rans.rs:190-197
static inline void to_range(bool bit, uint32_t pz, uint32_t& start, uint32_t& freq) {
    if (bit) { start = pz; freq = 256 - pz; } else { start = 0; freq = pz; }
}

// RansEncoder  (rans.rs:209-298)
struct RansEnc {
    struct Pending { bool bit; uint8_t p; };
    std::vector<Pending> stack;
    int sp = BATCH;
    std::vector<uint8_t>* sink;
    explicit RansEnc(std::vector<uint8_t>* s) : stack(BATCH), sink(s) {}
    static void put_sym(uint32_t& r, std::deque<uint16_t>& out, uint32_t start, uint32_t freq) {
        uint32_t x_max = ((RANS_L >> 8) << 16) * freq;   // rans.rs:148-157
        if (r >= x_max) { out.push_front((uint16_t)r); r >>= 16; }
        r = ((r / freq) << 8) | ((r % freq) + start);   // rans.rs:163-166
    }
    void drain() {   // rans.rs:231-269
        uint32_t r0 = RANS_L, r1 = RANS_L;
        std::deque<uint16_t> out;
        int i = sp;
        if (i & 1) {
            const Pending& s = stack[i];
            uint32_t st, fr; to_range(s.bit, s.p ? s.p : 1, st, fr);
            put_sym(r0, out, st, fr);
            i++;
            std::swap(r0, r1);
        }
        while (i < BATCH) {
            const Pending& s0 = stack[i];
            const Pending& s1 = stack[i + 1];
            uint32_t st0, fr0, st1, fr1;
            to_range(s0.bit, s0.p ? s0.p : 1, st0, fr0);
            to_range(s1.bit, s1.p ? s1.p : 1, st1, fr1);
            put_sym(r0, out, st0, fr0);
            put_sym(r1, out, st1, fr1);
            i += 2;
        }
        out.push_front((uint16_t)(r0 >> 16)); out.push_front((uint16_t)(r0 & 0xffff));
        out.push_front((uint16_t)(r1 >> 16)); out.push_front((uint16_t)(r1 & 0xffff));
        for (uint16_t w : out) { sink->push_back((uint8_t)(w & 0xff)); sink->push_back((uint8_t)(w >> 8)); }
        sp = BATCH;
    }
    void put(bool bit, BinProb& ctx) {   // rans.rs:272-279
        uint8_t p = ctx.prob_zero();
        ctx.observe(bit);
        if (sp == 0) drain();
        sp--;
        stack[sp] = Pending{bit, p};
    }
    void put_bypass(bool bit) {   // rans.rs:282-287
        if (sp == 0) drain();
        sp--;
        stack[sp] = Pending{bit, 128};
    }
    void finish() { drain(); }
};

struct StreamErr : std::runtime_error { using std::runtime_error::runtime_error; };

// RansDecoder  (rans.rs:306-354)
struct RansDec {
    const uint8_t* data; size_t len; size_t pos = 0;
    uint32_t r0 = 0, r1 = 0; int count = 0;
    RansDec(const uint8_t* d, size_t n) : data(d), len(n) { r0 = init(); r1 = init(); }
    uint16_t rd16() {
        if (pos + 2 > len) throw StreamErr("rans: read past end");
        uint16_t v = (uint16_t)(data[pos] | (data[pos + 1] << 8)); pos += 2; return v;
    }
    uint32_t init() { uint32_t lo = rd16(); uint32_t hi = rd16(); return lo | (hi << 16); }
    void maybe_reload() {
        if (count == BATCH) { count = 0; r0 = init(); r1 = init(); }
        count++;
    }
    void advance(uint32_t& r, uint32_t start, uint32_t freq) {
        r = freq * (r >> 8) + (r & 255) - start;
        if (r < RANS_L) r = (r << 16) | rd16();
    }
    bool get(BinProb& ctx) {
        maybe_reload();
        uint32_t r = r0; r0 = r1;
        uint32_t p = ctx.prob_zero();
        bool bit = (r & 255) >= p;
        ctx.observe(bit);
        uint32_t st, fr; to_range(bit, p, st, fr);
        advance(r, st, fr);
        r1 = r;
        return bit;
    }
    bool get_bypass() {
        maybe_reload();
        uint32_t r = r0; r0 = r1;
        uint32_t start = r & 0x80;
        advance(r, start, 128);
        r1 = r;
        return start != 0;
    }
};

// ---------------------------------------------------------------------------
// Residual coefficient coder  (reitero_residual/src/rans.rs:1-460)
// ---------------------------------------------------------------------------
static const int EOB_Y_CLASSES = 9, EOB_UV_CLASSES = 7, MAG_CLASSES = 13;
static inline int eob_class(int v) { if (v <= 1) return v; int b = 0; while ((1 << b) <= v) b++; return b; }   // rans.rs:43-50
static inline int luma_mag_band(int p) { return p == 0 ? 0 : (p < 16 ? 1 : (p < 64 ? 2 : 3)); }   // :52-58
static inline int chroma_mag_band(int p) { return p == 0 ? 0 : (p < 8 ? 1 : 2); }   // :60-65

// encode_magnitude / encode_eob share the same class scheme  (rans.rs:69-96, :128-155)
static void enc_class_value(RansEnc& w, int value, BinProb* ctx, int nctx) {
    int max_class = nctx - 1;
    int cls = eob_class(value);
    if (cls > max_class) throw StreamErr("class overflow");   // the reference indexes out of bounds (panic)
    for (int c = 0; c < cls; c++) w.put(true, ctx[c]);
    if (cls < max_class) w.put(false, ctx[cls]);
    if (cls >= 2) {
        int base = 1 << (cls - 1), off = value - base;
        for (int b = cls - 2; b >= 0; b--) w.put_bypass(((off >> b) & 1) != 0);
    }
}
static int dec_class_value(RansDec& r, BinProb* ctx, int nctx) {   // rans.rs:99-125, :158-183
    int max_class = nctx - 1, cls = 0;
    while (cls < max_class) { if (!r.get(ctx[cls])) break; cls++; }
    if (cls <= 1) return cls;
    int base = 1 << (cls - 1), off = 0;
    for (int i = 0; i < cls - 1; i++) off = (off << 1) | (r.get_bypass() ? 1 : 0);
    return base + off;
}

struct DctCtx {
    BinProb luma[256], chroma[64];
    BinProb luma_mag[4][MAG_CLASSES], chroma_mag[3][MAG_CLASSES];
    BinProb eob_y[EOB_Y_CLASSES], eob_uv[EOB_UV_CLASSES];
    int16_t prev_dc[3] = {0, 0, 0};
    bool dc_pred = false;
};

// DctRansEncoder  (rans.rs:189-327)
struct DctRansEnc {
    std::vector<uint8_t> out;
    RansEnc w;
    DctCtx c;
    explicit DctRansEnc(bool dc_prediction) : w(&out) { c.dc_pred = dc_prediction; }
    void encode_plane(const int16_t* plane, int size, bool luma, int pidx) {   // rans.rs:248-321
        int16_t dc_val; int last_nz;
        if (c.dc_pred) {
            int16_t dc_orig = plane[0];
            int16_t dc_delta = (int16_t)(dc_orig - c.prev_dc[pidx]);
            c.prev_dc[pidx] = dc_orig;
            last_nz = 0;
            for (int i = size - 1; i >= 1; i--) if (plane[i] != 0) { last_nz = i; break; }
            dc_val = dc_delta;
        } else {
            last_nz = 0;
            for (int i = size - 1; i >= 0; i--) if (plane[i] != 0) { last_nz = i; break; }
            dc_val = plane[0];
        }
        if (luma) enc_class_value(w, last_nz, c.eob_y, EOB_Y_CLASSES);
        else enc_class_value(w, last_nz, c.eob_uv, EOB_UV_CLASSES);
        BinProb* sig = luma ? c.luma : c.chroma;
        for (int i = 0; i <= last_nz; i++) {
            int16_t val = i == 0 ? dc_val : plane[i];
            w.put(val != 0, sig[i]);
            if (val != 0) {
                w.put_bypass(val < 0);
                int abs_val = std::abs((int)val) - 1;
                BinProb* mag = luma ? c.luma_mag[luma_mag_band(i)] : c.chroma_mag[chroma_mag_band(i)];
                enc_class_value(w, abs_val, mag, MAG_CLASSES);
            }
        }
    }
    void encode_block(const int16_t* y, const int16_t* u, const int16_t* v) {
        encode_plane(y, 256, true, 0); encode_plane(u, 64, false, 1); encode_plane(v, 64, false, 2);
    }
    std::vector<uint8_t> finish() { w.finish(); return out; }
};

// DctRansDecoder  (rans.rs:333-460)
struct DctRansDec {
    RansDec* r = nullptr;
    DctCtx c;
    explicit DctRansDec(bool dc_prediction) { c.dc_pred = dc_prediction; }
    ~DctRansDec() { delete r; }
    void consume(const uint8_t* d, size_t n) { delete r; r = new RansDec(d, n); }
    void decode_plane(int16_t* blk, int size, bool luma, int pidx) {
        if (!r) throw StreamErr("residual decoder used before consume");
        int16_t dc_pred = c.dc_pred ? c.prev_dc[pidx] : 0;
        int end = luma ? dec_class_value(*r, c.eob_y, EOB_Y_CLASSES) : dec_class_value(*r, c.eob_uv, EOB_UV_CLASSES);
        if (end >= size) throw StreamErr("EOB exceeds block size");   // rans.rs:417-422 (assert)
        BinProb* sig = luma ? c.luma : c.chroma;
        for (int i = 0; i <= end; i++) {
            if (r->get(sig[i])) {
                bool neg = r->get_bypass();
                BinProb* mag = luma ? c.luma_mag[luma_mag_band(i)] : c.chroma_mag[chroma_mag_band(i)];
                int16_t a = (int16_t)(dec_class_value(*r, mag, MAG_CLASSES) + 1);
                blk[i] = neg ? (int16_t)-a : a;
            }
        }
        if (c.dc_pred) { blk[0] = (int16_t)(blk[0] + dc_pred); c.prev_dc[pidx] = blk[0]; }
    }
    void decode_block(int16_t* y, int16_t* u, int16_t* v) {
        decode_plane(y, 256, true, 0); decode_plane(u, 64, false, 1); decode_plane(v, 64, false, 2);
    }
};

// ---------------------------------------------------------------------------
// MV modes + predictors  (reitero_residual/src/mv_predictor.rs)
// ---------------------------------------------------------------------------
enum Mode : uint8_t { ZERO = 0, NEAREST = 1, NEAR = 2, TOPRIGHT = 3, TOPLEFT = 4, NEWMV = 5, TEMPORAL = 6 };   // :5-33

struct Neigh { bool hl = false, ht = false, htr = false, htl = false, htemp = false; Q4 l{}, t{}, tr{}, tl{}, temp{}; };

// gather_mv_neighbor_set  (mv_predictor.rs:59-118)
static Neigh gather_neighbors(const std::vector<MV>& cur, const std::vector<MV>* prev, int bw, int bh, int bx, int by) {
    Neigh n; size_t idx = (size_t)by * bw + bx, mx = cur.size();
    if (bx > 0 && idx - 1 < mx) { n.hl = true; n.l = q4(cur[idx - 1]); }
    if (by > 0) {
        size_t t = idx - bw;
        if (t < mx) { n.ht = true; n.t = q4(cur[t]); }
        if (bx + 1 < bw && t + 1 < mx) { n.htr = true; n.tr = q4(cur[t + 1]); }
        if (bx > 0 && t - 1 < mx) { n.htl = true; n.tl = q4(cur[t - 1]); }
    }
    if (prev && prev->size() == (size_t)bw * bh && idx < prev->size()) { n.htemp = true; n.temp = q4((*prev)[idx]); }
    return n;
}

struct Preds { Q4 nearest{}, near{}, temporal{}; int unique = 0; };
// derive_mv_predictors_with_stats + gather_mv_candidates + calculate_temporal_predictor
// (mv_predictor.rs:135-272)
static Preds derive_predictors(const std::vector<MV>& cur, const std::vector<MV>* prev, int bw, int bh, int bx, int by) {
    std::vector<Q4> cand;
    auto add = [&](Q4 q) { for (auto& e : cand) if (e == q) return; cand.push_back(q); };
    size_t idx = (size_t)by * bw + bx, mx = cur.size();
    if (bx > 0 && idx - 1 < mx) add(q4(cur[idx - 1]));
    if (by > 0) {
        size_t t = idx - bw;
        if (t < mx) add(q4(cur[t]));
        if (bx + 1 < bw && t + 1 < mx) add(q4(cur[t + 1]));
        if (bx > 0 && t > 0 && t - 1 < mx) add(q4(cur[t - 1]));
    }
    if (prev && prev->size() == (size_t)bw * bh && idx < prev->size()) add(q4((*prev)[idx]));
    if (cand.empty()) cand.push_back(Q4{0, 0, 0, 0});
    Preds p; p.unique = (int)cand.size();
    if (cand.size() == 1) cand.push_back(Q4{0, 0, 0, 0});
    p.nearest = cand[0]; p.near = cand[1];
    p.temporal = Q4{0, 0, 0, 0};
    if (prev && idx < prev->size()) p.temporal = q4((*prev)[idx]);
    return p;
}

// mv_mode_context  (mv_predictor.rs:275-308)
static int mode_context(const std::vector<uint8_t>& modes, int bw, int bx, int by) {
    size_t idx = (size_t)by * bw + bx;
    int score = 0;
    auto sc = [](uint8_t m) { return m == ZERO ? 2 : (m == NEWMV ? 0 : 1); };
    if (bx > 0 && idx - 1 < modes.size()) score += sc(modes[idx - 1]);
    if (by > 0 && idx >= (size_t)bw && idx - bw < modes.size()) score += sc(modes[idx - bw]);
    return std::min(score, 3);
}

// ---------------------------------------------------------------------------
// MV entropy coder  (reitero_residual/src/mv_rans.rs, mv_rans/encoder.rs, mv_rans/decoder.rs)
// ---------------------------------------------------------------------------
static const int MODE_CTXS = 4, MV_MAX_CLASS = 10, MV_CLASS_TREE_DEPTH = 4, MV_MAX_MAG_BITS = 10, CANDIDATE_SIZE = 16, CANDIDATE_IDX_DEPTH = 4;
// Subpel ordinal order Zero < PlusHalf < MinusHalf (mv_rans.rs:25-29, derived Ord)
static inline int subpel_ord(int8_t s) { return s == 0 ? 0 : (s == 1 ? 1 : 2); }
static inline int8_t subpel_from_ord(int o) { return o == 0 ? 0 : (o == 1 ? 1 : -1); }

struct CodedBlock { uint8_t mode = ZERO, new_base = 0; int8_t delta_x = 0, delta_y = 0, sub_x = 0, sub_y = 0; bool skip = false; };   // mv_rans.rs:61-79

// mv_rans.rs:81-95
static int mv_class_from_magnitude(int mag) {
    if (mag == 0) return 0;
    int msb = 15; while (!((mag >> msb) & 1)) msb--;
    return std::min(msb + 1, MV_MAX_CLASS);
}
static int mv_class_base(int c) { return c == 0 ? 0 : (1 << (c - 1)); }
static int mv_class_bits(int c) { return c <= 1 ? 0 : c - 1; }

struct MvCtx {
    BinProb skip;
    BinProb is_new[MODE_CTXS], nonzero[MODE_CTXS], is_near[MODE_CTXS], is_temporal[MODE_CTXS], is_tr[MODE_CTXS], is_tl[MODE_CTXS];
    BinProb nb_nearest[MODE_CTXS], nb_near[MODE_CTXS], nb_tr[MODE_CTXS], nb_tl[MODE_CTXS];
    BinProb class_x[MV_CLASS_TREE_DEPTH], class_y[MV_CLASS_TREE_DEPTH];
    BinProb sign_x, sign_y;
    BinProb mag_x[MV_MAX_MAG_BITS], mag_y[MV_MAX_MAG_BITS];
    BinProb sub_x_has, sub_y_has;
    BinProb cand_hit[3];
    BinProb cand_idx[CANDIDATE_IDX_DEPTH];
    typedef std::tuple<int, int, int, int> Key;   // (dx, dy, subpel ordinal x, subpel ordinal y)
    std::vector<Key> candidates;
    std::map<Key, uint32_t> counts;
    uint32_t total_new_history = 0;
    void reset() { *this = MvCtx(); }
    static Key key(int8_t dx, int8_t dy, int8_t sx, int8_t sy) { return Key(dx, dy, subpel_ord(sx), subpel_ord(sy)); }
    int find(const Key& k) const { for (size_t i = 0; i < candidates.size(); i++) if (candidates[i] == k) return (int)i; return -1; }
    // mv_rans/encoder.rs:370-412 and mv_rans/decoder.rs:283-326
    void update_candidates(const std::vector<CodedBlock>& blocks, bool decoder_side) {
        for (auto& kv : counts) kv.second = (kv.second * 7) / 8;
        total_new_history = (total_new_history * 7) / 8;
        for (auto& b : blocks) {
            if (b.mode != NEWMV) continue;
            if (b.sub_x == -1 || b.sub_y == -1) {
                if (decoder_side) continue;
                throw StreamErr("internal invariant violated: -0.5 subpixel should never occur");
            }
            counts[key(b.delta_x, b.delta_y, b.sub_x, b.sub_y)] += 1;
            total_new_history += 1;
        }
        for (auto it = counts.begin(); it != counts.end();) { if (it->second == 0) it = counts.erase(it); else ++it; }
        if (total_new_history < 100) { candidates.clear(); return; }
        std::vector<std::pair<Key, uint32_t>> sorted(counts.begin(), counts.end());
        std::stable_sort(sorted.begin(), sorted.end(), [](const std::pair<Key, uint32_t>& a, const std::pair<Key, uint32_t>& b) {
            if (a.second != b.second) return a.second > b.second;
            return a.first < b.first;
        });
        uint32_t thr = total_new_history / 100;
        candidates.clear();
        for (auto& kv : sorted) { if (kv.second > thr) { candidates.push_back(kv.first); if ((int)candidates.size() == CANDIDATE_SIZE) break; } }
    }
};

// mv_rans/encoder.rs:12-55
static void enc_class_symbol(RansEnc& w, BinProb* ctx, int cls) {
    int lo = 0, hi = MV_MAX_CLASS, d = 0;
    while (lo < hi) { int mid = (lo + hi) / 2; bool g = cls > mid; w.put(g, ctx[d]); if (g) lo = mid + 1; else hi = mid; d++; }
}
static void enc_candidate_index(RansEnc& w, BinProb* ctx, int idx) {
    int lo = 0, hi = CANDIDATE_SIZE, d = 0;
    while (hi - lo > 1) { int mid = (lo + hi) / 2; bool g = idx >= mid; w.put(g, ctx[d]); if (g) lo = mid; else hi = mid; d++; }
}
// mv_rans/decoder.rs:12-54
static int dec_class_symbol(RansDec& r, BinProb* ctx) {
    int lo = 0, hi = MV_MAX_CLASS, d = 0;
    while (lo < hi) { int mid = (lo + hi) / 2; bool g = r.get(ctx[d]); if (g) lo = mid + 1; else hi = mid; d++; }
    return lo;
}
static int dec_candidate_index(RansDec& r, BinProb* ctx) {
    int lo = 0, hi = CANDIDATE_SIZE, d = 0;
    while (hi - lo > 1) { int mid = (lo + hi) / 2; bool g = r.get(ctx[d]); if (g) lo = mid; else hi = mid; d++; }
    return lo;
}
// encode_component  (mv_rans/encoder.rs:303-331)
static void enc_component(RansEnc& w, int8_t delta, BinProb* cls_ctx, BinProb& sign_ctx, BinProb* mag_ctx) {
    int mag = std::abs((int)delta);
    int cls = mv_class_from_magnitude(mag);
    enc_class_symbol(w, cls_ctx, cls);
    if (cls == 0) return;
    w.put(delta < 0, sign_ctx);
    int off = mag - mv_class_base(cls);
    for (int b = 0; b < mv_class_bits(cls); b++) w.put(((off >> b) & 1) != 0, mag_ctx[b]);
}
// decode_component  (mv_rans/decoder.rs:375-401)
static int8_t dec_component(RansDec& r, BinProb* cls_ctx, BinProb& sign_ctx, BinProb* mag_ctx) {
    int cls = dec_class_symbol(r, cls_ctx);
    if (cls == 0) return 0;
    bool sign = r.get(sign_ctx);
    int off = 0;
    for (int b = 0; b < mv_class_bits(cls); b++) if (r.get(mag_ctx[b])) off |= 1 << b;
    int16_t v = (int16_t)(mv_class_base(cls) + off);
    int16_t s = sign ? (int16_t)-v : v;
    return (int8_t)clampv<int16_t>(s, -128, 127);
}

// MvRansEncoder::encode_frame_and_get_data  (mv_rans/encoder.rs:145-255, :233-301, :351-368)
static std::vector<uint8_t> mv_encode_frame(MvCtx& c, const std::vector<CodedBlock>& blocks, int bw, int bh) {
    std::vector<uint8_t> out;
    RansEnc w(&out);
    std::vector<uint8_t> modes; std::vector<bool> hits;
    for (size_t i = 0; i < blocks.size(); i++) {
        int bx = (int)(i % bw), by = (int)(i / bw);
        const CodedBlock& b = blocks[i];
        w.put(b.skip, c.skip);
        int ctx = mode_context(modes, bw, bx, by);
        // encode_mode  (mv_rans/encoder.rs:257-301)
        bool is_new = b.mode == NEWMV;
        w.put(is_new, c.is_new[ctx]);
        if (!is_new) {
            bool nz = b.mode != ZERO;
            w.put(nz, c.nonzero[ctx]);
            if (nz) {
                bool nr = b.mode == NEAR; w.put(nr, c.is_near[ctx]);
                if (!nr) {
                    bool tp = b.mode == TEMPORAL; w.put(tp, c.is_temporal[ctx]);
                    if (!tp) {
                        bool trm = b.mode == TOPRIGHT; w.put(trm, c.is_tr[ctx]);
                        if (!trm) { bool tlm = b.mode == TOPLEFT; w.put(tlm, c.is_tl[ctx]); }
                    }
                }
            }
        }
        if (is_new) {
            bool n0 = b.new_base == 0; w.put(n0, c.nb_nearest[ctx]);
            if (!n0) {
                bool n1 = b.new_base == 1; w.put(n1, c.nb_near[ctx]);
                if (!n1) {
                    bool n2 = b.new_base == 2; w.put(n2, c.nb_tr[ctx]);
                    if (!n2) { bool n3 = b.new_base == 3; w.put(n3, c.nb_tl[ctx]); }
                }
            }
            bool frac_via_cand = false;
            if (!c.candidates.empty()) {
                int ci = c.find(MvCtx::key(b.delta_x, b.delta_y, b.sub_x, b.sub_y));
                bool hit = ci >= 0;
                bool lh = bx > 0 ? hits[i - 1] : false;
                bool th = by > 0 ? hits[i - bw] : false;
                w.put(hit, c.cand_hit[(int)lh + (int)th]);
                if (hit) { enc_candidate_index(w, c.cand_idx, ci); frac_via_cand = true; }
                else {
                    enc_component(w, b.delta_x, c.class_x, c.sign_x, c.mag_x);
                    enc_component(w, b.delta_y, c.class_y, c.sign_y, c.mag_y);
                }
                hits.push_back(hit);
            } else {
                enc_component(w, b.delta_x, c.class_x, c.sign_x, c.mag_x);
                enc_component(w, b.delta_y, c.class_y, c.sign_y, c.mag_y);
                hits.push_back(false);
            }
            if (!frac_via_cand) {   // encode_fractional  (mv_rans/encoder.rs:333-349)
                if (b.sub_x == -1 || b.sub_y == -1) throw StreamErr("raw -0.5 subpixel is disallowed");
                w.put(b.sub_x != 0, c.sub_x_has);
                w.put(b.sub_y != 0, c.sub_y_has);
            }
        } else {
            hits.push_back(false);
        }
        modes.push_back(b.mode);
    }
    w.finish();
    c.update_candidates(blocks, false);
    return out;
}

// MvRansDecoder::decode_frame  (mv_rans/decoder.rs:149-283, :328-373, :403-425)
static std::vector<CodedBlock> mv_decode_frame(MvCtx& c, const uint8_t* d, size_t n, int bw, int bh) {
    RansDec r(d, n);
    int nb = bw * bh;
    std::vector<CodedBlock> res; res.reserve(nb);
    std::vector<uint8_t> modes; std::vector<bool> hits;
    for (int i = 0; i < nb; i++) {
        int bx = i % bw, by = i / bw;
        CodedBlock b;
        b.skip = r.get(c.skip);
        int ctx = mode_context(modes, bw, bx, by);
        uint8_t mode;
        if (r.get(c.is_new[ctx])) mode = NEWMV;
        else if (!r.get(c.nonzero[ctx])) mode = ZERO;
        else if (r.get(c.is_near[ctx])) mode = NEAR;
        else if (r.get(c.is_temporal[ctx])) mode = TEMPORAL;
        else if (r.get(c.is_tr[ctx])) mode = TOPRIGHT;
        else if (r.get(c.is_tl[ctx])) mode = TOPLEFT;
        else mode = NEAREST;
        b.mode = mode;
        if (mode == NEWMV) {
            if (r.get(c.nb_nearest[ctx])) b.new_base = 0;
            else if (r.get(c.nb_near[ctx])) b.new_base = 1;
            else if (r.get(c.nb_tr[ctx])) b.new_base = 2;
            else b.new_base = r.get(c.nb_tl[ctx]) ? 3 : 4;
            bool frac_via_cand = false;
            if (!c.candidates.empty()) {
                bool lh = bx > 0 ? hits[i - 1] : false;
                bool th = by > 0 ? hits[i - bw] : false;
                bool hit = r.get(c.cand_hit[(int)lh + (int)th]);
                if (hit) {
                    int idx = dec_candidate_index(r, c.cand_idx);
                    if (idx < (int)c.candidates.size()) {
                        auto& k = c.candidates[idx];
                        b.delta_x = (int8_t)std::get<0>(k); b.delta_y = (int8_t)std::get<1>(k);
                        b.sub_x = subpel_from_ord(std::get<2>(k)); b.sub_y = subpel_from_ord(std::get<3>(k));
                        frac_via_cand = true;
                    }
                } else {
                    b.delta_x = dec_component(r, c.class_x, c.sign_x, c.mag_x);
                    b.delta_y = dec_component(r, c.class_y, c.sign_y, c.mag_y);
                }
                hits.push_back(hit);
            } else {
                b.delta_x = dec_component(r, c.class_x, c.sign_x, c.mag_x);
                b.delta_y = dec_component(r, c.class_y, c.sign_y, c.mag_y);
                hits.push_back(false);
            }
            if (!frac_via_cand) {
                b.sub_x = r.get(c.sub_x_has) ? 1 : 0;
                b.sub_y = r.get(c.sub_y_has) ? 1 : 0;
            }
        } else {
            hits.push_back(false);
        }
        res.push_back(b);
        modes.push_back(mode);
    }
    c.update_candidates(res, true);
    return res;
}

// ---------------------------------------------------------------------------
// Motion estimation  (reitero_encode/src/motion/scalar.rs, satd_scalar.rs)
// Luma is kept as u8: the reference's i32<<8 LumaPlane (luma.rs:31-38) only scales
// the samples, and its half-pel interpolation on the scaled values reduces to the
// floor averages used here (satd_scalar.rs:160-165).
// ---------------------------------------------------------------------------
static const int HEX[6][2] = {{2, 0}, {1, 2}, {-1, 2}, {-2, 0}, {-1, -2}, {1, -2}};   // scalar.rs:8
static const int CARD[4][2] = {{1, 0}, {-1, 0}, {0, 1}, {0, -1}};                     // scalar.rs:7

static inline int median3(int a, int b, int c) {   // scalar.rs:24-43
    if (a > b) { if (b > c) return b; else if (a > c) return c; else return a; }
    else { if (a > c) return a; else if (b > c) return c; else return b; }
}

// sad_block_int_limit (full SAD; the early exit is result-invariant)  satd_scalar.rs:4-99
static int64_t sad_int(const uint8_t* prev, const uint8_t* cur, int W, int H, int x0, int y0, int dx, int dy) {
    int64_t sad = 0;
    if (x0 >= 0 && y0 >= 0 && x0 + 16 <= W && y0 + 16 <= H && x0 + dx >= 0 && y0 + dy >= 0 && x0 + dx + 16 <= W && y0 + dy + 16 <= H) {
        // both blocks inside: no clamping applies (satd_simd.rs:15-111 in-bounds path)
        int acc = 0;
        for (int j = 0; j < 16; j++) {
            const uint8_t* c = cur + (size_t)(y0 + j) * W + x0;
            const uint8_t* r = prev + (size_t)(y0 + dy + j) * W + x0 + dx;
            for (int i = 0; i < 16; i++) acc += std::abs((int)c[i] - (int)r[i]);
        }
        return acc;
    }
    for (int by = 0; by < 4; by++)
        for (int bx = 0; bx < 4; bx++) {
            int sx0 = x0 + bx * 4, sy0 = y0 + by * 4;
            if (sx0 < 0 || sy0 < 0 || sx0 >= W || sy0 >= H) continue;
            int rx = sx0 + dx, ry = sy0 + dy;
            if (rx < 0) rx = 0;
            if (ry < 0) ry = 0;
            if (rx + 4 > W) rx = std::max(W - 4, 0);
            if (ry + 4 > H) ry = std::max(H - 4, 0);
            for (int j = 0; j < 4; j++)
                for (int i = 0; i < 4; i++) {
                    int cy = sy0 + j, cx = sx0 + i, yy = ry + j, xx = rx + i;
                    if (cy >= 0 && cy < H && cx >= 0 && cx < W && yy >= 0 && yy < H && xx >= 0 && xx < W)
                        sad += std::abs((int)cur[(size_t)cy * W + cx] - (int)prev[(size_t)yy * W + xx]);
                }
        }
    return sad;
}
// sample_luma_halfpel  (satd_scalar.rs:113-166) -- floor interpolation
static inline int sample_luma_hp_floor(const uint8_t* p, int W, int H, int xh, int yh) {
    xh = clampv(xh, 0, (W - 1) * 2); yh = clampv(yh, 0, (H - 1) * 2);
    int x0 = xh / 2, y0 = yh / 2;
    bool xo = xh & 1, yo = yh & 1;
    int x1 = x0 + 1 < W ? x0 + 1 : x0, y1 = y0 + 1 < H ? y0 + 1 : y0;
    int a = p[(size_t)y0 * W + x0];
    if (!xo && !yo) return a;
    int b = p[(size_t)y0 * W + x1], d = p[(size_t)y1 * W + x0], e = p[(size_t)y1 * W + x1];
    if (xo && !yo) return (a + b) >> 1;
    if (!xo && yo) return (a + d) >> 1;
    return (a + b + d + e) >> 2;
}
// sad_block_halfpel_limit_luma (full SAD)  satd_scalar.rs:168-223
static int64_t sad_hp(const uint8_t* prev, const uint8_t* cur, int W, int H, int x0, int y0, int dxh, int dyh) {
    int64_t sad = 0;
    if (2 * x0 + dxh >= 0 && 2 * (x0 + 15) + dxh + 1 <= 2 * (W - 1) && 2 * y0 + dyh >= 0 && 2 * (y0 + 15) + dyh + 1 <= 2 * (H - 1)) {
        // every sample and its interpolation partner inside: no clamping (satd_simd.rs:184-246 path)
        const int bx = (2 * x0 + dxh) >> 1, by = (2 * y0 + dyh) >> 1, ax = dxh & 1, ay = dyh & 1;
        int acc = 0;
        for (int j = 0; j < 16; j++) {
            const uint8_t* c = cur + (size_t)(y0 + j) * W + x0;
            const uint8_t* r0 = prev + (size_t)(by + j) * W + bx;
            const uint8_t* r1 = r0 + (ay ? W : 0);
            for (int i = 0; i < 16; i++) {
                int v = (r0[i] + r0[i + ax] + r1[i] + r1[i + ax]) >> 2;   // floor average of 1, 2 or 4 samples
                acc += std::abs((int)c[i] - v);
            }
        }
        return acc;
    }
    for (int yy = 0; yy < 16; yy++)
        for (int xx = 0; xx < 16; xx++) {
            int cx = x0 + xx, cy = y0 + yy;
            int r = sample_luma_hp_floor(prev, W, H, cx * 2 + dxh, cy * 2 + dyh);
            sad += std::abs((int)cur[(size_t)cy * W + cx] - r);
        }
    return sad;
}

// estimate_mv_cost  (scalar.rs:453-544)
static int64_t estimate_mv_cost(int dx, int dy, int nx, int ny, int rx, int ry, int mx, int my, double lambda) {
    if (lambda == 0.0) return 0;
    const double scale = 2.0;
    int64_t best = INT64_MAX;
    if (dx == nx && dy == ny) best = std::min(best, f64_to_i64_sat(lambda * scale * 1.5));
    if (dx == rx && dy == ry) best = std::min(best, f64_to_i64_sat(lambda * scale * 1.5));
    if (dx == mx && dy == my) best = std::min(best, f64_to_i64_sat(lambda * scale * 1.5));
    if (dx == 0 && dy == 0) best = std::min(best, f64_to_i64_sat(lambda * scale * 0.5));
    for (int k = 0; k < 2; k++) {
        int bxr = k == 0 ? nx : rx, byr = k == 0 ? ny : ry;
        int ddx = std::abs(dx - bxr), ddy = std::abs(dy - byr);
        double bx = ddx == 0 ? 1.0 : 2.0 * std::log2((double)ddx) + 2.0;
        double by = ddy == 0 ? 1.0 : 2.0 * std::log2((double)ddy) + 2.0;
        best = std::min(best, f64_to_i64_sat(lambda * scale * (6.0 + bx + by)));
    }
    return best;
}
static inline int64_t sat_sub(int64_t a, int64_t b) {
    if (b < 0 && a > INT64_MAX + b) return INT64_MAX;
    if (b > 0 && a < INT64_MIN + b) return INT64_MIN;
    return a - b;
}

struct MeResult { MV mv; int64_t sad; };

// process_block_hex + evaluate_seed_candidates + refine_integer_motion + refine_halfpel
// (scalar.rs:546-839)
static MeResult process_block_hex(int bx, int by, int W, int H, int sr, const std::vector<std::pair<int, int>>& cands,
                                  const uint8_t* prev, const uint8_t* cur, int64_t zero_thr, int64_t pred_thr, double lambda,
                                  int nx, int ny, int rx, int ry, int mx, int my) {
    // block_origin (scalar.rs:158-166): storage dims are multiples of 16
    int x0 = bx * 16 + 16 <= W ? bx * 16 : std::max(0, W - 16);
    int y0 = by * 16 + 16 <= H ? by * 16 : std::max(0, H - 16);
    int64_t zero_sad = sad_int(prev, cur, W, H, x0, y0, 0, 0);   // block_is_identical => 0 == SAD
    auto cost = [&](int dx, int dy) { return estimate_mv_cost(dx, dy, nx, ny, rx, ry, mx, my, lambda); };
    int sdx = 0, sdy = 0;
    int64_t scost = zero_sad + cost(0, 0);
    bool skip_int = zero_thr > 0 && zero_sad <= zero_thr;
    if (!skip_int) {
        for (auto& c : cands) {   // evaluate_seed_candidates (:638-678)
            int dx = c.first, dy = c.second;
            if ((dx == 0 && dy == 0) || std::abs(dx) > sr || std::abs(dy) > sr) continue;
            int64_t mc = cost(dx, dy);
            if (sat_sub(scost, mc) < 0) continue;
            int64_t t = sad_int(prev, cur, W, H, x0, y0, dx, dy) + mc;
            if (t < scost) { sdx = dx; sdy = dy; scost = t; }
        }
        int64_t cur_sad = scost - cost(sdx, sdy);
        if (pred_thr > 0 && cur_sad <= pred_thr) skip_int = true;
    }
    if (!skip_int) {   // refine_integer_motion (:680-757)
        for (;;) {
            bool improved = false;
            int cx = sdx, cy = sdy;
            for (auto& o : HEX) {
                int dx = cx + o[0], dy = cy + o[1];
                if (std::abs(dx) > sr || std::abs(dy) > sr) continue;
                int64_t mc = cost(dx, dy);
                if (sat_sub(scost, mc) < 0) continue;
                int64_t t = sad_int(prev, cur, W, H, x0, y0, dx, dy) + mc;
                if (t < scost) { sdx = dx; sdy = dy; scost = t; improved = true; }
            }
            if (!improved) break;
        }
        int cx = sdx, cy = sdy;
        for (auto& o : CARD) {
            int dx = cx + o[0], dy = cy + o[1];
            if (std::abs(dx) > sr || std::abs(dy) > sr) continue;
            int64_t mc = cost(dx, dy);
            if (sat_sub(scost, mc) < 0) continue;
            int64_t t = sad_int(prev, cur, W, H, x0, y0, dx, dy) + mc;
            if (t < scost) { sdx = dx; sdy = dy; scost = t; }
        }
    }
    // refine_halfpel (:759-839)
    int bdx = sdx * 2, bdy = sdy * 2, best_dx = bdx, best_dy = bdy;
    int64_t best_total = sad_hp(prev, cur, W, H, x0, y0, bdx, bdy), best_sad = best_total;
    for (int sy = -1; sy <= 1; sy++)
        for (int sx = -1; sx <= 1; sx++) {
            if (sx == 0 && sy == 0) continue;
            int dxh = bdx + sx, dyh = bdy + sy;
            if ((sdx == -128 && sx == -1) || (sdy == -128 && sy == -1)) continue;
            bool frac = (dxh % 2 != 0) || (dyh % 2 != 0);
            int64_t pen = frac ? f64_to_i64_sat(lambda * 2.0 * 1.5) : 0;
            if (sat_sub(best_total, pen) < 0) continue;
            int64_t s = sad_hp(prev, cur, W, H, x0, y0, dxh, dyh);
            int64_t t = s + pen;
            if (t < best_total) { best_total = t; best_sad = s; best_dx = dxh; best_dy = dyh; }
        }
    int dxh = clampv(best_dx, -255, 255), dyh = clampv(best_dy, -255, 255);
    MV m = MV::make((int8_t)clampv(dxh / 2, -128, 127), (int8_t)clampv(dyh / 2, -128, 127), (int8_t)(dxh % 2), (int8_t)(dyh % 2), false);
    return MeResult{m.canonicalized(), best_sad};
}

static void push_candidate(std::vector<std::pair<int, int>>& c, int dx, int dy, int sr) {   // scalar.rs:46-53
    if (std::abs(dx) > sr || std::abs(dy) > sr) return;
    for (auto& e : c) if (e.first == dx && e.second == dy) return;
    c.push_back({dx, dy});
}

// One ME step for block (bx, by) given a predictor-source accessor. `row_has_top` tells
// whether the row above is visible (false at a strip start in the `threads` emulation).
template <typename Get>
static MeResult me_block(Get get, bool row_has_top, const std::vector<MV>* prev_mvs, bool temporal_len_check, size_t nmv,
                         int bw, int bx, int by, int W, int H, int sr, const uint8_t* prev, const uint8_t* cur,
                         int64_t zthr, int64_t pthr, double lambda) {
    // gather_local_motion_candidates (scalar.rs:56-156) and the strip variant (:268-371)
    std::vector<std::pair<int, int>> preds;
    size_t idx = (size_t)by * bw + bx;
    if (bx > 0) { MV m = get(bx - 1, 0); preds.push_back({m.dx, m.dy}); }
    if (row_has_top) {
        MV t = get(bx, 1); preds.push_back({t.dx, t.dy});
        if (bx + 1 < bw) { MV m = get(bx + 1, 1); preds.push_back({m.dx, m.dy}); }
        if (bx > 0) { MV m = get(bx - 1, 1); preds.push_back({m.dx, m.dy}); }
    }
    if (prev_mvs && (!temporal_len_check || prev_mvs->size() == nmv) && idx < prev_mvs->size()) {
        const MV& m = (*prev_mvs)[idx]; preds.push_back({m.dx, m.dy});
    }
    std::vector<std::pair<int, int>> uniq;
    for (auto& p : preds) { bool f = false; for (auto& u : uniq) if (u == p) f = true; if (!f) uniq.push_back(p); }
    if (uniq.empty()) uniq.push_back({0, 0});
    if (uniq.size() == 1) uniq.push_back({0, 0});
    int lx = 0, ly = 0, tx = 0, ty = 0, trx = 0, try_ = 0;
    if (bx > 0) { MV m = get(bx - 1, 0); lx = m.dx; ly = m.dy; }
    if (row_has_top) { MV m = get(bx, 1); tx = m.dx; ty = m.dy; }
    if (row_has_top && bx + 1 < bw) { MV m = get(bx + 1, 1); trx = m.dx; try_ = m.dy; }
    else if (row_has_top && bx > 0) { MV m = get(bx - 1, 1); trx = m.dx; try_ = m.dy; }
    int mdx = median3(lx, tx, trx), mdy = median3(ly, ty, try_);
    std::vector<std::pair<int, int>> cands;
    push_candidate(cands, 0, 0, sr);
    for (auto& p : preds) push_candidate(cands, p.first, p.second, sr);
    return process_block_hex(bx, by, W, H, sr, cands, prev, cur, zthr, pthr, lambda,
                             uniq[0].first, uniq[0].second, uniq[1].first, uniq[1].second, mdx, mdy);
}

// hex_search_impl  (scalar.rs:233-449). threads == 0 -> serial raster order (reference
// default features); threads > 0 -> the `threads` feature's row strips (:250-407),
// target_strips = 4*threads, whose MVs differ from the serial ones.
static void hex_search(const uint8_t* prev, const uint8_t* cur, int W, int H, int sr, const std::vector<MV>* prev_mvs,
                       int64_t zthr, int64_t pthr, double lambda, int threads, std::vector<MV>& mvs, std::vector<int64_t>& sads) {
    int bw = (W + 15) / 16, bh = (H + 15) / 16;
    size_t n = (size_t)bw * bh;
    mvs.assign(n, MV()); sads.assign(n, 0);
    if (threads <= 0) {
        for (int by = 0; by < bh; by++)
            for (int bx = 0; bx < bw; bx++) {
                auto get = [&](int x, int up) -> MV { return mvs[(size_t)(by - up) * bw + x]; };
                MeResult r = me_block(get, by > 0, prev_mvs, true, n, bw, bx, by, W, H, sr, prev, cur, zthr, pthr, lambda);
                mvs[(size_t)by * bw + bx] = r.mv; sads[(size_t)by * bw + bx] = r.sad;
            }
        return;
    }
    int target = threads * 4;
    int rows_per_strip = std::max(1, (bh + target - 1) / target);
    int nstrips = (bh + rows_per_strip - 1) / rows_per_strip;
    auto work = [&](int s0, int s1) {
        for (int s = s0; s < s1; s++) {
            int r0 = s * rows_per_strip, r1 = std::min(bh, r0 + rows_per_strip);
            for (int by = r0; by < r1; by++)
                for (int bx = 0; bx < bw; bx++) {
                    auto get = [&](int x, int up) -> MV { return mvs[(size_t)(by - up) * bw + x]; };
                    MeResult r = me_block(get, by > r0, prev_mvs, false, n, bw, bx, by, W, H, sr, prev, cur, zthr, pthr, lambda);
                    mvs[(size_t)by * bw + bx] = r.mv; sads[(size_t)by * bw + bx] = r.sad;
                }
        }
    };
    std::vector<std::thread> pool;
    // strips are independent; hand them out round-robin (rayon's scheduling does not affect results)
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&, t]() { for (int s = t; s < nstrips; s += threads) work(s, s + 1); });
    for (auto& th : pool) th.join();
}

// ---------------------------------------------------------------------------
// Motion compensation  (reitero_video_common/src/fast_motion.rs:103-352 and the
// per-pixel samplers motion.rs:28-117; rounding interpolation)
// ---------------------------------------------------------------------------
static inline uint8_t sample_plane_hp(const uint8_t* p, int W, int H, int xh, int yh) {   // motion.rs:28-60
    xh = clampv(xh, 0, (W - 1) * 2); yh = clampv(yh, 0, (H - 1) * 2);
    int x0 = xh / 2, y0 = yh / 2;
    bool xo = xh & 1, yo = yh & 1;
    int x1 = x0 + 1 < W ? x0 + 1 : x0, y1 = y0 + 1 < H ? y0 + 1 : y0;
    int a = p[(size_t)y0 * W + x0];
    if (!xo && !yo) return (uint8_t)a;
    int b = p[(size_t)y0 * W + x1], d = p[(size_t)y1 * W + x0], e = p[(size_t)y1 * W + x1];
    if (xo && !yo) return (uint8_t)((a + b + 1) / 2);
    if (!xo && yo) return (uint8_t)((a + d + 1) / 2);
    return (uint8_t)((a + b + d + e + 2) / 4);
}
static void build_predicted(const Yuv& prev, const std::vector<MV>& mvs, Yuv& out) {
    int W = prev.w, H = prev.h, cw = W / 2, ch = H / 2;
    int bw = (W + 15) / 16, bh = (H + 15) / 16;
    out = Yuv(W, H);
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            const MV& mv = mvs[(size_t)by * bw + bx];
            int mvx = mv.dx_hp(), mvy = mv.dy_hp();
            int x0 = bx * 16, y0 = by * 16;
            if ((mvx & 1) == 0 && (mvy & 1) == 0) {   // integer fast path (fast_motion.rs:133-198)
                int sx0 = x0 + mvx / 2, sy0 = y0 + mvy / 2;
                if (sx0 >= 0 && sy0 >= 0 && sx0 + 15 < W && sy0 + 15 < H) {
                    for (int r = 0; r < 16; r++)
                        memcpy(&out.y[(size_t)(y0 + r) * W + x0], &prev.y[(size_t)(sy0 + r) * W + sx0], 16);
                    int csx = sx0 / 2, csy = sy0 / 2;   // chroma copied at floor((x0+dx)/2): the 1/2 chroma phase is dropped
                    for (int r = 0; r < 8; r++) {
                        memcpy(&out.u[(size_t)(y0 / 2 + r) * cw + x0 / 2], &prev.u[(size_t)(csy + r) * cw + csx], 8);
                        memcpy(&out.v[(size_t)(y0 / 2 + r) * cw + x0 / 2], &prev.v[(size_t)(csy + r) * cw + csx], 8);
                    }
                    continue;
                }
            }
            for (int yy = 0; yy < 16; yy++) {   // general path
                int y = y0 + yy;
                for (int xx = 0; xx < 16; xx++) {
                    int x = x0 + xx;
                    out.y[(size_t)y * W + x] = sample_plane_hp(prev.y.data(), W, H, 2 * x + mvx, 2 * y + mvy);
                    if ((y & 1) == 0 && (x & 1) == 0) {
                        int cxh = (2 * x + mvx) >> 1, cyh = (2 * y + mvy) >> 1;
                        out.u[(size_t)(y / 2) * cw + x / 2] = sample_plane_hp(prev.u.data(), cw, ch, cxh, cyh);
                        out.v[(size_t)(y / 2) * cw + x / 2] = sample_plane_hp(prev.v.data(), cw, ch, cxh, cyh);
                    }
                }
            }
        }
}

// ---------------------------------------------------------------------------
// Quantisation helpers  (reitero_residual/src/residual.rs)
// ---------------------------------------------------------------------------
static float quant_step_from_quality(uint8_t q) {   // residual.rs:13-23
    if (q == 100) return 1.0f;
    if (q == 1) return 112.0f;
    float qf = (float)q;
    float range = 112.0f - 1.0f;
    return 1.0f + (100.0f - qf) / 99.0f * range;
}
static const float JPEG_LUMA_8X8[64] = {   // residual.rs:32-41
    16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69, 56,
    14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81, 104, 113, 92,
    49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
static const float JPEG_CHROMA_8X8[64] = {   // residual.rs:45-54
    17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99, 24, 26, 56, 99, 99, 99, 99, 99,
    47, 66, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
    99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
static void build_intra_y_table(float qs, uint16_t* t) {   // residual.rs:58-85
    for (int v = 0; v < 16; v++)
        for (int u = 0; u < 16; u++) {
            float fx = (float)u * 7.0f / 15.0f, fy = (float)v * 7.0f / 15.0f;
            int x0 = (int)fx, y0 = (int)fy;
            int x1 = std::min(x0 + 1, 7), y1 = std::min(y0 + 1, 7);
            float dx = fx - (float)x0, dy = fy - (float)y0;
            float val = JPEG_LUMA_8X8[y0 * 8 + x0] * (1.0f - dx) * (1.0f - dy) + JPEG_LUMA_8X8[y0 * 8 + x1] * dx * (1.0f - dy) +
                        JPEG_LUMA_8X8[y1 * 8 + x0] * (1.0f - dx) * dy + JPEG_LUMA_8X8[y1 * 8 + x1] * dx * dy;
            float scaled = (val / JPEG_LUMA_8X8[0]) * qs;
            t[v * 16 + u] = f32_to_u16_sat(std::max(roundf(scaled), 1.0f));
        }
}
static void build_intra_uv_table(float qs, uint16_t* t) {   // residual.rs:89-96
    for (int i = 0; i < 64; i++) t[i] = f32_to_u16_sat(std::max(roundf((JPEG_CHROMA_8X8[i] / JPEG_CHROMA_8X8[0]) * qs), 1.0f));
}
// compute_aq_quant_steps  (residual.rs:116-167): glibc logf/exp2f, sequential f32 sum
static std::vector<float> compute_aq_steps(const uint8_t* y, int stride, int W, int H, float base) {
    int bw = W / 16, bh = H / 16, nb = bw * bh;
    std::vector<float> var(nb);
    for (int by = 0; by < bh; by++)
        for (int bx = 0; bx < bw; bx++) {
            uint64_t s = 0, s2 = 0;
            for (int j = 0; j < 16; j++)
                for (int i = 0; i < 16; i++) { uint64_t v = y[(size_t)(by * 16 + j) * stride + bx * 16 + i]; s += v; s2 += v * v; }
            float vv = (float)(s2 * 256 - s * s) / (float)(256 * 256);
            var[by * bw + bx] = std::max(vv, 4.0f);
        }
    float sum = 0.0f;
    for (int i = 0; i < nb; i++) sum += logf(var[i]);
    float avg = sum / (float)nb;
    std::vector<float> out(nb);
    for (int i = 0; i < nb; i++) {
        float off = 0.8f * (logf(var[i]) - avg);
        out[i] = base * exp2f(off / 6.0f);
    }
    return out;
}

// ---------------------------------------------------------------------------
// Transforms  (reitero_dct/src/dct16.rs, dct8.rs, fast_dct8.rs)
// Basis: round(2^18 * c_k * cos((2n+1)k*pi/2N)), c_0 = sqrt(1/N), c_k = sqrt(2/N); this
// closed form reproduces FDCT_TRANSPOSE_16 (dct16.rs:7-74), IDCT_ROWS_16 (its transpose,
// dct16.rs:78-145) and FDCT_TRANSPOSE_8 (dct8.rs:6-25) entry for entry.
// ---------------------------------------------------------------------------
static int32_t B16[16][16], B8[8][8];   // [n = sample][k = frequency]
static bool g_basis_init = false;
static void init_basis() {
    if (g_basis_init) return;
    for (int n = 0; n < 16; n++)
        for (int k = 0; k < 16; k++)
            B16[n][k] = (int32_t)llround(262144.0 * (k == 0 ? sqrt(1.0 / 16) : sqrt(2.0 / 16)) * cos((2 * n + 1) * k * M_PI / 32.0));
    for (int n = 0; n < 8; n++)
        for (int k = 0; k < 8; k++)
            B8[n][k] = (int32_t)llround(262144.0 * (k == 0 ? sqrt(1.0 / 8) : sqrt(2.0 / 8)) * cos((2 * n + 1) * k * M_PI / 16.0));
    g_basis_init = true;
}
// fdct_1d_16 / fdct_1d_8: i64 accumulation, +2^17 >> 18 (dct16.rs:151-194, dct8.rs:32-54)
template <int N>
static void fdct_1d(const int32_t* in, int32_t* out, const int32_t (*B)[N]) {
    for (int k = 0; k < N; k++) {
        int64_t acc = 0;
        for (int n = 0; n < N; n++) acc += (int64_t)in[n] * (int64_t)B[n][k];
        out[k] = (int32_t)clampv<int64_t>((acc + (1ll << 17)) >> 18, INT32_MIN, INT32_MAX);
    }
}
// rows, then columns; coefficient (u = horizontal, v = vertical) at out[v*N+u]
template <int N>
static void fdct_2d(const int16_t* blk, int stride, int32_t* coef, const int32_t (*B)[N]) {
    int32_t tmp[N][N], in[N], o[N];
    for (int y = 0; y < N; y++) { for (int x = 0; x < N; x++) in[x] = blk[y * stride + x]; fdct_1d<N>(in, tmp[y], B); }
    for (int u = 0; u < N; u++) {
        for (int y = 0; y < N; y++) in[y] = tmp[y][u];
        fdct_1d<N>(in, o, B);
        for (int v = 0; v < N; v++) coef[v * N + u] = o[v];
    }
}
// idct_1d_16: wrapping i32 MAC, +2^17 >> 18 (dct16.rs:201-233); rows first, then columns
static void idct16_2d(const int32_t* c, int16_t* out, int stride) {
    int32_t tmp[16][16];
    for (int v = 0; v < 16; v++)
        for (int x = 0; x < 16; x++) {
            int32_t acc = 0;
            for (int u = 0; u < 16; u++) acc = wadd(acc, wmul(c[v * 16 + u], B16[x][u]));
            tmp[v][x] = wadd(acc, 1 << 17) >> 18;
        }
    for (int x = 0; x < 16; x++)
        for (int y = 0; y < 16; y++) {
            int32_t acc = 0;
            for (int v = 0; v < 16; v++) acc = wadd(acc, wmul(tmp[v][x], B16[y][v]));
            out[y * stride + x] = (int16_t)clampv<int32_t>(wadd(acc, 1 << 17) >> 18, -32768, 32767);
        }
}
// AAN/stb 8x8 IDCT with the reference's F2F constants (fast_dct8.rs:123-174, :254-330)
static const int32_t F2F_0541 = 2218, F2F_1848 = -7572, F2F_0765 = 3136, F2F_1176 = 4817, F2F_0299 = 1223, F2F_2053 = 8411,
                     F2F_3073 = 12595, F2F_1501 = 6150, F2F_0900 = -3685, F2F_2563 = -10496, F2F_1962 = -8035, F2F_0390 = -1598;
static void aan_kernel(const int32_t* s, int32_t xsc, int32_t* x, int32_t* t) {
    // kernel_x_simd
    int32_t p2 = s[2], p3 = s[6];
    int32_t p1 = wmul(wadd(p2, p3), F2F_0541);
    int32_t t2 = wadd(p1, wmul(p3, F2F_1848));
    int32_t t3 = wadd(p1, wmul(p2, F2F_0765));
    p2 = s[0]; p3 = s[4];
    int32_t t0 = wshl(wadd(p2, p3), 12), t1 = wshl(wsub(p2, p3), 12);
    x[0] = wadd(wadd(t0, t3), xsc); x[3] = wadd(wsub(t0, t3), xsc);
    x[1] = wadd(wadd(t1, t2), xsc); x[2] = wadd(wsub(t1, t2), xsc);
    // kernel_t_simd
    int32_t a0 = s[7], a1 = s[5], a2 = s[3], a3 = s[1];
    int32_t q3 = wadd(a0, a2), q4_ = wadd(a1, a3), q1 = wadd(a0, a3), q2 = wadd(a1, a2);
    int32_t p5 = wmul(wadd(q3, q4_), F2F_1176);
    a0 = wmul(a0, F2F_0299); a1 = wmul(a1, F2F_2053); a2 = wmul(a2, F2F_3073); a3 = wmul(a3, F2F_1501);
    q1 = wadd(p5, wmul(q1, F2F_0900)); q2 = wadd(p5, wmul(q2, F2F_2563));
    q3 = wmul(q3, F2F_1962); q4_ = wmul(q4_, F2F_0390);
    a3 = wadd(wadd(a3, q1), q4_); a2 = wadd(wadd(a2, q2), q3); a1 = wadd(wadd(a1, q2), q4_); a0 = wadd(wadd(a0, q1), q3);
    t[0] = a0; t[1] = a1; t[2] = a2; t[3] = a3;
}
static void idct8_aan(const int16_t* coef, const uint16_t* q, int16_t* out, int stride) {
    int32_t tmp[8][8];   // tmp[x][r]: first pass over u for every row r
    for (int r = 0; r < 8; r++) {
        int32_t s[8], x[4], t[4];
        for (int u = 0; u < 8; u++) s[u] = wmul((int32_t)coef[r * 8 + u], (int32_t)q[r * 8 + u]);
        aan_kernel(s, 512, x, t);
        tmp[0][r] = wadd(x[0], t[3]) >> 10; tmp[7][r] = wsub(x[0], t[3]) >> 10;
        tmp[1][r] = wadd(x[1], t[2]) >> 10; tmp[6][r] = wsub(x[1], t[2]) >> 10;
        tmp[2][r] = wadd(x[2], t[1]) >> 10; tmp[5][r] = wsub(x[2], t[1]) >> 10;
        tmp[3][r] = wadd(x[3], t[0]) >> 10; tmp[4][r] = wsub(x[3], t[0]) >> 10;
    }
    for (int xcol = 0; xcol < 8; xcol++) {   // second pass over r for every column
        int32_t s[8], x[4], t[4], o[8];
        for (int r = 0; r < 8; r++) s[r] = tmp[xcol][r];
        aan_kernel(s, 65536, x, t);
        o[0] = wadd(x[0], t[3]) >> 17; o[1] = wadd(x[1], t[2]) >> 17; o[2] = wadd(x[2], t[1]) >> 17; o[3] = wadd(x[3], t[0]) >> 17;
        o[4] = wsub(x[3], t[0]) >> 17; o[5] = wsub(x[2], t[1]) >> 17; o[6] = wsub(x[1], t[2]) >> 17; o[7] = wsub(x[0], t[3]) >> 17;
        for (int y = 0; y < 8; y++) out[y * stride + xcol] = (int16_t)clampv<int32_t>(o[y], -32768, 32767);
    }
}
// encode_plane_16x16_aq / encode_plane_8x8_aq quantiser (dct16.rs:321-333, dct8.rs:122-129)
static inline int16_t quant_aq(int32_t c, float qs, float dz) {
    float r = (float)c / qs;
    return fabsf(r) < dz ? 0 : f32_to_i16_sat(roundf(r));
}
// encode_plane_*_matrix quantiser (dct16.rs:467-480, dct8.rs:183-188)
static inline int16_t quant_matrix(int32_t c, uint16_t q) {
    float qs = (float)std::max<uint16_t>(q, 1);
    return f32_to_i16_sat(roundf((float)c / qs));
}
// decode_plane_16x16_aq dequant (dct16.rs:371-381)
static void dequant16_aq(const int16_t* c, float qs, int32_t* o) {
    int32_t qfp = f32_to_i32_sat(roundf(qs * 16384.0f));
    for (int i = 0; i < 256; i++) o[i] = wadd(wmul((int32_t)c[i], qfp), 1 << 13) >> 14;
}

// ---------------------------------------------------------------------------
// RDO  (reitero_encode/src/rdo.rs)
// ---------------------------------------------------------------------------
struct Lambda { double lambda, dr, slope, intercept, skip_bits, keep_bits; };
static Lambda lambda_for_quality(uint8_t quality, double mult) {   // rdo.rs:23-49, :75-86
    quality = clampv<uint8_t>(quality, 1, 100);
    uint32_t inv = 100u - quality;
    int qp = clampv((int)round((double)inv * 0.6), 0, 60);
    Lambda e;
    e.lambda = mult * pow(2.0, ((double)qp - 12.0) / 3.0);
    double nq = (double)quality / 100.0;
    double q = (double)quant_step_from_quality(quality);
    e.slope = 0.5 / q;
    e.intercept = 5.0 + 0.56 * q;
    e.dr = 0.15 + ((q - 1.0) / 111.0) * 0.55;
    e.skip_bits = 0.55 - nq * 0.15;
    e.keep_bits = 1.20 + nq * 0.35;
    return e;
}
// RdoContext::decide  (rdo.rs:140-178)
static bool rdo_decide(const Lambda& e, uint8_t skip_threshold, int64_t satd_i, bool integer_aligned) {
    if (skip_threshold == 0) return false;
    if (!integer_aligned) return false;
    double satd = (double)std::max<int64_t>(satd_i, 0);
    double guard = 768.0 * (double)skip_threshold;
    if (satd > guard) return false;
    double skip_cost = satd + e.lambda * e.skip_bits;
    double est = e.intercept + e.slope * satd;
    double rate = e.keep_bits + std::max(est, 0.0);
    double coded = satd * e.dr + e.lambda * rate;
    return skip_cost <= coded;
}

// ---------------------------------------------------------------------------
// Deblock  (reitero_video_common/src/deblock.rs)
// ---------------------------------------------------------------------------
static int deblock_level_from_qs(float qs) { float r = roundf(qs * 0.6f); r = clampv(r, 0.0f, 63.0f); return (int)r; }   // :34-36
static void thresholds(int level, bool intra, int& e, int& il, int& hev) {   // :40-58
    il = std::max(level, 1);
    e = (level + 2) * 2 + il;
    if (intra) hev = level >= 40 ? 2 : (level >= 15 ? 1 : 0);
    else hev = level >= 40 ? 3 : (level >= 20 ? 2 : (level >= 15 ? 1 : 0));
}
static inline int c8(int x) { return clampv(x, -128, 127); }
static void mb_filter_line(int* px, int E, int I, int HEV) {   // :68-115
    int p3 = px[0], p2 = px[1], p1 = px[2], p0 = px[3], q0 = px[4], q1 = px[5], q2 = px[6], q3 = px[7];
    bool mask = 2 * std::abs(p0 - q0) + std::abs(q1 - p1) / 2 <= E && std::abs(p3 - p2) <= I && std::abs(p2 - p1) <= I &&
                std::abs(p1 - p0) <= I && std::abs(q1 - q0) <= I && std::abs(q2 - q1) <= I && std::abs(q3 - q2) <= I;
    if (!mask) return;
    int ps2 = p2 - 128, ps1 = p1 - 128, ps0 = p0 - 128, qs0 = q0 - 128, qs1 = q1 - 128, qs2 = q2 - 128;
    bool hev = std::abs(p1 - p0) > HEV || std::abs(q1 - q0) > HEV;
    int w = c8(c8(ps1 - qs1) + 3 * (qs0 - ps0));
    if (hev) {
        int f1 = c8(w + 4) >> 3, f2 = c8(w + 3) >> 3;
        px[4] = clampv(c8(qs0 - f1) + 128, 0, 255);
        px[3] = clampv(c8(ps0 + f2) + 128, 0, 255);
    } else {
        int a = c8((27 * w + 63) >> 7);
        int nq0 = c8(qs0 - a), np0 = c8(ps0 + a);
        a = c8((18 * w + 63) >> 7);
        int nq1 = c8(qs1 - a), np1 = c8(ps1 + a);
        a = c8((9 * w + 63) >> 7);
        int nq2 = c8(qs2 - a), np2 = c8(ps2 + a);
        px[1] = clampv(np2 + 128, 0, 255); px[2] = clampv(np1 + 128, 0, 255); px[3] = clampv(np0 + 128, 0, 255);
        px[4] = clampv(nq0 + 128, 0, 255); px[5] = clampv(nq1 + 128, 0, 255); px[6] = clampv(nq2 + 128, 0, 255);
    }
}
static void deblock_plane(uint8_t* p, int w, int h, int grid, int E, int I, int HEV, const std::vector<uint8_t>* mask) {   // :129-177
    int bw = w / grid;
    auto active = [&](int ax, int ay, int bx, int by) { return !mask || (*mask)[ay * bw + ax] || (*mask)[by * bw + bx]; };
    int px[8];
    for (int x = grid; x < w; x += grid) {
        int bx = x / grid;
        for (int y = 0; y < h; y++) {
            if (!active(bx - 1, y / grid, bx, y / grid)) continue;
            size_t i = (size_t)y * w + x;
            for (int k = 0; k < 8; k++) px[k] = p[i - 4 + k];
            mb_filter_line(px, E, I, HEV);
            for (int k = 0; k < 8; k++) p[i - 4 + k] = (uint8_t)px[k];
        }
    }
    for (int y = grid; y < h; y += grid) {
        int by = y / grid;
        for (int x = 0; x < w; x++) {
            if (!active(x / grid, by - 1, x / grid, by)) continue;
            for (int k = 0; k < 8; k++) px[k] = p[(size_t)(y - 4 + k) * w + x];
            mb_filter_line(px, E, I, HEV);
            for (int k = 0; k < 8; k++) p[(size_t)(y - 4 + k) * w + x] = (uint8_t)px[k];
        }
    }
}
static void deblock_yuv420(Yuv& f, int level, bool intra, const std::vector<uint8_t>* mask) {   // :189-202
    if (level <= 0) return;
    int E, I, HEV; thresholds(level, intra, E, I, HEV);
    deblock_plane(f.y.data(), f.w, f.h, 16, E, I, HEV, mask);
    deblock_plane(f.u.data(), f.w / 2, f.h / 2, 8, E, I, HEV, mask);
    deblock_plane(f.v.data(), f.w / 2, f.h / 2, 8, E, I, HEV, mask);
}

// ---------------------------------------------------------------------------
// Residual encode/decode glue  (reitero_residual/src/residual.rs)
// ---------------------------------------------------------------------------
// ResidualEncoder::encode_intra  (residual.rs:267-388)
static std::vector<uint8_t> encode_intra(const Yuv& cur, uint8_t quality, Yuv& recon, std::vector<int16_t>& coeffs) {
    int W = cur.w, H = cur.h, cw = W / 2, bw = W / 16, bh = H / 16, nb = bw * bh;
    float qs = quant_step_from_quality(quality);
    uint16_t yt[256], ct[64];
    build_intra_y_table(qs, yt); build_intra_uv_table(qs, ct);
    coeffs.assign((size_t)nb * 384, 0);
    DctRansEnc enc(true);
    recon = Yuv(W, H);
    for (int b = 0; b < nb; b++) {
        int bx = b % bw, by = b / bw;
        int16_t res[256], cu[64], cv[64];
        for (int j = 0; j < 16; j++) for (int i = 0; i < 16; i++) res[j * 16 + i] = (int16_t)(cur.y[(size_t)(by * 16 + j) * W + bx * 16 + i] - 128);
        for (int j = 0; j < 8; j++) for (int i = 0; i < 8; i++) {
            size_t o = (size_t)(by * 8 + j) * cw + bx * 8 + i;
            cu[j * 8 + i] = (int16_t)(cur.u[o] - 128); cv[j * 8 + i] = (int16_t)(cur.v[o] - 128);
        }
        int16_t* cb = &coeffs[(size_t)b * 384];
        int32_t c[256];
        fdct_2d<16>(res, 16, c, B16);
        for (int i = 0; i < 256; i++) cb[i] = quant_matrix(c[i], yt[i]);
        fdct_2d<8>(cu, 8, c, B8);
        for (int i = 0; i < 64; i++) cb[256 + i] = quant_matrix(c[i], ct[i]);
        fdct_2d<8>(cv, 8, c, B8);
        for (int i = 0; i < 64; i++) cb[320 + i] = quant_matrix(c[i], ct[i]);
        enc.encode_block(cb, cb + 256, cb + 320);
    }
    for (int b = 0; b < nb; b++) {   // decode_plane_*_matrix, then +128
        int bx = b % bw, by = b / bw;
        const int16_t* cb = &coeffs[(size_t)b * 384];
        int32_t d[256]; int16_t r[256];
        for (int i = 0; i < 256; i++) d[i] = wmul((int32_t)cb[i], (int32_t)std::max<uint16_t>(yt[i], 1));   // dct16.rs:526-534
        idct16_2d(d, r, 16);
        for (int j = 0; j < 16; j++) for (int i = 0; i < 16; i++)
            recon.y[(size_t)(by * 16 + j) * W + bx * 16 + i] = (uint8_t)clampv(r[j * 16 + i] + 128, 0, 255);
        for (int pl = 0; pl < 2; pl++) {
            int16_t rc[64];
            idct8_aan(cb + 256 + pl * 64, ct, rc, 8);
            std::vector<uint8_t>& dst = pl == 0 ? recon.u : recon.v;
            for (int j = 0; j < 8; j++) for (int i = 0; i < 8; i++) dst[(size_t)(by * 8 + j) * cw + bx * 8 + i] = (uint8_t)clampv(rc[j * 8 + i] + 128, 0, 255);
        }
    }
    return enc.finish();
}

// residual reconstruction for one inter MB: dequant + IDCT + add/clamp
static void recon_inter_block(const int16_t* cb, float qs, const Yuv& pred, Yuv& rec, int bx, int by) {
    int W = pred.w, cw = W / 2;
    int32_t d[256]; int16_t r[256];
    dequant16_aq(cb, qs, d);
    idct16_2d(d, r, 16);
    for (int j = 0; j < 16; j++) for (int i = 0; i < 16; i++) {
        size_t o = (size_t)(by * 16 + j) * W + bx * 16 + i;
        rec.y[o] = (uint8_t)clampv((int)pred.y[o] + r[j * 16 + i], 0, 255);
    }
    uint16_t q = f32_to_u16_sat(std::max(roundf(qs), 1.0f));   // fast_dct8.rs:95
    uint16_t qt[64]; for (int i = 0; i < 64; i++) qt[i] = q;
    for (int pl = 0; pl < 2; pl++) {
        int16_t rc[64];
        idct8_aan(cb + 256 + pl * 64, qt, rc, 8);
        const std::vector<uint8_t>& ps = pl == 0 ? pred.u : pred.v;
        std::vector<uint8_t>& dst = pl == 0 ? rec.u : rec.v;
        for (int j = 0; j < 8; j++) for (int i = 0; i < 8; i++) {
            size_t o = (size_t)(by * 8 + j) * cw + bx * 8 + i;
            dst[o] = (uint8_t)clampv((int)ps[o] + rc[j * 8 + i], 0, 255);
        }
    }
}

// ResidualEncoder::encode_inter  (residual.rs:396-809)
static std::vector<uint8_t> encode_inter(const Yuv& cur, const Yuv& pred, const std::vector<uint8_t>& skip_mask, uint8_t q, float dz,
                                         Yuv& recon, std::vector<uint8_t>& opt_skip, std::vector<int16_t>& coeffs, std::vector<float>& aq) {
    int W = cur.w, H = cur.h, cw = W / 2, bw = W / 16, bh = H / 16, nb = bw * bh;
    float base = quant_step_from_quality(q);
    aq = compute_aq_steps(pred.y.data(), W, W, H, base);
    coeffs.assign((size_t)nb * 384, 0);
    opt_skip.assign(nb, 0);
    DctRansEnc* enc = nullptr;
    for (int b = 0; b < nb; b++) {
        if (skip_mask[b]) { opt_skip[b] = 1; continue; }
        int bx = b % bw, by = b / bw;
        int16_t ry[256], ru[64], rv[64];
        for (int j = 0; j < 16; j++) for (int i = 0; i < 16; i++) {
            size_t o = (size_t)(by * 16 + j) * W + bx * 16 + i;
            ry[j * 16 + i] = (int16_t)((int)cur.y[o] - (int)pred.y[o]);
        }
        for (int j = 0; j < 8; j++) for (int i = 0; i < 8; i++) {
            size_t o = (size_t)(by * 8 + j) * cw + bx * 8 + i;
            ru[j * 8 + i] = (int16_t)((int)cur.u[o] - (int)pred.u[o]);
            rv[j * 8 + i] = (int16_t)((int)cur.v[o] - (int)pred.v[o]);
        }
        int16_t* cb = &coeffs[(size_t)b * 384];
        int32_t c[256];
        float qs = aq[b];
        fdct_2d<16>(ry, 16, c, B16);
        for (int i = 0; i < 256; i++) cb[i] = quant_aq(c[i], qs, dz);
        fdct_2d<8>(ru, 8, c, B8);
        for (int i = 0; i < 64; i++) cb[256 + i] = quant_aq(c[i], qs, dz);
        fdct_2d<8>(rv, 8, c, B8);
        for (int i = 0; i < 64; i++) cb[320 + i] = quant_aq(c[i], qs, dz);
        bool zero = true;
        for (int i = 0; i < 384; i++) if (cb[i]) { zero = false; break; }
        if (zero) { opt_skip[b] = 1; continue; }
        if (!enc) enc = new DctRansEnc(false);
        enc->encode_block(cb, cb + 256, cb + 320);
    }
    recon = pred;
    for (int b = 0; b < nb; b++) if (!opt_skip[b]) recon_inter_block(&coeffs[(size_t)b * 384], aq[b], pred, recon, b % bw, b / bw);
    std::vector<uint8_t> data;
    if (enc) { data = enc->finish(); delete enc; }
    return data;
}

// ---------------------------------------------------------------------------
// Container  (reitero_video_common/src/format.rs)
// ---------------------------------------------------------------------------
static void put_u32(std::vector<uint8_t>& b, uint32_t v) { for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i))); }
static void put_u64(std::vector<uint8_t>& b, uint64_t v) { for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i))); }
static uint32_t get_u32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
static uint64_t get_u64(const uint8_t* p) { return (uint64_t)get_u32(p) | ((uint64_t)get_u32(p + 4) << 32); }
struct Header { uint32_t dw, dh, sw, sh, fps; uint64_t frame_count; };
static std::vector<uint8_t> header_bytes(const Header& h) {   // format.rs:64-88
    std::vector<uint8_t> b = {'R', 'I', 'V', 0};
    put_u32(b, 5); put_u32(b, h.dw); put_u32(b, h.dh); put_u32(b, h.sw); put_u32(b, h.sh); put_u32(b, h.fps); put_u64(b, h.frame_count);
    return b;
}

}  // namespace

// ===========================================================================
// C API (ctypes)
// ===========================================================================
extern "C" {

typedef struct {
    uint32_t display_width, display_height, fps;
    int32_t keyframe_mode;        // 0 = Fixed(keyframe_interval), 1 = AllIntra, 2 = Automatic
    uint64_t keyframe_interval;
    uint8_t intra_quality, inter_quality, search_range, skip_threshold;
    int64_t me_zero_mv_threshold, me_predictor_threshold;
    double rdo_lambda_mult;
    float inter_dead_zone;
    int32_t me_threads;           // 0 = serial raster ME (bit-exact); T > 0 = `threads` strip emulation
} ro_config;

typedef struct {
    uint64_t frame_index;
    int32_t frame_type;           // 1 = intra, 2 = inter
    uint32_t total_bytes, mv_bytes, residual_bytes;
    uint32_t blocks_total, blocks_skipped, rdo_skips;
    uint32_t mode_counts[7];
    int32_t bias_dx, bias_dy;
    uint32_t new_blocks;
} ro_frame_stats;

struct ro_encoder {
    ro_config cfg;
    Header hdr;
    std::vector<uint8_t> out;
    bool have_prev = false;
    Yuv prev_recon;
    std::vector<MV> prev_mvs; bool have_prev_mvs = false;
    MvCtx mvctx;
    uint64_t frame_count = 0;
    // last-frame dumps for parity tests
    Yuv cur, pred;
    std::vector<MV> search_mvs, recon_mvs;
    std::vector<int64_t> sads;
    std::vector<uint8_t> modes, new_base, rdo_skip, opt_skip;
    std::vector<int8_t> raw_dx, raw_dy;
    std::vector<float> aq;
    std::vector<int16_t> coeffs;
    std::vector<uint8_t> last_frame_bytes;
};


// Mode decision over the reconstructed prefix, in raster order (encoder.rs:447-733): mode by
// exact half-pel equality in the order Zero, Nearest, Near, TopRight, TopLeft, Temporal, else NEW;
// the NEW base by estimated bits (encoder.rs:128-159, :567-661); the reconstructed MV keeps the
// predictor's representation (:667-718).
static void decide_modes(const std::vector<MV>& smv, const std::vector<MV>* pm, int bw, int bh, std::vector<uint8_t>& modes,
                         std::vector<uint8_t>& new_base, std::vector<int8_t>& raw_dx, std::vector<int8_t>& raw_dy, std::vector<MV>& coded) {
    const int nb = bw * bh;
    coded.clear(); coded.reserve(nb);
    modes.assign(nb, 0); new_base.assign(nb, 0); raw_dx.assign(nb, 0); raw_dy.assign(nb, 0);
    for (int i = 0; i < nb; i++) {
        int bx = i % bw, by = i / bw;
        const MV& mv = smv[i];
        Preds p = derive_predictors(coded, pm, bw, bh, bx, by);
        Neigh n = gather_neighbors(coded, pm, bw, bh, bx, by);
        bool zero_el = mv.dx == 0 && mv.dy == 0 && mv.sx == 0 && mv.sy == 0;
        uint8_t mode;
        auto hp = [](const Q4& q, int& x, int& y) { x = q.dx * 2 + q.sx; y = q.dy * 2 + q.sy; };
        if (zero_el) mode = ZERO;
        else {
            int mx = mv.dx_hp(), my = mv.dy_hp(), ax, ay;
            hp(p.nearest, ax, ay);
            if (mx == ax && my == ay) mode = NEAREST;
            else {
                hp(p.near, ax, ay);
                if (mx == ax && my == ay) mode = NEAR;
                else {
                    bool trm = false, tlm = false;
                    if (n.htr) { hp(n.tr, ax, ay); trm = mx == ax && my == ay; }
                    if (trm) mode = TOPRIGHT;
                    else {
                        if (n.htl) { hp(n.tl, ax, ay); tlm = mx == ax && my == ay; }
                        if (tlm) mode = TOPLEFT;
                        else { hp(p.temporal, ax, ay); mode = (mx == ax && my == ay) ? TEMPORAL : NEWMV; }
                    }
                }
            }
        }
        uint8_t chosen = 0; int8_t rdx = 0, rdy = 0;
        if (mode == NEWMV) {
            int adx = mv.dx, ady = mv.dy;
            bool have[5] = {true, true, n.htr, n.htl, true};
            int bxs[5] = {p.nearest.dx, p.near.dx, n.tr.dx, n.tl.dx, p.temporal.dx};
            int bys[5] = {p.nearest.dy, p.near.dy, n.tr.dy, n.tl.dy, p.temporal.dy};
            static const uint32_t selbits[5] = {1, 2, 3, 4, 4};
            uint32_t best = UINT32_MAX; int bi = 0; int bbx = p.nearest.dx, bby = p.nearest.dy;
            for (int k = 0; k < 5; k++) {
                if (!have[k]) continue;
                int8_t ddx = (int8_t)clampv(adx - bxs[k], -128, 127), ddy = (int8_t)clampv(ady - bys[k], -128, 127);
                auto cbits = [](int8_t d) { int cl = mv_class_from_magnitude(std::abs((int)d)); return cl == 0 ? 4u : 4u + (uint32_t)cl; };
                uint32_t cost = selbits[k] + cbits(ddx) + cbits(ddy);
                if (cost < best) { best = cost; bi = k; bbx = bxs[k]; bby = bys[k]; }
            }
            chosen = (uint8_t)bi;
            rdx = (int8_t)clampv(adx - bbx, -128, 127); rdy = (int8_t)clampv(ady - bby, -128, 127);
        }
        modes[i] = mode; new_base[i] = chosen; raw_dx[i] = rdx; raw_dy[i] = rdy;
        MV rec = mv;
        auto from = [&](const Q4& q) { return MV::make(q.dx, q.dy, q.sx, q.sy, mv.skip); };
        if (mode == NEAREST) rec = from(p.nearest);
        else if (mode == NEAR) rec = from(p.near);
        else if (mode == TOPRIGHT) { if (n.htr) rec = from(n.tr); }
        else if (mode == TOPLEFT) { if (n.htl) rec = from(n.tl); }
        else if (mode == TEMPORAL) rec = from(p.temporal);
        else if (mode == NEWMV) {
            Q4 base = chosen == 0 ? p.nearest : chosen == 1 ? p.near : chosen == 2 ? (n.htr ? n.tr : p.nearest)
                      : chosen == 3 ? (n.htl ? n.tl : p.nearest) : p.temporal;
            rec.dx = (int8_t)clampv((int)base.dx + rdx, -128, 127);
            rec.dy = (int8_t)clampv((int)base.dy + rdy, -128, 127);
        }
        coded.push_back(rec);
    }
}

// compute_delta_bias_for_new / compute_bias_inner (encoder.rs:984-1048)
static void new_delta_bias(const std::vector<uint8_t>& modes, const std::vector<int8_t>& rdx, const std::vector<int8_t>& rdy,
                           int& bias_dx, int& bias_dy) {
    bias_dx = 0; bias_dy = 0;
    std::vector<std::pair<int8_t, int8_t>> f;
    for (size_t i = 0; i < modes.size(); i++) if (modes[i] == NEWMV) f.push_back({rdx[i], rdy[i]});
    if (f.empty()) return;
    int64_t bzx = -1, bzy = -1, bcx = INT64_MAX, bcy = INT64_MAX;
    for (int b = -24; b <= 24; b++) {
        int64_t zx = 0, cx = 0, zy = 0, cy = 0;
        for (auto& d : f) {
            int8_t ax = (int8_t)clampv((int)d.first - b, -128, 127);
            if (ax == 0) zx++;
            cx += std::abs((int)ax);
            int8_t ay = (int8_t)clampv((int)d.second - b, -128, 127);
            if (ay == 0) zy++;
            cy += std::abs((int)ay);
        }
        if (zx > bzx || (zx == bzx && cx < bcx)) { bzx = zx; bcx = cx; bias_dx = b; }
        if (zy > bzy || (zy == bzy && cy < bcy)) { bzy = zy; bcy = cy; bias_dy = b; }
    }
}

static int me_frame_encode(ro_encoder* e, const uint8_t* rgb, uint64_t ts, ro_frame_stats* st) {
    const ro_config& c = e->cfg;
    int SW = (int)e->hdr.sw, SH = (int)e->hdr.sh, bw = SW / 16, bh = SH / 16, nb = bw * bh;
    // keyframe decision  (encoder.rs:265-291)
    bool intra;
    if (c.keyframe_mode == 1) intra = true;
    else if (c.keyframe_mode == 2) intra = e->frame_count % 30 == 0;
    else intra = e->frame_count % c.keyframe_interval == 0;
    if (!e->have_prev) intra = true;
    std::vector<uint8_t> padded = pad_rgb24(rgb, (int)c.display_width, (int)c.display_height, SW, SH);
    rgb_to_yuv420(padded.data(), SW, SH, e->cur);
    memset(st, 0, sizeof(*st));
    st->frame_index = e->frame_count;
    st->blocks_total = (uint32_t)nb;
    std::vector<uint8_t> bytes;
    if (intra) {   // encoder.rs:308-387
        Yuv recon;
        std::vector<uint8_t> data = encode_intra(e->cur, c.intra_quality, recon, e->coeffs);
        deblock_yuv420(recon, deblock_level_from_qs(quant_step_from_quality(c.intra_quality)), true, nullptr);
        e->prev_recon = recon; e->have_prev = true;
        put_u64(bytes, ts); bytes.push_back(1); bytes.push_back(c.intra_quality); put_u32(bytes, (uint32_t)data.size());
        bytes.insert(bytes.end(), data.begin(), data.end());
        e->have_prev_mvs = false; e->prev_mvs.clear();
        e->mvctx.reset();
        st->frame_type = 1; st->residual_bytes = (uint32_t)data.size();
        e->pred = Yuv(); e->search_mvs.clear(); e->recon_mvs.clear(); e->sads.clear(); e->modes.clear();
        e->new_base.clear(); e->rdo_skip.clear(); e->opt_skip.assign(nb, 0); e->raw_dx.clear(); e->raw_dy.clear(); e->aq.clear();
    } else {   // encoder.rs:388-945
        const Yuv& prev = e->prev_recon;
        Lambda lam = lambda_for_quality(c.inter_quality, c.rdo_lambda_mult);
        const std::vector<MV>* pm = e->have_prev_mvs ? &e->prev_mvs : nullptr;
        hex_search(prev.y.data(), e->cur.y.data(), SW, SH, c.search_range, pm, c.me_zero_mv_threshold, c.me_predictor_threshold,
                   lam.lambda, c.me_threads, e->search_mvs, e->sads);
        build_predicted(prev, e->search_mvs, e->pred);
        // mode decision  (encoder.rs:447-733)
        std::vector<MV> coded;
        e->rdo_skip.assign(nb, 0);
        std::vector<uint8_t> flags(nb, 0);
        uint32_t mode_counts[7] = {0};
        decide_modes(e->search_mvs, pm, bw, bh, e->modes, e->new_base, e->raw_dx, e->raw_dy, coded);
        for (int i = 0; i < nb; i++) {
            const MV& mv = e->search_mvs[i];
            bool ia = mv.sx == 0 && mv.sy == 0;
            bool skipc = rdo_decide(lam, c.skip_threshold, e->sads[i], ia);
            e->rdo_skip[i] = skipc;
            mode_counts[e->modes[i]]++;
            uint8_t f = e->modes[i] == NEWMV ? (uint8_t)(mv.to_flags() & 0x0f) : 0;
            if (skipc) f |= 0x40;
            flags[i] = f;
        }
        Yuv recon;
        std::vector<uint8_t> res = encode_inter(e->cur, e->pred, e->rdo_skip, c.inter_quality, c.inter_dead_zone, recon, e->opt_skip, e->coeffs, e->aq);
        // bias  (encoder.rs:965-1048)
        int bias_dx = 0, bias_dy = 0;
        new_delta_bias(e->modes, e->raw_dx, e->raw_dy, bias_dx, bias_dy);
        std::vector<CodedBlock> blocks(nb);
        uint32_t nnew = 0, nskip = 0, nrdo = 0;
        for (int i = 0; i < nb; i++) {
            CodedBlock& b = blocks[i];
            b.mode = e->modes[i]; b.new_base = e->new_base[i];
            if (b.mode == NEWMV) {
                b.delta_x = (int8_t)clampv((int)e->raw_dx[i] - bias_dx, -128, 127);
                b.delta_y = (int8_t)clampv((int)e->raw_dy[i] - bias_dy, -128, 127);
                b.sub_x = e->search_mvs[i].sx; b.sub_y = e->search_mvs[i].sy;
                nnew++;
            }
            b.skip = e->opt_skip[i] != 0;
            nskip += b.skip; nrdo += e->rdo_skip[i];
        }
        std::vector<uint8_t> mvb = mv_encode_frame(e->mvctx, blocks, bw, bh);
        std::vector<uint8_t> fm(nb);   // filter mask (encoder.rs:831-845)
        for (int i = 0; i < nb; i++) { const MV& m = coded[i]; fm[i] = !e->opt_skip[i] || m.dx != 0 || m.dy != 0 || m.sx != 0 || m.sy != 0; }
        deblock_yuv420(recon, deblock_level_from_qs(quant_step_from_quality(c.inter_quality)), false, &fm);
        e->prev_recon = recon;
        e->prev_mvs = coded; e->have_prev_mvs = true;
        e->recon_mvs = coded;
        // PackedFrame inter  (format.rs:183-205)
        put_u64(bytes, ts); bytes.push_back(2); bytes.push_back(c.inter_quality);
        MV g = MV::make((int8_t)bias_dx, (int8_t)bias_dy, 0, 0, false);
        bytes.push_back((uint8_t)g.dx); bytes.push_back((uint8_t)g.dy); bytes.push_back(g.to_flags());
        put_u32(bytes, (uint32_t)mvb.size()); bytes.insert(bytes.end(), mvb.begin(), mvb.end());
        put_u32(bytes, (uint32_t)res.size()); bytes.insert(bytes.end(), res.begin(), res.end());
        st->frame_type = 2; st->mv_bytes = (uint32_t)mvb.size(); st->residual_bytes = (uint32_t)res.size();
        st->blocks_skipped = nskip; st->rdo_skips = nrdo; st->bias_dx = bias_dx; st->bias_dy = bias_dy; st->new_blocks = nnew;
        memcpy(st->mode_counts, mode_counts, sizeof(mode_counts));
    }
    st->total_bytes = (uint32_t)bytes.size();
    e->last_frame_bytes = bytes;
    e->out.insert(e->out.end(), bytes.begin(), bytes.end());
    e->frame_count++;
    e->hdr.frame_count = e->frame_count;
    return 0;
}

int ro_encoder_create(const ro_config* cfg, ro_encoder** out) {
    init_zero_prob(); init_basis();
    // EncoderConfig::validate  (config.rs:129-164)
    if (cfg->display_width == 0 || cfg->display_height == 0 || cfg->fps == 0 || cfg->search_range > 127) return 2;
    if (cfg->keyframe_mode == 0 && cfg->keyframe_interval == 0) return 2;
    ro_encoder* e = new ro_encoder();
    e->cfg = *cfg;
    e->hdr.dw = cfg->display_width; e->hdr.dh = cfg->display_height;
    e->hdr.sw = (cfg->display_width + 15) / 16 * 16; e->hdr.sh = (cfg->display_height + 15) / 16 * 16;
    e->hdr.fps = cfg->fps; e->hdr.frame_count = 0;
    e->out = header_bytes(e->hdr);
    *out = e;
    return 0;
}
int ro_encode_frame(ro_encoder* e, const uint8_t* rgb, size_t len, uint32_t w, uint32_t h, uint64_t ts, ro_frame_stats* st) {
    if (w != e->cfg.display_width || h != e->cfg.display_height) return 5;
    if (len != (size_t)w * h * 3) return 5;
    try { return me_frame_encode(e, rgb, ts, st); } catch (const std::exception&) { return 3; }
}
int ro_encoder_finish(ro_encoder* e) {   // encoder.rs:950-961: patch frame_count in the header
    std::vector<uint8_t> h = header_bytes(e->hdr);
    memcpy(e->out.data(), h.data(), h.size());
    return 0;
}
size_t ro_encoder_output(ro_encoder* e, const uint8_t** data) { *data = e->out.data(); return e->out.size(); }
void ro_encoder_destroy(ro_encoder* e) { delete e; }

// last-frame dumps
uint32_t ro_encoder_blocks(ro_encoder* e) { return (e->hdr.sw / 16) * (e->hdr.sh / 16); }
void ro_encoder_last_frame_bytes(ro_encoder* e, const uint8_t** d, size_t* n) { *d = e->last_frame_bytes.data(); *n = e->last_frame_bytes.size(); }
void ro_encoder_last_yuv(ro_encoder* e, int which, uint8_t* y, uint8_t* u, uint8_t* v) {
    const Yuv& f = which == 0 ? e->cur : (which == 1 ? e->pred : e->prev_recon);
    if (f.y.empty()) return;
    memcpy(y, f.y.data(), f.y.size()); memcpy(u, f.u.data(), f.u.size()); memcpy(v, f.v.data(), f.v.size());
}
// per-block record: search mv (dx,dy,sx,sy), recon mv (dx,dy,sx,sy), mode, new_base, raw_dx, raw_dy, rdo_skip, opt_skip
void ro_encoder_last_blocks(ro_encoder* e, int8_t* rec12, int64_t* sads) {
    size_t nb = e->search_mvs.size();
    for (size_t i = 0; i < nb; i++) {
        int8_t* r = rec12 + i * 12;
        const MV& s = e->search_mvs[i]; const MV& c = e->recon_mvs[i];
        r[0] = s.dx; r[1] = s.dy; r[2] = s.sx; r[3] = s.sy; r[4] = c.dx; r[5] = c.dy; r[6] = c.sx; r[7] = c.sy;
        r[8] = (int8_t)e->modes[i]; r[9] = (int8_t)e->new_base[i]; r[10] = e->raw_dx[i]; r[11] = e->raw_dy[i];
        sads[i] = e->sads[i];
    }
}
void ro_encoder_last_skip(ro_encoder* e, uint8_t* rdo, uint8_t* opt) {
    if (!e->rdo_skip.empty()) memcpy(rdo, e->rdo_skip.data(), e->rdo_skip.size());
    memcpy(opt, e->opt_skip.data(), e->opt_skip.size());
}
void ro_encoder_last_aq(ro_encoder* e, float* out) { if (!e->aq.empty()) memcpy(out, e->aq.data(), e->aq.size() * 4); }
void ro_encoder_last_coeffs(ro_encoder* e, int16_t* out) { memcpy(out, e->coeffs.data(), e->coeffs.size() * 2); }

// ---------------------------------------------------------------------------
// Decoder  (reitero_decode/src/decoder.rs)
// ---------------------------------------------------------------------------
struct ro_decoder {
    std::vector<uint8_t> data; size_t pos = 0;
    Header hdr;
    uint64_t current = 0;
    bool have_prev = false; Yuv prev;
    std::vector<MV> prev_mvs; bool have_prev_mvs = false;
    MvCtx mvctx;
    Yuv last;
    bool skip_residuals = false;   // Decoder::set_skip_residuals (decoder.rs:492-494)
};
typedef struct { uint64_t timestamp, frame_index; int32_t frame_type; uint32_t width, height; } ro_decoded_info;

int ro_decoder_create(const uint8_t* d, size_t n, ro_decoder** out) {   // decoder.rs:112-171
    init_zero_prob(); init_basis();
    if (n < 36) return 14;   // EndOfStream
    if (!(d[0] == 'R' && d[1] == 'I' && d[2] == 'V' && d[3] == 0)) return 11;
    if (get_u32(d + 4) != 5) return 11;
    ro_decoder* x = new ro_decoder();
    x->data.assign(d, d + n);
    x->hdr.dw = get_u32(d + 8); x->hdr.dh = get_u32(d + 12); x->hdr.sw = get_u32(d + 16); x->hdr.sh = get_u32(d + 20);
    x->hdr.fps = get_u32(d + 24); x->hdr.frame_count = get_u64(d + 28);
    x->pos = 36;
    *out = x;
    return 0;
}
static int decode_next(ro_decoder* x, uint8_t* rgb_out, ro_decoded_info* info) {
    const uint8_t* d = x->data.data(); size_t n = x->data.size();
    auto need = [&](size_t k) { if (x->pos + k > n) throw 14; };
    need(9);
    uint64_t ts = get_u64(d + x->pos); uint8_t ft = d[x->pos + 8]; x->pos += 9;
    if (ft != 1 && ft != 2) return 13;
    int SW = (int)x->hdr.sw, SH = (int)x->hdr.sh, bw = SW / 16, bh = SH / 16, nb = bw * bh;
    Yuv cur;
    if (ft == 1) {   // decoder.rs:261-281 / residual.rs:815-887
        need(5); uint8_t q = d[x->pos]; uint32_t sz = get_u32(d + x->pos + 1); x->pos += 5;
        need(sz); const uint8_t* payload = d + x->pos; x->pos += sz;
        float qs = quant_step_from_quality(q);
        uint16_t yt[256], ct[64]; build_intra_y_table(qs, yt); build_intra_uv_table(qs, ct);
        DctRansDec dec(true);
        if (sz) dec.consume(payload, sz);
        std::vector<int16_t> coeffs((size_t)nb * 384, 0);
        for (int b = 0; b < nb; b++) { int16_t* cb = &coeffs[(size_t)b * 384]; dec.decode_block(cb, cb + 256, cb + 320); }
        cur = Yuv(SW, SH); int cw = SW / 2;
        for (int b = 0; b < nb; b++) {
            int bx = b % bw, by = b / bw; const int16_t* cb = &coeffs[(size_t)b * 384];
            int32_t dd[256]; int16_t r[256];
            for (int i = 0; i < 256; i++) dd[i] = wmul((int32_t)cb[i], (int32_t)std::max<uint16_t>(yt[i], 1));
            idct16_2d(dd, r, 16);
            for (int j = 0; j < 16; j++) for (int i = 0; i < 16; i++) cur.y[(size_t)(by * 16 + j) * SW + bx * 16 + i] = (uint8_t)clampv(r[j * 16 + i] + 128, 0, 255);
            for (int pl = 0; pl < 2; pl++) {
                int16_t rc[64]; idct8_aan(cb + 256 + pl * 64, ct, rc, 8);
                std::vector<uint8_t>& dst = pl == 0 ? cur.u : cur.v;
                for (int j = 0; j < 8; j++) for (int i = 0; i < 8; i++) dst[(size_t)(by * 8 + j) * cw + bx * 8 + i] = (uint8_t)clampv(rc[j * 8 + i] + 128, 0, 255);
            }
        }
        if (!x->skip_residuals) deblock_yuv420(cur, deblock_level_from_qs(qs), true, nullptr);   // decoder.rs:269-274
        x->prev = cur; x->have_prev = true;
        x->have_prev_mvs = false; x->prev_mvs.clear();
        x->mvctx.reset();
    } else {   // decoder.rs:282-441
        need(8); uint8_t q = d[x->pos]; int8_t gdx = (int8_t)d[x->pos + 1], gdy = (int8_t)d[x->pos + 2];
        uint32_t mvs = get_u32(d + x->pos + 4); x->pos += 8;
        need(mvs); const uint8_t* mvp = d + x->pos; x->pos += mvs;
        need(4); uint32_t rs = get_u32(d + x->pos); x->pos += 4;
        need(rs); const uint8_t* rp = d + x->pos; x->pos += rs;
        if (!x->have_prev) return 13;
        std::vector<CodedBlock> blocks = mv_decode_frame(x->mvctx, mvp, mvs, bw, bh);
        const std::vector<MV>* pm = x->have_prev_mvs ? &x->prev_mvs : nullptr;
        std::vector<MV> mvsv; mvsv.reserve(nb);
        std::vector<uint8_t> skip(nb);
        for (int i = 0; i < nb; i++) {
            int bx = i % bw, by = i / bw; const CodedBlock& b = blocks[i];
            int ddx = b.mode == NEWMV ? clampv((int)b.delta_x + (int)gdx, -128, 127) : 0;
            int ddy = b.mode == NEWMV ? clampv((int)b.delta_y + (int)gdy, -128, 127) : 0;
            Preds p = derive_predictors(mvsv, pm, bw, bh, bx, by);
            Neigh nn = gather_neighbors(mvsv, pm, bw, bh, bx, by);
            Q4 base;
            switch (b.mode) {
                case ZERO: base = Q4{0, 0, 0, 0}; break;
                case NEAREST: base = p.nearest; break;
                case NEAR: base = p.near; break;
                case TOPRIGHT: base = nn.htr ? nn.tr : p.nearest; break;
                case TOPLEFT: base = nn.htl ? nn.tl : p.nearest; break;
                case TEMPORAL: base = p.temporal; break;
                default:
                    base = b.new_base == 0 ? p.nearest : b.new_base == 1 ? p.near : b.new_base == 2 ? (nn.htr ? nn.tr : p.nearest)
                           : b.new_base == 3 ? (nn.htl ? nn.tl : p.nearest) : b.new_base == 4 ? p.temporal : p.nearest;
            }
            int8_t dx = (int8_t)clampv((int)base.dx + ddx, -128, 127), dy = (int8_t)clampv((int)base.dy + ddy, -128, 127);
            MV m = b.mode == NEWMV ? MV::make(dx, dy, b.sub_x, b.sub_y, b.skip) : MV::make(dx, dy, base.sx, base.sy, b.skip);
            mvsv.push_back(m); skip[i] = b.skip;
        }
        Yuv pred; build_predicted(x->prev, mvsv, pred);
        if (x->skip_residuals) {   // the prediction alone, unfiltered (residual.rs:918-921, decoder.rs:407)
            x->last = pred;
            x->prev = pred;
            x->prev_mvs = mvsv; x->have_prev_mvs = true;
            std::vector<uint8_t> rgb((size_t)SW * SH * 3);
            yuv420_to_rgb(pred, rgb.data());
            crop_rgb24(rgb.data(), SW, (int)x->hdr.dw, (int)x->hdr.dh, rgb_out);
            info->timestamp = ts; info->frame_index = x->current; info->frame_type = ft; info->width = x->hdr.dw; info->height = x->hdr.dh;
            return 0;
        }
        // ResidualDecoder::decode_inter (residual.rs:889-1079)
        DctRansDec dec(false);
        if (rs) dec.consume(rp, rs);
        std::vector<int16_t> coeffs((size_t)nb * 384, 0);
        for (int i = 0; i < nb; i++) if (!skip[i]) { int16_t* cb = &coeffs[(size_t)i * 384]; dec.decode_block(cb, cb + 256, cb + 320); }
        float base_qs = quant_step_from_quality(q);
        std::vector<float> aq = compute_aq_steps(pred.y.data(), SW, SW, SH, base_qs);
        cur = pred;
        for (int i = 0; i < nb; i++) if (!skip[i]) recon_inter_block(&coeffs[(size_t)i * 384], aq[i], pred, cur, i % bw, i / bw);
        std::vector<uint8_t> fm(nb);
        for (int i = 0; i < nb; i++) { const MV& m = mvsv[i]; fm[i] = !m.skip || m.dx != 0 || m.dy != 0 || m.sx != 0 || m.sy != 0; }
        deblock_yuv420(cur, deblock_level_from_qs(base_qs), false, &fm);
        x->prev = cur;
        x->prev_mvs = mvsv; x->have_prev_mvs = true;
    }
    x->last = cur;
    std::vector<uint8_t> rgb((size_t)SW * SH * 3);
    yuv420_to_rgb(cur, rgb.data());
    crop_rgb24(rgb.data(), SW, (int)x->hdr.dw, (int)x->hdr.dh, rgb_out);
    info->timestamp = ts; info->frame_index = x->current; info->frame_type = ft; info->width = x->hdr.dw; info->height = x->hdr.dh;
    return 0;
}
int ro_decode_frame(ro_decoder* x, uint8_t* rgb_out, size_t cap, ro_decoded_info* info) {
    if (x->current >= x->hdr.frame_count) return 14;
    if (cap < (size_t)x->hdr.dw * x->hdr.dh * 3) return 12;
    try {
        int r = decode_next(x, rgb_out, info);
        if (r) return r;
    } catch (int code) { return code; } catch (const std::exception&) { return 12; }
    x->current++;
    return 0;
}
void ro_decoder_header(ro_decoder* x, uint32_t* out6) { out6[0] = x->hdr.dw; out6[1] = x->hdr.dh; out6[2] = x->hdr.sw; out6[3] = x->hdr.sh; out6[4] = x->hdr.fps; out6[5] = (uint32_t)x->hdr.frame_count; }
void ro_decoder_last_yuv(ro_decoder* x, uint8_t* y, uint8_t* u, uint8_t* v) {
    memcpy(y, x->last.y.data(), x->last.y.size()); memcpy(u, x->last.u.data(), x->last.u.size()); memcpy(v, x->last.v.data(), x->last.v.size());
}
void ro_decoder_set_skip_residuals(ro_decoder* x, int skip) { x->skip_residuals = skip != 0; }
void ro_decoder_destroy(ro_decoder* x) { delete x; }

// ---------------------------------------------------------------------------
// Stage entry points for unit tests
// ---------------------------------------------------------------------------
// encode_plane_8x8_aq + decode_plane_8x8_aq on one plane (reitero_dct KAT, tests/roundtrip.rs:366-400)
void ro_dct8_aq_plane(const int16_t* plane, int w, int h, const float* qs, float dz, int16_t* coeffs, int16_t* recon) {
    init_basis();
    int bw = w / 8, bh = h / 8;
    for (int by = 0; by < bh; by++) for (int bx = 0; bx < bw; bx++) {
        int b = by * bw + bx; int32_t c[64];
        fdct_2d<8>(plane + (size_t)by * 8 * w + bx * 8, w, c, B8);
        for (int i = 0; i < 64; i++) coeffs[b * 64 + i] = quant_aq(c[i], qs[b], dz);
        uint16_t q = f32_to_u16_sat(std::max(roundf(qs[b]), 1.0f)); uint16_t qt[64]; for (int i = 0; i < 64; i++) qt[i] = q;
        idct8_aan(coeffs + b * 64, qt, recon + (size_t)by * 8 * w + bx * 8, w);
    }
}
void ro_dct16_aq_plane(const int16_t* plane, int w, int h, const float* qs, float dz, int16_t* coeffs, int16_t* recon) {
    init_basis();
    int bw = w / 16, bh = h / 16;
    for (int by = 0; by < bh; by++) for (int bx = 0; bx < bw; bx++) {
        int b = by * bw + bx; int32_t c[256];
        fdct_2d<16>(plane + (size_t)by * 16 * w + bx * 16, w, c, B16);
        for (int i = 0; i < 256; i++) coeffs[b * 256 + i] = quant_aq(c[i], qs[b], dz);
        int32_t d[256]; dequant16_aq(coeffs + b * 256, qs[b], d);
        idct16_2d(d, recon + (size_t)by * 16 * w + bx * 16, w);
    }
}
void ro_basis(int32_t* b16, int32_t* b8) { init_basis(); memcpy(b16, B16, sizeof(B16)); memcpy(b8, B8, sizeof(B8)); }
void ro_intra_tables(float qs, uint16_t* yt, uint16_t* ct) { build_intra_y_table(qs, yt); build_intra_uv_table(qs, ct); }
float ro_quant_step(uint8_t q) { return quant_step_from_quality(q); }
int ro_deblock_level(float qs) { return deblock_level_from_qs(qs); }
void ro_deblock(uint8_t* y, uint8_t* u, uint8_t* v, int w, int h, int level, int intra, const uint8_t* mask) {
    Yuv f(w, h);
    memcpy(f.y.data(), y, f.y.size()); memcpy(f.u.data(), u, f.u.size()); memcpy(f.v.data(), v, f.v.size());
    std::vector<uint8_t> m; if (mask) m.assign(mask, mask + (w / 16) * (h / 16));
    deblock_yuv420(f, level, intra != 0, mask ? &m : nullptr);
    memcpy(y, f.y.data(), f.y.size()); memcpy(u, f.u.data(), f.u.size()); memcpy(v, f.v.data(), f.v.size());
}
void ro_rgb_to_yuv420(const uint8_t* rgb, int w, int h, uint8_t* y, uint8_t* u, uint8_t* v) {
    Yuv f; rgb_to_yuv420(rgb, w, h, f);
    memcpy(y, f.y.data(), f.y.size()); memcpy(u, f.u.data(), f.u.size()); memcpy(v, f.v.data(), f.v.size());
}
void ro_yuv420_to_rgb(const uint8_t* y, const uint8_t* u, const uint8_t* v, int w, int h, uint8_t* rgb) {
    Yuv f(w, h);
    memcpy(f.y.data(), y, f.y.size()); memcpy(f.u.data(), u, f.u.size()); memcpy(f.v.data(), v, f.v.size());
    yuv420_to_rgb(f, rgb);
}
// MC with explicit MVs (dx,dy,sx,sy per block)
void ro_build_predicted(const uint8_t* y, const uint8_t* u, const uint8_t* v, int w, int h, const int8_t* mv4, uint8_t* oy, uint8_t* ou, uint8_t* ov) {
    Yuv f(w, h);
    memcpy(f.y.data(), y, f.y.size()); memcpy(f.u.data(), u, f.u.size()); memcpy(f.v.data(), v, f.v.size());
    int nb = (w / 16) * (h / 16);
    std::vector<MV> mvs(nb);
    for (int i = 0; i < nb; i++) mvs[i] = MV::make(mv4[4 * i], mv4[4 * i + 1], mv4[4 * i + 2], mv4[4 * i + 3], false);
    Yuv o; build_predicted(f, mvs, o);
    memcpy(oy, o.y.data(), o.y.size()); memcpy(ou, o.u.data(), o.u.size()); memcpy(ov, o.v.data(), o.v.size());
}
// ME on two luma planes; prev_mv4 may be NULL (no temporal predictor)
void ro_hex_search(const uint8_t* prev, const uint8_t* cur, int w, int h, int sr, const int8_t* prev_mv4, int64_t zthr, int64_t pthr,
                   double lambda, int threads, int8_t* out_mv4, int64_t* out_sad) {
    init_basis();
    int nb = ((w + 15) / 16) * ((h + 15) / 16);
    std::vector<MV> pm;
    if (prev_mv4) { pm.resize(nb); for (int i = 0; i < nb; i++) pm[i] = MV::make(prev_mv4[4 * i], prev_mv4[4 * i + 1], prev_mv4[4 * i + 2], prev_mv4[4 * i + 3], false); }
    std::vector<MV> mvs; std::vector<int64_t> sads;
    hex_search(prev, cur, w, h, sr, prev_mv4 ? &pm : nullptr, zthr, pthr, lambda, threads, mvs, sads);
    for (int i = 0; i < nb; i++) { out_mv4[4 * i] = mvs[i].dx; out_mv4[4 * i + 1] = mvs[i].dy; out_mv4[4 * i + 2] = mvs[i].sx; out_mv4[4 * i + 3] = mvs[i].sy; out_sad[i] = sads[i]; }
}
int64_t ro_sad_int(const uint8_t* prev, const uint8_t* cur, int w, int h, int x0, int y0, int dx, int dy) { return sad_int(prev, cur, w, h, x0, y0, dx, dy); }
int64_t ro_sad_hp(const uint8_t* prev, const uint8_t* cur, int w, int h, int x0, int y0, int dxh, int dyh) { return sad_hp(prev, cur, w, h, x0, y0, dxh, dyh); }
int64_t ro_mv_cost(int dx, int dy, int nx, int ny, int rx, int ry, int mx, int my, double lambda) { return estimate_mv_cost(dx, dy, nx, ny, rx, ry, mx, my, lambda); }
double ro_lambda(uint8_t q, double mult) { return lambda_for_quality(q, mult).lambda; }
int ro_rdo_skip(uint8_t q, double mult, uint8_t thr, int64_t sad, int aligned) { return rdo_decide(lambda_for_quality(q, mult), thr, sad, aligned != 0); }
void ro_aq_steps(const uint8_t* y, int w, int h, float base, float* out) { std::vector<float> a = compute_aq_steps(y, w, w, h, base); memcpy(out, a.data(), a.size() * 4); }
uint16_t ro_binprob_observe(uint16_t packed, int bit) { init_zero_prob(); BinProb b; b.packed = packed; b.observe(bit != 0); return b.packed; }
uint8_t ro_binprob_p0(uint16_t packed) { init_zero_prob(); return g_zero_prob[packed]; }
// rANS round trip over a bit sequence with one adaptive context per `ctx_ids` value (<64), or bypass when ctx id < 0
size_t ro_rans_encode(const uint8_t* bits, const int8_t* ctx_ids, size_t n, uint8_t* out, size_t cap) {
    init_zero_prob();
    std::vector<uint8_t> o; RansEnc w(&o); BinProb ctx[64];
    for (size_t i = 0; i < n; i++) { if (ctx_ids[i] < 0) w.put_bypass(bits[i] != 0); else w.put(bits[i] != 0, ctx[ctx_ids[i]]); }
    w.finish();
    if (o.size() <= cap) memcpy(out, o.data(), o.size());
    return o.size();
}
int ro_rans_decode(const uint8_t* data, size_t len, const int8_t* ctx_ids, size_t n, uint8_t* bits) {
    init_zero_prob();
    try {
        RansDec r(data, len); BinProb ctx[64];
        for (size_t i = 0; i < n; i++) bits[i] = ctx_ids[i] < 0 ? r.get_bypass() : r.get(ctx[ctx_ids[i]]);
    } catch (const std::exception&) { return 1; }
    return 0;
}


// ---- seams for independent pins (tests only) ----
// Mode decision and NEW-delta bias of one frame: smv4 = nb search MVs (dx, dy, sx, sy), prev4 =
// nb previous reconstructed MVs or null; out12 per block as ro_encoder_last_blocks' records
// (search MV, reconstructed MV, mode, NEW base, raw NEW delta), bias2 = (bias_dx, bias_dy).
void ro_mode_decision(const int8_t* smv4, const int8_t* prev4, int bw, int bh, int8_t* out12, int32_t* bias2) {
    const int nb = bw * bh;
    std::vector<MV> smv(nb), prev;
    for (int i = 0; i < nb; i++) smv[i] = MV::make(smv4[4 * i], smv4[4 * i + 1], smv4[4 * i + 2], smv4[4 * i + 3], false);
    if (prev4) { prev.resize(nb); for (int i = 0; i < nb; i++) prev[i] = MV::make(prev4[4 * i], prev4[4 * i + 1], prev4[4 * i + 2], prev4[4 * i + 3], false); }
    std::vector<uint8_t> modes, nbase; std::vector<int8_t> rdx, rdy; std::vector<MV> coded;
    decide_modes(smv, prev4 ? &prev : nullptr, bw, bh, modes, nbase, rdx, rdy, coded);
    for (int i = 0; i < nb; i++) {
        int8_t* r = out12 + 12 * i;
        r[0] = smv[i].dx; r[1] = smv[i].dy; r[2] = smv[i].sx; r[3] = smv[i].sy;
        r[4] = coded[i].dx; r[5] = coded[i].dy; r[6] = coded[i].sx; r[7] = coded[i].sy;
        r[8] = (int8_t)modes[i]; r[9] = (int8_t)nbase[i]; r[10] = rdx[i]; r[11] = rdy[i];
    }
    int bx = 0, by = 0;
    new_delta_bias(modes, rdx, rdy, bx, by);
    bias2[0] = bx; bias2[1] = by;
}
// The residual coder (DctRansEncoder) over nb blocks of 384 coefficients (Y 256, U 64, V 64),
// the ones with coded[b] != 0, with or without intra DC prediction; bytes into out (cap), size returned.
size_t ro_encode_residual_blocks(const int16_t* coeffs, const uint8_t* coded, int nb, int dc_pred, uint8_t* out, size_t cap) {
    init_zero_prob();
    bool any = false;
    for (int b = 0; b < nb; b++) any |= coded[b] != 0;
    if (!any) return 0;
    DctRansEnc enc(dc_pred != 0);
    for (int b = 0; b < nb; b++) if (coded[b]) { const int16_t* q = coeffs + (size_t)b * 384; enc.encode_block(q, q + 256, q + 320); }
    std::vector<uint8_t> bytes = enc.finish();
    if (bytes.size() <= cap) memcpy(out, bytes.data(), bytes.size());
    return bytes.size();
}
// The MV coder over nframes consecutive P-frames of one stream (contexts and candidate history
// carried from frame to frame): blocks7 = nframes * nb blocks of (mode, new_base, delta_x,
// delta_y, sub_x, sub_y, skip); each frame's bytes appended to out, sizes[f] its length.
size_t ro_mv_encode_sequence(const int8_t* blocks7, int nframes, int bw, int bh, uint8_t* out, size_t cap, uint32_t* sizes) {
    init_zero_prob();
    const int nb = bw * bh;
    MvCtx ctx;
    size_t total = 0;
    for (int f = 0; f < nframes; f++) {
        std::vector<CodedBlock> blocks(nb);
        for (int i = 0; i < nb; i++) {
            const int8_t* q = blocks7 + ((size_t)f * nb + i) * 7;
            CodedBlock& b = blocks[i];
            b.mode = (uint8_t)q[0]; b.new_base = (uint8_t)q[1]; b.delta_x = q[2]; b.delta_y = q[3]; b.sub_x = q[4]; b.sub_y = q[5]; b.skip = q[6] != 0;
        }
        std::vector<uint8_t> bytes = mv_encode_frame(ctx, blocks, bw, bh);
        if (total + bytes.size() <= cap) memcpy(out + total, bytes.data(), bytes.size());
        total += bytes.size();
        sizes[f] = (uint32_t)bytes.size();
    }
    return total;
}
// IDCT16 of one dequantised block (decode_plane_16x16_aq after its dequantisation, dct16.rs:345-412)
void ro_idct16(const int32_t* d256, int16_t* out256) { init_basis(); idct16_2d(d256, out256, 16); }
// IDCT8 of one block with its quantisation table (idct_block_8x8_residual_simd, fast_dct8.rs:123-174)
void ro_idct8(const int16_t* coef64, const uint16_t* q64, int16_t* out64) { idct8_aan(coef64, q64, out64, 8); }
}  // extern "C"
