// Host-side entropy coding of the RIV bitstream (see entropy.hpp). Byte-exact with the
// reference coders; citations are to the reference tree under reitero_video/.
#include "entropy.hpp"
#include <algorithm>
#include <cstring>
#include <cstdlib>

namespace rivh {

namespace {
// P(0) table of the packed counter pair (reitero_video_common/src/rans.rs:30-44):
// zeros*256/(zeros+ones) truncated to u8; entries that truncate to 0 read as 1.
struct P0Table {
    uint8_t t[65536];
    P0Table() {
        t[0] = 1;
        for (uint32_t i = 1; i < 65536; i++) {
            uint32_t z = i >> 8, o = i & 0xff;
            uint8_t v = (uint8_t)((z << 8) / (z + o));
            t[i] = v ? v : 1;
        }
    }
};
const P0Table& p0table() { static P0Table tab; return tab; }
inline uint16_t swap_bytes(uint16_t v) { return (uint16_t)((v << 8) | (v >> 8)); }
// x / d for x < 2^32 and 1 <= d <= 256 without a divide: with M = floor(2^40 / d) + 1,
// x*M / 2^40 = x/d + e where 0 < e < x / 2^40 < 1/256 <= 1/d, and frac(x/d) <= 1 - 1/d, so the
// floor is exact.
struct RcpTable {
    uint64_t m[257];
    RcpTable() { m[0] = 0; for (int d = 1; d <= 256; d++) m[d] = (((uint64_t)1) << 40) / (uint64_t)d + 1; }
};
const RcpTable g_rcp;
inline uint32_t div_small(uint32_t x, uint32_t d) { return (uint32_t)(((unsigned __int128)x * g_rcp.m[d]) >> 40); }
constexpr uint32_t kLow = 1u << 16;   // normalisation interval [2^16, 2^24)
}  // namespace

uint8_t Ctx::p0() const { return p0table().t[c]; }

// BinProb::observe (rans.rs:73-84): the counter of the observed symbol sits in the high
// byte; on overflow both counters are halved (rounding up), the observed one re-based.
void Ctx::update(bool bit) {
    uint16_t v = bit ? swap_bytes(c) : c;
    if (v >= 0xff00) {
        uint16_t keep = v == 0xff01 ? 0xff00 : 0x8100;
        v = (uint16_t)((uint16_t)(v + 0x101) >> 1) | keep;
    } else {
        v = (uint16_t)(v + 0x100);
    }
    c = bit ? swap_bytes(v) : v;
}

RansWriter::RansWriter(std::vector<uint8_t>& out) : out_(out), sym_(kBatch), bits_(kBatch) {}

// Encode the buffered block last-symbol-first with two alternating states; emitted u16
// words are written in reverse emission order after both final states (state 1 first),
// which is the byte order of the reference's front-pushed deque (rans.rs:231-269).
void RansWriter::drain() {
    uint32_t st[2] = {kLow, kLow};
    std::vector<uint16_t> emitted;
    emitted.reserve(n_ / 4 + 8);
    auto code = [&](uint32_t& x, int k) {
        uint32_t p = sym_[k] ? sym_[k] : 1u;
        uint32_t start = bits_[k] ? p : 0u, freq = bits_[k] ? 256u - p : p;
        if (x >= (freq << 24)) { emitted.push_back((uint16_t)x); x >>= 16; }
        const uint32_t q = div_small(x, freq);
        x = (q << 8) + (x - q * freq) + start;
    };
    int k = n_ - 1;
    int cur = 0;
    if (n_ & 1) { code(st[0], k--); cur = 1; }   // odd block: first coded symbol alone, then roles swap
    while (k >= 1) { code(st[cur], k); code(st[cur ^ 1], k - 1); k -= 2; }
    // after an odd start the states are swapped: st[cur] plays r0, st[cur^1] plays r1
    uint32_t r0 = st[cur], r1 = st[cur ^ 1];
    auto put16 = [&](uint16_t w) { out_.push_back((uint8_t)w); out_.push_back((uint8_t)(w >> 8)); };
    put16((uint16_t)r1); put16((uint16_t)(r1 >> 16));
    put16((uint16_t)r0); put16((uint16_t)(r0 >> 16));
    for (size_t i = emitted.size(); i-- > 0;) put16(emitted[i]);
    n_ = 0;
}
void RansWriter::flush() { drain(); }

RansReader::RansReader(const uint8_t* data, size_t len) : d_(data), len_(len) { a_ = state(); b_ = state(); }
uint16_t RansReader::word() {
    if (pos_ + 2 > len_) throw CorruptStream("rANS stream truncated");
    uint16_t w = (uint16_t)(d_[pos_] | (d_[pos_ + 1] << 8));
    pos_ += 2;
    return w;
}
uint32_t RansReader::state() { uint32_t lo = word(); return lo | ((uint32_t)word() << 16); }
void RansReader::step() {   // reload both states every block of 16386 symbols (rans.rs:318-326)
    if (count_ == 16386) { count_ = 0; a_ = state(); b_ = state(); }
    count_++;
}
bool RansReader::bit(Ctx& c) {
    step();
    uint32_t x = a_; a_ = b_;
    uint32_t p = c.p0();
    bool bit = (x & 255) >= p;
    c.update(bit);
    uint32_t start = bit ? p : 0, freq = bit ? 256 - p : p;
    x = freq * (x >> 8) + (x & 255) - start;
    if (x < kLow) x = (x << 16) | word();
    b_ = x;
    return bit;
}
bool RansReader::bypass() {
    step();
    uint32_t x = a_; a_ = b_;
    uint32_t start = x & 0x80;
    x = 128 * (x >> 8) + (x & 255) - start;
    if (x < kLow) x = (x << 16) | word();
    b_ = x;
    return start != 0;
}

// ----------------------------------------------------------------------------------------
// Residual coefficients (reitero_residual/src/rans.rs)
// ----------------------------------------------------------------------------------------
namespace {
// value class: 0, 1, then k for [2^(k-1), 2^k - 1] (rans.rs:43-50)
inline int value_class(int v) { return v <= 1 ? v : 32 - __builtin_clz((unsigned)v); }
// truncated-unary class over `n` contexts, then the offset MSB-first as bypass bits
// (encode_eob / encode_magnitude, rans.rs:69-96, :128-155)
void put_class_value(RansWriter& w, int v, Ctx* ctx, int n) {
    int k = value_class(v);
    if (k > n - 1) throw CorruptStream("coefficient magnitude outside the class table");
    for (int i = 0; i < k; i++) w.bit(true, ctx[i]);
    if (k < n - 1) w.bit(false, ctx[k]);
    for (int b = k - 2; b >= 0; b--) w.bypass(((v - (1 << (k - 1))) >> b) & 1);
}
int get_class_value(RansReader& r, Ctx* ctx, int n) {
    int k = 0;
    while (k < n - 1 && r.bit(ctx[k])) k++;
    if (k <= 1) return k;
    int v = 1;
    for (int i = 0; i < k - 1; i++) v = (v << 1) | (int)r.bypass();
    return v;
}
inline int band_y(int i) { return i == 0 ? 0 : i < 16 ? 1 : i < 64 ? 2 : 3; }
inline int band_c(int i) { return i == 0 ? 0 : i < 8 ? 1 : 2; }

void put_plane(RansWriter& w, ResidualCtx& c, const int16_t* q, int n, int plane, bool dc_pred) {
    const bool luma = plane == 0;
    int16_t first = q[0];
    int last = 0;
    if (dc_pred) {   // intra DC prediction: DC delta to the previous block of the plane (rans.rs:250-266)
        first = (int16_t)(q[0] - c.prev_dc[plane]);
        c.prev_dc[plane] = q[0];
        for (int i = n - 1; i >= 1; i--) if (q[i]) { last = i; break; }
    } else {
        for (int i = n - 1; i >= 0; i--) if (q[i]) { last = i; break; }
    }
    put_class_value(w, last, luma ? c.eob_y : c.eob_c, luma ? 9 : 7);
    Ctx* sig = luma ? c.sig_y : c.sig_c;
    for (int i = 0; i <= last; i++) {
        int v = i == 0 ? first : q[i];
        w.bit(v != 0, sig[i]);
        if (v) {
            w.bypass(v < 0);
            put_class_value(w, std::abs(v) - 1, luma ? c.mag_y[band_y(i)] : c.mag_c[band_c(i)], 13);
        }
    }
}
void get_plane(RansReader& r, ResidualCtx& c, int16_t* q, int n, int plane, bool dc_pred) {
    const bool luma = plane == 0;
    int last = get_class_value(r, luma ? c.eob_y : c.eob_c, luma ? 9 : 7);
    if (last >= n) throw CorruptStream("EOB beyond block size");
    Ctx* sig = luma ? c.sig_y : c.sig_c;
    for (int i = 0; i <= last; i++) {
        if (!r.bit(sig[i])) continue;
        bool neg = r.bypass();
        int a = get_class_value(r, luma ? c.mag_y[band_y(i)] : c.mag_c[band_c(i)], 13) + 1;
        q[i] = (int16_t)(neg ? -a : a);
    }
    if (dc_pred) { q[0] = (int16_t)(q[0] + c.prev_dc[plane]); c.prev_dc[plane] = q[0]; }
}
}  // namespace

std::vector<uint8_t> encode_residuals(const int16_t* coeffs, const uint8_t* coded, int nb, bool dc_pred) {
    std::vector<uint8_t> out;
    bool any = false;
    for (int b = 0; b < nb && !any; b++) any = coded[b] != 0;
    if (!any) return out;   // no coder instantiated when every block is skipped (residual.rs:670-674, :792-798)
    ResidualCtx ctx;
    RansWriter w(out);
    for (int b = 0; b < nb; b++) {
        if (!coded[b]) continue;
        const int16_t* q = coeffs + (size_t)b * 384;
        put_plane(w, ctx, q, 256, 0, dc_pred);
        put_plane(w, ctx, q + 256, 64, 1, dc_pred);
        put_plane(w, ctx, q + 320, 64, 2, dc_pred);
    }
    w.flush();
    return out;
}

void unpack_coeffs(const uint8_t* slot, int16_t* q) {
    std::memset(q, 0, 384 * sizeof(int16_t));
    const int16_t* v = (const int16_t*)(slot + 48);
    for (int k = 0; k < 6; k++) {
        uint64_t m;
        std::memcpy(&m, slot + 8 * k, 8);
        while (m) {
            const int bit = __builtin_ctzll(m);
            q[64 * k + bit] = *v++;
            m &= m - 1;
        }
    }
}

std::vector<uint8_t> encode_residuals_packed(const uint8_t* pack, size_t stride, const uint8_t* coded, int nb, bool dc_pred) {
    std::vector<uint8_t> out;
    bool any = false;
    for (int b = 0; b < nb && !any; b++) any = coded[b] != 0;
    if (!any) return out;
    ResidualCtx ctx;
    RansWriter w(out);
    int16_t q[384];
    for (int b = 0; b < nb; b++) {
        if (!coded[b]) continue;
        unpack_coeffs(pack + (size_t)b * stride, q);
        put_plane(w, ctx, q, 256, 0, dc_pred);
        put_plane(w, ctx, q + 256, 64, 1, dc_pred);
        put_plane(w, ctx, q + 320, 64, 2, dc_pred);
    }
    w.flush();
    return out;
}

void decode_residuals(const uint8_t* data, size_t len, const uint8_t* coded, int nb, bool dc_pred, int16_t* coeffs) {
    bool any = false;
    for (int b = 0; b < nb && !any; b++) any = coded[b] != 0;
    if (!any) return;
    if (len == 0) throw CorruptStream("residual payload missing");
    ResidualCtx ctx;
    RansReader r(data, len);
    for (int b = 0; b < nb; b++) {
        if (!coded[b]) continue;
        int16_t* q = coeffs + (size_t)b * 384;
        get_plane(r, ctx, q, 256, 0, dc_pred);
        get_plane(r, ctx, q + 256, 64, 1, dc_pred);
        get_plane(r, ctx, q + 320, 64, 2, dc_pred);
    }
}

// ----------------------------------------------------------------------------------------
// Motion vectors (reitero_residual/src/mv_rans.rs, mv_rans/encoder.rs, mv_rans/decoder.rs)
// ----------------------------------------------------------------------------------------
int mv_class(int m) { return m == 0 ? 0 : std::min(32 - __builtin_clz((unsigned)m), 10); }

namespace {
enum { MZERO = 0, MNEAREST = 1, MNEAR = 2, MTR = 3, MTL = 4, MNEW = 5, MTEMP = 6 };
// mv_mode_context (mv_predictor.rs:275-308): left + top, Zero = 2, New = 0, others 1, capped at 3
int mode_ctx(const std::vector<uint8_t>& modes, int bw, int i) {
    int bx = i % bw, s = 0;
    auto score = [](uint8_t m) { return m == MZERO ? 2 : m == MNEW ? 0 : 1; };
    if (bx > 0) s += score(modes[i - 1]);
    if (i >= bw) s += score(modes[i - bw]);
    return std::min(s, 3);
}
inline int subpel_rank(int8_t s) { return s == 0 ? 0 : s == 1 ? 1 : 2; }   // derived Ord of Subpel
inline int8_t subpel_of_rank(int r) { return r == 0 ? 0 : r == 1 ? 1 : -1; }
}  // namespace

void MvCoder::reset() { *this = MvCoder(); }

int MvCoder::lookup(int8_t dx, int8_t dy, int8_t sx, int8_t sy) const {
    std::array<int, 4> k = {dx, dy, subpel_rank(sx), subpel_rank(sy)};
    for (size_t i = 0; i < cand_.size(); i++) if (cand_[i] == k) return (int)i;
    return -1;
}

// update_candidates (mv_rans/encoder.rs:370-412, decoder.rs:283-326): decay every count by
// 7/8, count this frame's NEW (delta, sub-pel) tuples, drop zero counts; below 100 NEW blocks
// of history the list is empty, else the top 16 by count (ties by key) above total/100.
void MvCoder::refresh_history(const std::vector<MvBlock>& blocks, bool from_decoder) {
    for (auto& e : counts_) e.second = e.second * 7 / 8;
    total_ = total_ * 7 / 8;
    for (const MvBlock& b : blocks) {
        if (b.mode != MNEW) continue;
        if (b.sx == -1 || b.sy == -1) {
            if (from_decoder) continue;
            throw CorruptStream("-0.5 sub-pel in a NEW vector");
        }
        std::array<int, 4> k = {b.dx, b.dy, subpel_rank(b.sx), subpel_rank(b.sy)};
        bool found = false;
        for (auto& e : counts_) if (e.first == k) { e.second++; found = true; break; }
        if (!found) counts_.push_back({k, 1});
        total_++;
    }
    counts_.erase(std::remove_if(counts_.begin(), counts_.end(), [](const std::pair<std::array<int, 4>, uint32_t>& e) { return e.second == 0; }),
                  counts_.end());
    cand_.clear();
    if (total_ < 100) return;
    std::vector<std::pair<std::array<int, 4>, uint32_t>> order = counts_;
    std::sort(order.begin(), order.end(), [](const std::pair<std::array<int, 4>, uint32_t>& a, const std::pair<std::array<int, 4>, uint32_t>& b) {
        return a.second != b.second ? a.second > b.second : a.first < b.first;
    });
    const uint32_t floor_count = total_ / 100;
    for (auto& e : order) {
        if (e.second <= floor_count) continue;
        cand_.push_back(e.first);
        if (cand_.size() == 16) break;
    }
}

std::vector<uint8_t> MvCoder::encode(const std::vector<MvBlock>& blocks, int bw, int bh) {
    (void)bh;
    std::vector<uint8_t> out;
    RansWriter w(out);
    std::vector<uint8_t> modes;
    std::vector<uint8_t> hits;
    modes.reserve(blocks.size()); hits.reserve(blocks.size());
    auto component = [&](int8_t d, int axis) {   // encode_component (mv_rans/encoder.rs:303-331)
        int mag = std::abs((int)d), k = mv_class(mag);
        int lo = 0, hi = 10, depth = 0;
        while (lo < hi) { int mid = (lo + hi) / 2; bool up = k > mid; w.bit(up, cls_[axis][depth++]); if (up) lo = mid + 1; else hi = mid; }
        if (k == 0) return;
        w.bit(d < 0, sign_[axis]);
        int off = mag - (1 << (k - 1));
        for (int b = 0; b < k - 1; b++) w.bit((off >> b) & 1, mag_[axis][b]);
    };
    for (size_t i = 0; i < blocks.size(); i++) {
        const MvBlock& b = blocks[i];
        const int bx = (int)(i % bw);
        w.bit(b.skip, skip_);
        const int ctx = mode_ctx(modes, bw, (int)i);
        // mode tree: New | Zero | Near | Temporal | TopRight | TopLeft | Nearest (mv_rans/encoder.rs:257-301)
        static const uint8_t order[6] = {MNEW, 0xff, MNEAR, MTEMP, MTR, MTL};
        for (int lvl = 0; lvl < 6; lvl++) {
            bool stop;
            if (lvl == 1) { bool nz = b.mode != MZERO; w.bit(nz, mode_[1][ctx]); stop = !nz; }
            else { bool is = b.mode == order[lvl]; w.bit(is, mode_[lvl][ctx]); stop = is; }
            if (stop) break;
        }
        bool hit = false;
        if (b.mode == MNEW) {
            for (int lvl = 0; lvl < 4; lvl++) {   // base selector (nearest, near, top-right, top-left | temporal)
                bool is = b.new_base == lvl;
                w.bit(is, base_[lvl][ctx]);
                if (is) break;
            }
            bool frac_done = false;
            if (!cand_.empty()) {
                int ci = lookup(b.dx, b.dy, b.sx, b.sy);
                hit = ci >= 0;
                int hctx = (bx > 0 ? hits[i - 1] : 0) + (i >= (size_t)bw ? hits[i - bw] : 0);
                w.bit(hit, hit_[hctx]);
                if (hit) {
                    int lo = 0, hi = 16, depth = 0;
                    while (hi - lo > 1) { int mid = (lo + hi) / 2; bool up = ci >= mid; w.bit(up, idx_[depth++]); if (up) lo = mid; else hi = mid; }
                    frac_done = true;
                } else {
                    component(b.dx, 0); component(b.dy, 1);
                }
            } else {
                component(b.dx, 0); component(b.dy, 1);
            }
            if (!frac_done) {
                if (b.sx == -1 || b.sy == -1) throw CorruptStream("-0.5 sub-pel in a NEW vector");
                w.bit(b.sx != 0, frac_[0]);
                w.bit(b.sy != 0, frac_[1]);
            }
        }
        hits.push_back(hit);
        modes.push_back(b.mode);
    }
    w.flush();
    refresh_history(blocks, false);
    return out;
}

std::vector<MvBlock> MvCoder::decode(const uint8_t* data, size_t len, int bw, int bh) {
    RansReader r(data, len);
    const int nb = bw * bh;
    std::vector<MvBlock> res(nb);
    std::vector<uint8_t> modes, hits;
    modes.reserve(nb); hits.reserve(nb);
    auto component = [&](int axis) -> int8_t {   // decode_component (mv_rans/decoder.rs:375-401)
        int lo = 0, hi = 10, depth = 0;
        while (lo < hi) { int mid = (lo + hi) / 2; if (r.bit(cls_[axis][depth++])) lo = mid + 1; else hi = mid; }
        int k = lo;
        if (k == 0) return 0;
        bool neg = r.bit(sign_[axis]);
        int off = 0;
        for (int b = 0; b < k - 1; b++) if (r.bit(mag_[axis][b])) off |= 1 << b;
        int v = (1 << (k - 1)) + off;
        v = neg ? -v : v;
        return (int8_t)std::max(-128, std::min(127, v));
    };
    for (int i = 0; i < nb; i++) {
        MvBlock b{};
        const int bx = i % bw;
        b.skip = r.bit(skip_);
        const int ctx = mode_ctx(modes, bw, i);
        if (r.bit(mode_[0][ctx])) b.mode = MNEW;
        else if (!r.bit(mode_[1][ctx])) b.mode = MZERO;
        else if (r.bit(mode_[2][ctx])) b.mode = MNEAR;
        else if (r.bit(mode_[3][ctx])) b.mode = MTEMP;
        else if (r.bit(mode_[4][ctx])) b.mode = MTR;
        else if (r.bit(mode_[5][ctx])) b.mode = MTL;
        else b.mode = MNEAREST;
        bool hit = false;
        if (b.mode == MNEW) {
            b.new_base = 4;
            for (int lvl = 0; lvl < 4; lvl++) if (r.bit(base_[lvl][ctx])) { b.new_base = (uint8_t)lvl; break; }
            bool frac_done = false;
            if (!cand_.empty()) {
                int hctx = (bx > 0 ? hits[i - 1] : 0) + (i >= bw ? hits[i - bw] : 0);
                hit = r.bit(hit_[hctx]);
                if (hit) {
                    int lo = 0, hi = 16, depth = 0;
                    while (hi - lo > 1) { int mid = (lo + hi) / 2; if (r.bit(idx_[depth++])) lo = mid; else hi = mid; }
                    if (lo < (int)cand_.size()) {
                        const auto& k = cand_[lo];
                        b.dx = (int8_t)k[0]; b.dy = (int8_t)k[1]; b.sx = subpel_of_rank(k[2]); b.sy = subpel_of_rank(k[3]);
                        frac_done = true;
                    }
                } else {
                    b.dx = component(0); b.dy = component(1);
                }
            } else {
                b.dx = component(0); b.dy = component(1);
            }
            if (!frac_done) { b.sx = r.bit(frac_[0]) ? 1 : 0; b.sy = r.bit(frac_[1]) ? 1 : 0; }
        }
        hits.push_back(hit);
        modes.push_back(b.mode);
        res[i] = b;
    }
    refresh_history(res, true);
    return res;
}

}  // namespace rivh
