// Host-side entropy coding of the RIV bitstream (see entropy.hpp). Byte-exact with the
// reference coders; citations are to the reference tree under reitero_video/.
#include "entropy.hpp"
#include "riv_common.hpp"
#include <algorithm>
#include <cstring>
#include <cstdlib>
#include <immintrin.h>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>

namespace rivh {

namespace {
struct P0Rcp {
    uint32_t m[512];
    constexpr P0Rcp() : m() {
        for (int d = 1; d < 512; d++) m[d] = (uint32_t)((((uint64_t)1) << 32) / (uint64_t)d + 1);
    }
};
constexpr P0Rcp kRcpInit{};
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
const bool g_avx2_rans = __builtin_cpu_supports("avx2");
}  // namespace

const uint32_t kP0Rcp[512] = {
#define R8(i) kRcpInit.m[i], kRcpInit.m[i + 1], kRcpInit.m[i + 2], kRcpInit.m[i + 3], kRcpInit.m[i + 4], kRcpInit.m[i + 5], kRcpInit.m[i + 6], kRcpInit.m[i + 7]
#define R64(i) R8(i), R8(i + 8), R8(i + 16), R8(i + 24), R8(i + 32), R8(i + 40), R8(i + 48), R8(i + 56)
    R64(0), R64(64), R64(128), R64(192), R64(256), R64(320), R64(384), R64(448)
#undef R64
#undef R8
};

namespace {
struct WriterBuffers { std::vector<uint16_t> buf, tmp[4], tmp8[16]; };
thread_local std::vector<WriterBuffers> tl_pool;
}  // namespace

RansWriter::RansWriter(std::vector<uint8_t>& out) : out_(out) {
    if (!tl_pool.empty()) {
        WriterBuffers& b = tl_pool.back();
        buf_.swap(b.buf);
        for (int j = 0; j < kLanes; j++) tmp_[j].swap(b.tmp[j]);
        for (int j = 0; j < 16; j++) tmp8_[j].swap(b.tmp8[j]);
        tl_pool.pop_back();
        sym_ = buf_.data(); cap_ = buf_.size();
    }
}
RansWriter::~RansWriter() {
    if (tl_pool.size() < 4) {
        tl_pool.emplace_back();
        WriterBuffers& b = tl_pool.back();
        b.buf.swap(buf_);
        for (int j = 0; j < kLanes; j++) b.tmp[j].swap(tmp_[j]);
        for (int j = 0; j < 16; j++) b.tmp8[j].swap(tmp8_[j]);
    }
}

void RansWriter::prewarm(size_t symbols) {
    if (!tl_pool.empty()) return;
    tl_pool.emplace_back();
    WriterBuffers& b = tl_pool.back();
    size_t cap = 1u << 16;
    while (cap < symbols) cap *= 2;
    b.buf.assign(cap, 0);
    for (auto& t : b.tmp) t.assign((size_t)kBatch + 8, 0);
    for (auto& t : b.tmp8) t.assign((size_t)kBatch + 8, 0);
}

void RansWriter::grow() {
    buf_.resize(std::max<size_t>(cap_ * 2, 1u << 16));
    sym_ = buf_.data();
    cap_ = buf_.size();
}

// Codes blocks last-symbol-first with two alternating states; a block's emitted u16 words go
// out in reverse emission order after both final states (state 1 first), which is the byte order
// of the reference's front-pushed deque (rans.rs:231-269). Renormalisation is branch-free: the
// low word is always stored below the write pointer, which moves only when the word is emitted.
// Blocks are independent, so `nblk` of them (at most kLanes) are coded in one interleaved loop to
// give the CPU independent state chains to overlap; block j holds n[j] symbols at sym[j].
namespace {
struct BlockCoder {
    const uint16_t* sym;
    uint16_t* wp;       // next emitted word goes to wp[-1]
    uint32_t A, B;      // A codes the higher symbol of each pair, B the lower; they end as r0, r1
    int k;              // next symbol to code
    void start(const uint16_t* s, int n, uint16_t* end) {
        sym = s; wp = end; A = kLow; B = kLow; k = n - 1;
        if (n & 1) { code(B, k); k--; }   // odd block: first symbol alone, then the roles swap
    }
    inline void code(uint32_t& x, int i) {
        const uint32_t s = sym[i], freq = s & 0xffu, start = s >> 8;
        const uint32_t e = x >= (freq << 24);
        wp[-1] = (uint16_t)x;
        wp -= e;
        x = e ? x >> 16 : x;
        const uint32_t q = div_small(x, freq);
        x += q * (256u - freq) + start;
    }
    inline void pair() { code(A, k); code(B, k - 1); k -= 2; }
};
}  // namespace

// A block's output: state 1 then state 0 as two u16 words each, then the emitted words [wp, end).
void RansWriter::emit_block(uint32_t r0, uint32_t r1, const uint16_t* wp, const uint16_t* end) {
    const size_t ne = (size_t)(end - wp), o = out_.size();
    out_.resize(o + 8 + 2 * ne);
    uint8_t* d = out_.data() + o;
    const uint16_t head[4] = {(uint16_t)r1, (uint16_t)(r1 >> 16), (uint16_t)r0, (uint16_t)(r0 >> 16)};
    for (int i = 0; i < 4; i++) { d[2 * i] = (uint8_t)head[i]; d[2 * i + 1] = (uint8_t)(head[i] >> 8); }
    for (size_t i = 0; i < ne; i++) { d[8 + 2 * i] = (uint8_t)wp[i]; d[9 + 2 * i] = (uint8_t)(wp[i] >> 8); }
}

void RansWriter::drain(const uint16_t* const* sym, const int* n, int nblk) {
    BlockCoder bc[kLanes];
    for (int j = 0; j < nblk; j++) {
        if (tmp_[j].size() < (size_t)n[j] + 8) tmp_[j].resize((size_t)n[j] + 8);
        bc[j].start(sym[j], n[j], tmp_[j].data() + tmp_[j].size());
    }
    if (nblk == kLanes) {   // full blocks have equal lengths: step all of them together
        while (bc[0].k >= 1 && bc[kLanes - 1].k >= 1) {
#pragma GCC unroll 4
            for (int j = 0; j < kLanes; j++) bc[j].pair();
        }
    }
    for (int j = 0; j < nblk; j++)
        while (bc[j].k >= 1) bc[j].pair();
    for (int j = 0; j < nblk; j++) emit_block(bc[j].A, bc[j].B, bc[j].wp, tmp_[j].data() + tmp_[j].size());
}

namespace {
// 4*NV full blocks (16386 symbols, an even count, so every step codes one pair per block) with
// AVX2, 8*NV state chains in NV vectors: vector v holds blocks 4v .. 4v+3, lane 2j block 4v+j's
// B state (the lower symbol of each pair), lane 2j+1 its A state. More vectors keep more
// independent chains in flight: a step's chain (renormalise, divide, multiply, add) is long.
// q = x / freq is a double-precision division: x < 2^32 and freq <= 255 convert exactly, the
// quotient is below 2^24, and a correctly rounded quotient of x/freq never crosses an integer
// (a non-integer x/freq lies at least 1/freq below the next one).
// Emitted words go below wp[j], block by block, A before B within a step (the scalar order).
template <int NV>
__attribute__((target("avx2"))) void drain_avx2(const uint16_t* const* sym, int n, uint16_t** wp, uint32_t* r0, uint32_t* r1) {
    const __m256i sign = _mm256_set1_epi32((int)0x80000000u), ff = _mm256_set1_epi32(0xff), c256 = _mm256_set1_epi32(256);
    const __m256d two31 = _mm256_set1_pd(2147483648.0);
    __m256i X[NV];
    for (int v = 0; v < NV; v++) X[v] = _mm256_set1_epi32((int)kLow);
    alignas(32) uint32_t xs[8 * NV];
    for (int k = n - 2; k >= 0; k -= 2) {   // pair (k+1, k): A codes k+1, B codes k
        __m256i S[NV], e[NV];
        int m[NV], any = 0;
#pragma GCC unroll 4
        for (int v = 0; v < NV; v++) {
            uint32_t w[4];
            for (int j = 0; j < 4; j++) std::memcpy(&w[j], sym[4 * v + j] + k, 4);   // low half: symbol k (B), high: k+1 (A)
            S[v] = _mm256_cvtepu16_epi32(_mm_set_epi32((int)w[3], (int)w[2], (int)w[1], (int)w[0]));
            // renormalise where x >= freq << 24 (unsigned compare through the sign flip)
            const __m256i freq = _mm256_and_si256(S[v], ff);
            e[v] = _mm256_xor_si256(_mm256_cmpgt_epi32(_mm256_xor_si256(_mm256_slli_epi32(freq, 24), sign),
                                                       _mm256_xor_si256(X[v], sign)), _mm256_set1_epi32(-1));
            m[v] = _mm256_movemask_ps(_mm256_castsi256_ps(e[v]));
            any |= m[v];
        }
        if (any) {
            for (int v = 0; v < NV; v++) {
                if (!m[v]) continue;
                _mm256_store_si256((__m256i*)(xs + 8 * v), X[v]);
                for (int j = 0; j < 4; j++) {
                    if (m[v] & (2 << (2 * j))) *--wp[4 * v + j] = (uint16_t)xs[8 * v + 2 * j + 1];   // A first
                    if (m[v] & (1 << (2 * j))) *--wp[4 * v + j] = (uint16_t)xs[8 * v + 2 * j];
                }
            }
        }
#pragma GCC unroll 4
        for (int v = 0; v < NV; v++) {
            const __m256i freq = _mm256_and_si256(S[v], ff), start = _mm256_srli_epi32(S[v], 8);
            const __m256i x = _mm256_blendv_epi8(X[v], _mm256_srli_epi32(X[v], 16), e[v]);
            const __m256i xb = _mm256_xor_si256(x, sign);
            const __m256d xl = _mm256_add_pd(_mm256_cvtepi32_pd(_mm256_castsi256_si128(xb)), two31);
            const __m256d xh = _mm256_add_pd(_mm256_cvtepi32_pd(_mm256_extracti128_si256(xb, 1)), two31);
            const __m256d fl = _mm256_cvtepi32_pd(_mm256_castsi256_si128(freq)), fh = _mm256_cvtepi32_pd(_mm256_extracti128_si256(freq, 1));
            const __m256i q = _mm256_set_m128i(_mm256_cvttpd_epi32(_mm256_div_pd(xh, fh)), _mm256_cvttpd_epi32(_mm256_div_pd(xl, fl)));
            X[v] = _mm256_add_epi32(_mm256_add_epi32(x, _mm256_mullo_epi32(q, _mm256_sub_epi32(c256, freq))), start);
        }
    }
    for (int v = 0; v < NV; v++) _mm256_store_si256((__m256i*)(xs + 8 * v), X[v]);
    for (int j = 0; j < 4 * NV; j++) { r1[j] = xs[2 * j]; r0[j] = xs[2 * j + 1]; }
}
}  // namespace

void RansWriter::drain_group(size_t b, int nblocks) {
    const uint16_t* sg[16];
    uint16_t* wp[16];
    uint32_t r0[16], r1[16];
    for (int j = 0; j < nblocks; j++) {
        sg[j] = sym_ + b + (size_t)j * kBatch;
        if (tmp8_[j].size() < (size_t)kBatch + 8) tmp8_[j].resize((size_t)kBatch + 8);
        wp[j] = tmp8_[j].data() + tmp8_[j].size();
    }
    if (nblocks == 16) drain_avx2<4>(sg, kBatch, wp, r0, r1);
    else drain_avx2<2>(sg, kBatch, wp, r0, r1);
    for (int j = 0; j < nblocks; j++) emit_block(r0[j], r1[j], wp[j], tmp8_[j].data() + tmp8_[j].size());
}

void RansWriter::make_room(size_t n) {
    if (g_avx2_rans && n_ >= 16 * (size_t)kBatch) {   // code the first sixteen blocks, keep the rest
        drain_group(0, 16);
        emitted_ = true;
        const size_t done = 16 * (size_t)kBatch;
        std::memmove(sym_, sym_ + done, (n_ - done) * sizeof(uint16_t));
        n_ -= done;
    }
    while (cap_ - n_ < n) grow();
}

void RansWriter::flush() {
    size_t b = 0;
    if (g_avx2_rans) {   // groups of sixteen full blocks, then of eight, then the rest below
        for (; n_ - b >= 16 * (size_t)kBatch; b += 16 * (size_t)kBatch) drain_group(b, 16);
        for (; n_ - b >= 8 * (size_t)kBatch; b += 8 * (size_t)kBatch) drain_group(b, 8);
        if (b == n_ && (b > 0 || emitted_)) { n_ = 0; emitted_ = false; return; }
    }
    do {   // an empty stream still carries its two initial states
        const uint16_t* s[kLanes];
        int n[kLanes], nblk = 0;
        for (; nblk < kLanes && (nblk == 0 || b < n_); nblk++) {
            s[nblk] = sym_ + b;
            n[nblk] = (int)std::min<size_t>(kBatch, n_ - b);
            b += (size_t)n[nblk];
        }
        drain(s, n, nblk);
    } while (b < n_);
    n_ = 0;
    emitted_ = false;
}

void RansWriter::flush_parallel(const std::function<void(int, const std::function<void(int)>&)>& pfor) {
    const size_t nblk = (n_ + kBatch - 1) / kBatch;
    if (!pfor || nblk < 2 || nblk >= 8 || emitted_) { flush(); return; }
    std::vector<std::vector<uint8_t>> part(nblk);
    pfor((int)nblk, [&](int j) {
        RansWriter w(part[j]);
        const size_t b = (size_t)j * kBatch;
        w.code_symbols(sym_ + b, std::min((size_t)kBatch, n_ - b));
    });
    for (const auto& p : part) out_.insert(out_.end(), p.begin(), p.end());
    n_ = 0;
}

void RansWriter::code_symbols(const uint16_t* sym, size_t n) {
    if (n_) throw std::logic_error("code_symbols on a writer with buffered symbols");
    // the blocks of flush(), read from the caller's array (flush only reads sym_)
    sym_ = const_cast<uint16_t*>(sym);
    n_ = n; cap_ = n;
    flush();
    sym_ = buf_.data(); cap_ = buf_.size();
}

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

namespace {
// Significance symbols of positions 0..last of one plane, into out[0..last], with the contexts
// updated as the scalar coder would leave them: each position has its own context, so 16
// positions are modelled at once. P(0) = (zeros << 8) / (zeros + ones) in single precision is
// exact here: the numerator is below 2^16 and the divisor at most 510, so a non-integer quotient
// lies at least 1/510 from the next integer, far beyond the division's rounding error.
__attribute__((target("avx2"))) void sig_symbols_avx2(Ctx* sig, const uint64_t* mask, int last, uint16_t* out) {
    const __m256i bsel = _mm256_setr_epi16(1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192, 16384, (short)0x8000);
    const __m256i iota = _mm256_setr_epi16(0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15);
    const __m256i swap = _mm256_setr_epi8(1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14, 1, 0, 3, 2, 5, 4, 7, 6, 9, 8, 11, 10, 13, 12, 15, 14);
    const __m256i ff = _mm256_set1_epi16(0xff), one = _mm256_set1_epi16(1), c256 = _mm256_set1_epi16(256);
    const __m256i xff00 = _mm256_set1_epi16((short)0xff00), xff01 = _mm256_set1_epi16((short)0xff01), x8100 = _mm256_set1_epi16((short)0x8100);
    const __m256i x100 = _mm256_set1_epi16(0x100), x101 = _mm256_set1_epi16(0x101);
    const __m256i lim = _mm256_set1_epi16((short)(last + 1));
    for (int g = 0; 16 * g <= last; g++) {
        const __m256i st = _mm256_loadu_si256((const __m256i*)(sig + 16 * g));
        const uint32_t b16 = (uint32_t)(mask[g >> 2] >> (16 * (g & 3))) & 0xffffu;
        const __m256i bit = _mm256_cmpeq_epi16(_mm256_and_si256(_mm256_set1_epi16((short)b16), bsel), bsel);
        const __m256i act = _mm256_cmpgt_epi16(lim, _mm256_add_epi16(_mm256_set1_epi16((short)(16 * g)), iota));
        // P(0): (z << 8) / (z + o) truncated to u8, 0 read as 1 (0/0 gives NaN -> 0 -> 1, as the table)
        const __m256i z = _mm256_srli_epi16(st, 8), d = _mm256_add_epi16(z, _mm256_and_si256(st, ff));
        const __m256i zl = _mm256_cvtepu16_epi32(_mm256_castsi256_si128(z)), zh = _mm256_cvtepu16_epi32(_mm256_extracti128_si256(z, 1));
        const __m256i dl = _mm256_cvtepu16_epi32(_mm256_castsi256_si128(d)), dh = _mm256_cvtepu16_epi32(_mm256_extracti128_si256(d, 1));
        const __m256i ql = _mm256_cvttps_epi32(_mm256_div_ps(_mm256_cvtepi32_ps(_mm256_slli_epi32(zl, 8)), _mm256_cvtepi32_ps(dl)));
        const __m256i qh = _mm256_cvttps_epi32(_mm256_div_ps(_mm256_cvtepi32_ps(_mm256_slli_epi32(zh, 8)), _mm256_cvtepi32_ps(dh)));
        const __m256i m8 = _mm256_set1_epi32(0xff);
        __m256i p = _mm256_permute4x64_epi64(_mm256_packus_epi32(_mm256_and_si256(ql, m8), _mm256_and_si256(qh, m8)), 0xD8);
        p = _mm256_blendv_epi8(p, one, _mm256_cmpeq_epi16(p, _mm256_setzero_si256()));
        // symbol: bit 0 -> (start 0, freq p); bit 1 -> (start p, freq 256 - p)
        const __m256i s1 = _mm256_or_si256(_mm256_slli_epi16(p, 8), _mm256_sub_epi16(c256, p));
        _mm256_storeu_si256((__m256i*)(out + 16 * g), _mm256_blendv_epi8(p, s1, bit));
        // BinProb::observe
        const __m256i v = _mm256_blendv_epi8(st, _mm256_shuffle_epi8(st, swap), bit);
        const __m256i sat = _mm256_cmpeq_epi16(_mm256_max_epu16(v, xff00), v);
        const __m256i keep = _mm256_blendv_epi8(x8100, xff00, _mm256_cmpeq_epi16(v, xff01));
        const __m256i halved = _mm256_or_si256(_mm256_srli_epi16(_mm256_add_epi16(v, x101), 1), keep);
        const __m256i n = _mm256_blendv_epi8(_mm256_add_epi16(v, x100), halved, sat);
        const __m256i nn = _mm256_blendv_epi8(n, _mm256_shuffle_epi8(n, swap), bit);
        _mm256_storeu_si256((__m256i*)(sig + 16 * g), _mm256_blendv_epi8(st, nn, act));
    }
}
void sig_symbols_scalar(Ctx* sig, const uint64_t* mask, int last, uint16_t* out) {
    for (int i = 0; i <= last; i++) {
        const bool b = (mask[i >> 6] >> (i & 63)) & 1u;
        const uint32_t p = sig[i].p0();
        sig[i].update(b);
        out[i] = (uint16_t)(b ? ((p << 8) | (256u - p)) : p);
    }
}
const bool g_avx2 = __builtin_cpu_supports("avx2");

inline void emit_bit(uint16_t*& o, bool b, Ctx& c) {
    const uint32_t p = c.p0();
    c.update(b);
    *o++ = (uint16_t)(b ? ((p << 8) | (256u - p)) : p);
}
inline void emit_bypass(uint16_t*& o, bool b) { *o++ = (uint16_t)(b ? 0x8080u : 0x80u); }
inline void emit_class_value(uint16_t*& o, int v, Ctx* ctx, int n) {   // put_class_value, into o
    const int k = value_class(v);
    if (k > n - 1) throw CorruptStream("coefficient magnitude outside the class table");
    for (int i = 0; i < k; i++) emit_bit(o, true, ctx[i]);
    if (k < n - 1) emit_bit(o, false, ctx[k]);
    for (int b = k - 2; b >= 0; b--) emit_bypass(o, ((v - (1 << (k - 1))) >> b) & 1);
}

// One inter plane (no DC prediction) straight from its non-zero masks (nw words) and values.
// Symbol order of put_plane: EOB class value, then per position 0..last the significance bit,
// and after a non-zero one its sign (bypass) and magnitude class value.
void put_plane_packed(uint16_t*& o, ResidualCtx& c, const uint64_t* mask, int nw, const int16_t*& vals, bool luma) {
    int last = 0;
    for (int k = nw - 1; k >= 0; k--) if (mask[k]) { last = 64 * k + 63 - __builtin_clzll(mask[k]); break; }
    emit_class_value(o, last, luma ? c.eob_y : c.eob_c, luma ? 9 : 7);
    uint16_t sym[256];
    Ctx* sig = luma ? c.sig_y : c.sig_c;
    if (g_avx2) sig_symbols_avx2(sig, mask, last, sym);
    else sig_symbols_scalar(sig, mask, last, sym);
    int prev = 0;
    for (int k = 0; k < nw; k++) {
        for (uint64_t m = mask[k]; m; m &= m - 1) {
            const int i = 64 * k + __builtin_ctzll(m);
            std::memcpy(o, sym + prev, sizeof(uint16_t) * (size_t)(i + 1 - prev));
            o += i + 1 - prev;
            prev = i + 1;
            const int v = *vals++;
            emit_bypass(o, v < 0);
            emit_class_value(o, std::abs(v) - 1, luma ? c.mag_y[band_y(i)] : c.mag_c[band_c(i)], 13);
        }
    }
    std::memcpy(o, sym + prev, sizeof(uint16_t) * (size_t)(last + 1 - prev));
    o += last + 1 - prev;
}
}  // namespace

namespace {
// Inter residuals (no DC prediction) from each coded MB's packed record: record(b, m) copies MB b's
// six masks to m and returns its values.
template <typename Rec>
std::vector<uint8_t> encode_inter_records(const uint8_t* coded, int nb, Rec record) {
    std::vector<uint8_t> out;
    bool any = false;
    for (int b = 0; b < nb && !any; b++) any = coded[b] != 0;
    if (!any) return out;   // no coder instantiated when every block is skipped (residual.rs:670-674, :792-798)
    ResidualCtx ctx;
    RansWriter w(out);
    // a block yields at most 3 EOB codes of 16 symbols and 384 positions of 27 symbols
    constexpr size_t kMaxBlockSymbols = 3 * 16 + 384 * 27;
    for (int b = 0; b < nb; b++) {
        if (!coded[b]) continue;
        uint64_t m[6];
        const int16_t* vals = record(b, m);
        uint16_t* o = w.reserve(kMaxBlockSymbols);
        put_plane_packed(o, ctx, m, 4, vals, true);
        put_plane_packed(o, ctx, m + 4, 1, vals, false);
        put_plane_packed(o, ctx, m + 5, 1, vals, false);
        w.commit(o);
    }
    w.flush();
    return out;
}
}  // namespace

std::vector<uint8_t> encode_residuals_packed(const uint8_t* pack, size_t stride, const uint8_t* coded, int nb, bool dc_pred) {
    if (!dc_pred)
        return encode_inter_records(coded, nb, [&](int b, uint64_t* m) {
            const uint8_t* slot = pack + (size_t)b * stride;
            std::memcpy(m, slot, 48);
            return (const int16_t*)(slot + 48);
        });
    std::vector<uint8_t> out;
    bool any = false;
    for (int b = 0; b < nb && !any; b++) any = coded[b] != 0;
    if (!any) return out;
    ResidualCtx ctx;
    RansWriter w(out);
    // intra: DC prediction changes position 0; rare, coded from unpacked blocks
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

namespace {
inline const int16_t* split_record(const uint8_t* main, const uint8_t* over, int b, uint64_t* m) {
    const uint8_t* slot = main + (size_t)b * riv::PACK_MAIN;
    std::memcpy(m, slot, 48);
    int nnz = 0;
    for (int k = 0; k < 6; k++) nnz += __builtin_popcountll(m[k]);
    return (const int16_t*)(nnz <= riv::PACK_MAIN_VALUES ? slot + 48 : over + (size_t)b * riv::PACK_STRIDE + 48);
}
}  // namespace

std::vector<uint8_t> encode_residuals_split(const uint8_t* main, const uint8_t* over, const uint8_t* coded, int nb) {
    return encode_inter_records(coded, nb, [&](int b, uint64_t* m) { return split_record(main, over, b, m); });
}

namespace {
// One position range of the parallel coder: a plane kind (luma, or chroma U then V), positions
// [lo, hi), and whether it codes the plane's EOB class value first.
struct CoefRange { bool luma; int lo, hi; bool eob; };
constexpr int kRanges = 5;
constexpr CoefRange kCoefRanges[kRanges] = {{true, 0, 16, true}, {true, 16, 64, false}, {true, 64, 256, false},
                                            {false, 0, 8, true}, {false, 8, 64, false}};
// Segments per coded MB in the merged order: Y ranges 0, 1, 2; U ranges 3, 4; V ranges 3, 4.
constexpr int kSegs = 7;
constexpr int kSegTask[kSegs] = {0, 1, 2, 3, 4, 3, 4};
constexpr int kSegPart[kSegs] = {0, 0, 0, 0, 0, 1, 1};   // chroma tasks write U (0) then V (1) per MB

// bits lo .. lo + 64 * nout - 1 of an nw-word mask, as nout words
inline void shift_mask(const uint64_t* m, int nw, int lo, uint64_t* out, int nout) {
    for (int k = 0; k < nout; k++) {
        const int bit = lo + 64 * k, w = bit >> 6, sh = bit & 63;
        const uint64_t a = w < nw ? m[w] >> sh : 0u;
        const uint64_t b = (sh && w + 1 < nw) ? m[w + 1] << (64 - sh) : 0u;
        out[k] = a | b;
    }
}

// Positions lo .. min(last, hi - 1) of one plane (EOB first when r.eob), in put_plane's order;
// vals: the plane's values from its first non-zero at or after lo.
void put_range(uint16_t*& o, ResidualCtx& c, const CoefRange& r, const uint64_t* mask, int nw, int last, const int16_t* vals) {
    if (r.eob) emit_class_value(o, last, r.luma ? c.eob_y : c.eob_c, r.luma ? 9 : 7);
    if (last < r.lo) return;
    const int end = std::min(last, r.hi - 1) - r.lo;   // last position of the range, relative to lo
    uint64_t sm[4];
    shift_mask(mask, nw, r.lo, sm, (end >> 6) + 1);
    uint16_t sym[256 + 16];
    Ctx* sig = (r.luma ? c.sig_y : c.sig_c) + r.lo;
    if (g_avx2) sig_symbols_avx2(sig, sm, end, sym);
    else sig_symbols_scalar(sig, sm, end, sym);
    int prev = 0;
    for (int k = 0; 64 * k <= end; k++) {
        uint64_t m = sm[k];
        if (64 * k + 63 > end) m &= end - 64 * k == 63 ? ~0ull : ((2ull << (end - 64 * k)) - 1u);
        for (; m; m &= m - 1) {
            const int i = 64 * k + __builtin_ctzll(m);
            std::memcpy(o, sym + prev, sizeof(uint16_t) * (size_t)(i + 1 - prev));
            o += i + 1 - prev;
            prev = i + 1;
            const int v = *vals++;
            const int pos = i + r.lo;
            emit_bypass(o, v < 0);
            emit_class_value(o, std::abs(v) - 1, r.luma ? c.mag_y[band_y(pos)] : c.mag_c[band_c(pos)], 13);
        }
    }
    std::memcpy(o, sym + prev, sizeof(uint16_t) * (size_t)(end + 1 - prev));
    o += end + 1 - prev;
}

inline int plane_last(const uint64_t* m, int nw) {
    for (int k = nw - 1; k >= 0; k--) if (m[k]) return 64 * k + 63 - __builtin_clzll(m[k]);
    return 0;
}
inline int popcount_below(const uint64_t* m, int lo) {
    int n = 0;
    for (int k = 0; k < (lo >> 6); k++) n += __builtin_popcountll(m[k]);
    if (lo & 63) n += __builtin_popcountll(m[lo >> 6] & ((1ull << (lo & 63)) - 1u));
    return n;
}

// Buffers of one parallel coding, reused from call to call.
struct ParScratch {
    std::vector<int> mbs;                       // coded MBs in raster order
    std::vector<uint16_t> sym[kRanges];         // each range task's symbols
    std::vector<uint32_t> seg[kRanges];         // each task's segment ends: one per MB (luma), two (chroma)
    std::vector<uint32_t> dst;                  // merged start of each coded MB
    std::vector<uint16_t> merged;
    std::vector<std::vector<uint8_t>> chunk;    // each rANS group's bytes
};
std::mutex g_par_mu;
std::vector<std::unique_ptr<ParScratch>> g_par_free;
}  // namespace

std::vector<uint8_t> encode_residuals_split_par(const uint8_t* main, const uint8_t* over, const uint8_t* coded, int nb,
                                                const ParallelFor& pfor) {
    std::unique_ptr<ParScratch> S;
    {
        std::lock_guard<std::mutex> lk(g_par_mu);
        if (!g_par_free.empty()) { S = std::move(g_par_free.back()); g_par_free.pop_back(); }
    }
    if (!S) S.reset(new ParScratch());
    S->mbs.clear();
    for (int b = 0; b < nb; b++) if (coded[b]) S->mbs.push_back(b);
    std::vector<uint8_t> out;
    const int n = (int)S->mbs.size();
    if (n == 0) {   // no coder instantiated when every block is skipped (residual.rs:670-674, :792-798)
        std::lock_guard<std::mutex> lk(g_par_mu);
        g_par_free.push_back(std::move(S));
        return out;
    }
    // 1. each range over every coded MB, with its own contexts
    std::vector<std::string> errs(kRanges);
    pfor(kRanges, [&](int t) {
        try {
            const CoefRange& r = kCoefRanges[t];
            ResidualCtx c;
            std::vector<uint16_t>& buf = S->sym[t];
            std::vector<uint32_t>& seg = S->seg[t];
            seg.resize((size_t)n * (r.luma ? 1 : 2));
            const size_t bound = (size_t)(r.hi - r.lo) * 27 + 32;   // symbols of one plane's range at most
            if (buf.size() < 4 * bound) buf.resize(4 * bound);
            size_t used = 0;
            for (int k = 0; k < n; k++) {
                uint64_t m[6];
                const int16_t* vals = split_record(main, over, S->mbs[k], m);
                const int ny = __builtin_popcountll(m[0]) + __builtin_popcountll(m[1]) + __builtin_popcountll(m[2]) +
                               __builtin_popcountll(m[3]);
                for (int part = 0; part < (r.luma ? 1 : 2); part++) {
                    if (buf.size() - used < bound) buf.resize(std::max(2 * buf.size(), used + bound));
                    uint16_t* o = buf.data() + used;
                    const uint64_t* pm = r.luma ? m : m + 4 + part;
                    const int nw = r.luma ? 4 : 1;
                    const int16_t* pv = r.luma ? vals : vals + ny + (part ? __builtin_popcountll(m[4]) : 0);
                    put_range(o, c, r, pm, nw, plane_last(pm, nw), pv + popcount_below(pm, r.lo));
                    used = (size_t)(o - buf.data());
                    seg[(size_t)k * (r.luma ? 1 : 2) + part] = (uint32_t)used;
                }
            }
        } catch (const std::exception& ex) {
            errs[t] = ex.what();
        }
    });
    for (const std::string& e : errs) if (!e.empty()) throw CorruptStream(e);
    // 2. the merged order (per MB the seven segments; starts by prefix sum), built and rANS-coded
    // in groups of sixteen blocks of 16,386 symbols (the stream's blocks are independent): each
    // group task copies the part of the merged order it covers, then codes it
    auto seg_bounds = [&](int k, int s, uint32_t& a, uint32_t& b) {
        const int t = kSegTask[s], per = kCoefRanges[t].luma ? 1 : 2, i = k * per + kSegPart[s];
        const std::vector<uint32_t>& e = S->seg[t];
        a = i ? e[(size_t)i - 1] : 0u;
        b = e[(size_t)i];
    };
    S->dst.resize((size_t)n + 1);
    uint32_t total = 0;
    for (int k = 0; k < n; k++) {
        S->dst[k] = total;
        for (int s = 0; s < kSegs; s++) { uint32_t a, b; seg_bounds(k, s, a, b); total += b - a; }
    }
    S->dst[n] = total;
    if (S->merged.size() < total) S->merged.resize(total);
    constexpr size_t kGroup = 16 * 16386;
    const int ng = (int)((total + kGroup - 1) / kGroup);
    if ((int)S->chunk.size() < ng) S->chunk.resize(ng);
    pfor(ng, [&](int g) {
        const uint32_t lo = (uint32_t)((size_t)g * kGroup), hi = (uint32_t)std::min((size_t)total, (size_t)(g + 1) * kGroup);
        // the MBs overlapping [lo, hi): from the last one starting at or before lo
        int k = (int)(std::upper_bound(S->dst.begin(), S->dst.begin() + n, lo) - S->dst.begin()) - 1;
        for (; k < n && S->dst[k] < hi; k++) {
            uint32_t at = S->dst[k];
            for (int s = 0; s < kSegs; s++) {
                uint32_t a, b;
                seg_bounds(k, s, a, b);
                const uint32_t c0 = std::max(at, lo), c1 = std::min(at + (b - a), hi);   // this segment's part in range
                if (c0 < c1) std::memcpy(S->merged.data() + c0, S->sym[kSegTask[s]].data() + a + (c0 - at), sizeof(uint16_t) * (c1 - c0));
                at += b - a;
            }
        }
        std::vector<uint8_t>& o = S->chunk[g];
        o.clear();
        RansWriter w(o);
        w.code_symbols(S->merged.data() + lo, hi - lo);
    });
    size_t bytes = 0;
    for (int g = 0; g < ng; g++) bytes += S->chunk[g].size();
    out.reserve(bytes);
    for (int g = 0; g < ng; g++) out.insert(out.end(), S->chunk[g].begin(), S->chunk[g].end());
    std::lock_guard<std::mutex> lk(g_par_mu);
    g_par_free.push_back(std::move(S));
    return out;
}

std::vector<uint8_t> code_symbol_stream(const uint16_t* sym, size_t n, const ParallelFor& pfor, int group_blocks) {
    std::vector<uint8_t> out;
    if (n == 0) return out;
    if (group_blocks != 16 && group_blocks != 8) throw std::invalid_argument("rANS block groups are of 16 or 8 blocks");
    const size_t kGroup = (size_t)group_blocks * 16386;
    const int ng = (int)((n + kGroup - 1) / kGroup);
    if (ng == 1 || !pfor) {
        RansWriter w(out);
        w.code_symbols(sym, n);
        return out;
    }
    std::vector<std::vector<uint8_t>> chunk(ng);
    pfor(ng, [&](int g) {
        RansWriter w(chunk[g]);
        const size_t b = (size_t)g * kGroup;
        w.code_symbols(sym + b, std::min(kGroup, n - b));
    });
    size_t bytes = 0;
    for (const auto& c : chunk) bytes += c.size();
    out.reserve(bytes);
    for (const auto& c : chunk) out.insert(out.end(), c.begin(), c.end());
    return out;
}

void unpack_split(const uint8_t* main, const uint8_t* over, int b, int16_t* q) {
    uint64_t m[6];
    const int16_t* v = split_record(main, over, b, m);
    std::memset(q, 0, 384 * sizeof(int16_t));
    for (int k = 0; k < 6; k++)
        for (uint64_t w = m[k]; w; w &= w - 1) q[64 * k + __builtin_ctzll(w)] = *v++;
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
// The decisions of the mode tree (New | Zero | Near | Temporal | TopRight | TopLeft | Nearest,
// mv_rans/encoder.rs:257-301) per mode, and of the class tree (a binary search over 0..10,
// encode_component, mv_rans/encoder.rs:303-331) per class: the count, and the bits by level.
struct MvTrees {
    uint8_t mode_n[8], mode_bits[8], cls_n[11], cls_bits[11];
    MvTrees() {
        static const uint8_t order[6] = {MNEW, 0xff, MNEAR, MTEMP, MTR, MTL};
        for (int m = 0; m < 8; m++) {
            int n = 0, bits = 0;
            for (int lvl = 0; lvl < 6; lvl++) {
                const bool bit = lvl == 1 ? m != MZERO : m == order[lvl];
                bits |= (int)bit << lvl;
                n++;
                if (lvl == 1 ? !bit : bit) break;
            }
            mode_n[m] = (uint8_t)n; mode_bits[m] = (uint8_t)bits;
        }
        for (int k = 0; k <= 10; k++) {
            int lo = 0, hi = 10, depth = 0, bits = 0;
            while (lo < hi) { const int mid = (lo + hi) / 2; const bool up = k > mid; bits |= (int)up << depth++; if (up) lo = mid + 1; else hi = mid; }
            cls_n[k] = (uint8_t)depth; cls_bits[k] = (uint8_t)bits;
        }
    }
};
const MvTrees kMvTrees;
// For every context state c and bit b, one word: the state after observing b (Ctx::update) in the
// high half, the rANS symbol that codes b at c (start << 8 | freq, from Ctx::p0) in the low half.
// One load per decision and no branch: the saturated-counter rule is in the table too.
struct CtxStepTable {
    uint32_t w[2][65536];
    CtxStepTable() {
        for (uint32_t c = 0; c < 65536; c++)
            for (uint32_t b = 0; b < 2; b++) {
                Ctx t; t.c = (uint16_t)c;
                const uint32_t pz = t.p0();
                t.update(b != 0);
                w[b][c] = ((uint32_t)t.c << 16) | (b ? ((pz << 8) | (256u - pz)) : pz);
            }
    }
};
const CtxStepTable& ctx_steps() { static const CtxStepTable t; return t; }
// one coded decision: context index | bit << 8
inline uint16_t mv_event(int ctx, bool bit) { return (uint16_t)(ctx | ((int)bit << 8)); }
constexpr int kMvMaxEvents = 42;   // per block: skip 1, mode 6, base 4, hit 1, 2 x (class 4, sign 1, magnitude 9), frac 2

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
void MvCoder::reset_history() {
    cand_.clear(); live_.clear(); index_.clear(); total_ = 0;
    for (uint32_t& k : cand_key_) k = ~0u;
}
void MvCoder::reset_contexts() { for (Ctx& c : ctx_) c = Ctx(); }

namespace {
inline uint32_t mv_key(int dx, int dy, int rx, int ry) { return ((uint32_t)(dx + 128) << 24) | ((uint32_t)(dy + 128) << 16) | ((uint32_t)rx << 8) | (uint32_t)ry; }
inline uint32_t mv_hash(uint32_t k) { return (k * 0x9E3779B1u) >> 7; }
}  // namespace

int MvCoder::lookup(int8_t dx, int8_t dy, int8_t sx, int8_t sy) const {
    if (cand_.empty()) return -1;
    const uint32_t k = mv_key(dx, dy, subpel_rank(sx), subpel_rank(sy));
    int r = -1;
    for (int i = 15; i >= 0; i--) r = cand_key_[i] == k ? i : r;   // the first match
    return r;
}

// The count of key (nullptr when absent and !insert; a new entry counts 0).
uint32_t* MvCoder::find_count(uint32_t key, bool insert) {
    const uint32_t mask = (uint32_t)index_.size() - 1;
    for (uint32_t h = mv_hash(key) & mask;; h = (h + 1) & mask) {
        const uint32_t s = index_[h];
        if (s == 0) {
            if (!insert) return nullptr;
            live_.push_back({key, 0u});
            index_[h] = (uint32_t)live_.size();
            return &live_.back().second;
        }
        if (live_[s - 1].first == key) return &live_[s - 1].second;
    }
}

// update_candidates (mv_rans/encoder.rs:370-412, decoder.rs:283-326): decay every count by
// 7/8, count this frame's NEW (delta, sub-pel) tuples, drop zero counts; below 100 NEW blocks
// of history the list is empty, else the top 16 by count (ties by key) above total/100.
void MvCoder::refresh_history(const std::vector<MvBlock>& blocks, bool from_decoder) {
    for (auto& e : live_) e.second = e.second * 7 / 8;
    total_ = total_ * 7 / 8;
    // index sized for every live key plus one new key per block at most half full
    size_t need = 64;
    while (need < 2 * (live_.size() + blocks.size())) need <<= 1;
    if (index_.size() != need) index_.assign(need, 0u);
    for (size_t i = 0; i < live_.size(); i++) {   // index the live keys (each once)
        const uint32_t mask = (uint32_t)need - 1;
        uint32_t h = mv_hash(live_[i].first) & mask;
        while (index_[h]) h = (h + 1) & mask;
        index_[h] = (uint32_t)i + 1;
    }
    for (const MvBlock& b : blocks) {
        if (b.mode != MNEW) continue;
        if (b.sx == -1 || b.sy == -1) {
            if (from_decoder) continue;
            throw CorruptStream("-0.5 sub-pel in a NEW vector");
        }
        (*find_count(mv_key(b.dx, b.dy, subpel_rank(b.sx), subpel_rank(b.sy)), true))++;
        total_++;
    }
    live_.erase(std::remove_if(live_.begin(), live_.end(), [](const std::pair<uint32_t, uint32_t>& e) { return e.second == 0; }), live_.end());
    std::fill(index_.begin(), index_.end(), 0u);   // re-indexed from the live list at the next refresh
    cand_.clear();
    for (uint32_t& k : cand_key_) k = ~0u;
    if (total_ < 100) return;
    const uint32_t floor_count = total_ / 100;
    std::vector<std::pair<uint32_t, uint32_t>> order;   // (count, key) above the floor
    for (const auto& e : live_) if (e.second > floor_count) order.push_back({e.second, e.first});
    std::sort(order.begin(), order.end(), [](const std::pair<uint32_t, uint32_t>& a, const std::pair<uint32_t, uint32_t>& b) {
        return a.first != b.first ? a.first > b.first : a.second < b.second;
    });
    for (const auto& e : order) {
        const uint32_t k = e.second;
        cand_key_[cand_.size()] = k;
        cand_.push_back({(int)(k >> 24) - 128, (int)((k >> 16) & 0xff) - 128, (int)((k >> 8) & 0xff), (int)(k & 0xff)});
        if (cand_.size() == 16) break;
    }
}

std::vector<uint8_t> MvCoder::encode(const std::vector<MvBlock>& blocks, int bw, int bh) {
    std::vector<uint8_t> out;
    RansWriter w(out);
    model(blocks, bw, bh, w);
    w.flush();
    return out;
}

void MvCoder::model(const std::vector<MvBlock>& blocks, int bw, int bh, RansWriter& w) {
    (void)bh;
    thread_local std::vector<uint16_t> evs;
    const size_t n = events(blocks, bw, evs);
    contexts(evs.data(), n, w);
}

size_t MvCoder::events(const std::vector<MvBlock>& blocks, int bw, std::vector<uint16_t>& evs) {
    const int nb = (int)blocks.size();
    // the frame's decisions in coding order, as (context, bit) events; every tree writes its full
    // depth and advances by the decisions it takes, so the loop has few branches
    thread_local std::vector<uint8_t> hitv;
    if (evs.size() < (size_t)nb * kMvMaxEvents + 16) evs.resize((size_t)nb * kMvMaxEvents + 16);
    hitv.assign((size_t)nb, 0);
    uint16_t* e = evs.data();
    const MvTrees& T = kMvTrees;
    auto score = [](uint8_t m) { return m == MZERO ? 2 : m == MNEW ? 0 : 1; };
    auto component = [&T](uint16_t* e, int8_t d, int axis) {   // encode_component (mv_rans/encoder.rs:303-331)
        const int mag = std::abs((int)d), k = mv_class(mag), cb = T.cls_bits[k];
        for (int depth = 0; depth < 4; depth++) e[depth] = mv_event(kCls + axis * 4 + depth, (cb >> depth) & 1);
        e += T.cls_n[k];
        if (k == 0) return e;
        *e++ = mv_event(kSign + axis, d < 0);
        const int off = mag - (1 << (k - 1));
        for (int b = 0; b < 9; b++) e[b] = mv_event(kMag + axis * 10 + b, (off >> b) & 1);
        return e + (k - 1);
    };
    const bool have_cand = !cand_.empty();
    for (int i = 0, bx = 0; i < nb; i++, bx = bx + 1 == bw ? 0 : bx + 1) {
        const MvBlock& b = blocks[i];
        *e++ = mv_event(kSkip, b.skip);
        const int ctx = std::min((bx > 0 ? score(blocks[i - 1].mode) : 0) + (i >= bw ? score(blocks[i - bw].mode) : 0), 3);
        const int m = b.mode < 8 ? b.mode : 7, mb = T.mode_bits[m];
        for (int lvl = 0; lvl < 6; lvl++) e[lvl] = mv_event(kMode + lvl * 4 + ctx, (mb >> lvl) & 1);
        e += T.mode_n[m];
        if (b.mode != MNEW) continue;
        // base selector (nearest, near, top-right, top-left | temporal)
        for (int lvl = 0; lvl < 4; lvl++) e[lvl] = mv_event(kBase + lvl * 4 + ctx, b.new_base == lvl);
        e += b.new_base < 4 ? b.new_base + 1 : 4;
        bool hit = false;
        if (have_cand) {
            const int ci = lookup(b.dx, b.dy, b.sx, b.sy);
            hit = ci >= 0;
            *e++ = mv_event(kHit + (bx > 0 ? hitv[i - 1] : 0) + (i >= bw ? hitv[i - bw] : 0), hit);
            if (hit) {   // the index, a binary search over 16: its 4 bits, most significant first
                for (int d = 0; d < 4; d++) e[d] = mv_event(kIdx + d, (ci >> (3 - d)) & 1);
                e += 4;
            }
        }
        if (!hit) {
            e = component(e, b.dx, 0);
            e = component(e, b.dy, 1);
            if (b.sx == -1 || b.sy == -1) throw CorruptStream("-0.5 sub-pel in a NEW vector");
            e[0] = mv_event(kFrac, b.sx != 0);
            e[1] = mv_event(kFrac + 1, b.sy != 0);
            e += 2;
        }
        hitv[i] = hit;
    }
    refresh_history(blocks, false);
    return (size_t)(e - evs.data());
}

void MvCoder::contexts(const uint16_t* ev, size_t n, RansWriter& w) {
    // the adaptive contexts over the events, in order, each giving its rANS symbol: the states
    // live in 32-bit words here (any 8-bit context index stays in bounds), and each decision is
    // one table word (the next state and the symbol), 2.6x faster than updating 16-bit Ctx
    // states in place with the saturated case branched (measured on 1080p P-frames' events)
    const CtxStepTable& T = ctx_steps();
    uint32_t cs[256];
    for (int i = 0; i < 256; i++) cs[i] = i < kNumCtx ? ctx_[i].c : 0x0101u;
    uint16_t* o = w.reserve(n);
    for (size_t j = 0; j < n; j++) {
        const uint32_t v = ev[j], t = T.w[v >> 8][cs[v & 0xffu]];
        cs[v & 0xffu] = t >> 16;
        o[j] = (uint16_t)t;
    }
    w.commit(o + n);
    for (int i = 0; i < kNumCtx; i++) ctx_[i].c = (uint16_t)cs[i];
}

std::vector<MvBlock> MvCoder::decode(const uint8_t* data, size_t len, int bw, int bh) {
    RansReader r(data, len);
    const int nb = bw * bh;
    std::vector<MvBlock> res(nb);
    std::vector<uint8_t> modes, hits;
    modes.reserve(nb); hits.reserve(nb);
    auto component = [&](int axis) -> int8_t {   // decode_component (mv_rans/decoder.rs:375-401)
        int lo = 0, hi = 10, depth = 0;
        while (lo < hi) { int mid = (lo + hi) / 2; if (r.bit(ctx_[kCls + axis * 4 + depth++])) lo = mid + 1; else hi = mid; }
        int k = lo;
        if (k == 0) return 0;
        bool neg = r.bit(ctx_[kSign + axis]);
        int off = 0;
        for (int b = 0; b < k - 1; b++) if (r.bit(ctx_[kMag + axis * 10 + b])) off |= 1 << b;
        int v = (1 << (k - 1)) + off;
        v = neg ? -v : v;
        return (int8_t)std::max(-128, std::min(127, v));
    };
    for (int i = 0; i < nb; i++) {
        MvBlock b{};
        const int bx = i % bw;
        b.skip = r.bit(ctx_[kSkip]);
        const int ctx = mode_ctx(modes, bw, i);
        if (r.bit(ctx_[kMode + ctx])) b.mode = MNEW;
        else if (!r.bit(ctx_[kMode + 4 + ctx])) b.mode = MZERO;
        else if (r.bit(ctx_[kMode + 8 + ctx])) b.mode = MNEAR;
        else if (r.bit(ctx_[kMode + 12 + ctx])) b.mode = MTEMP;
        else if (r.bit(ctx_[kMode + 16 + ctx])) b.mode = MTR;
        else if (r.bit(ctx_[kMode + 20 + ctx])) b.mode = MTL;
        else b.mode = MNEAREST;
        bool hit = false;
        if (b.mode == MNEW) {
            b.new_base = 4;
            for (int lvl = 0; lvl < 4; lvl++) if (r.bit(ctx_[kBase + lvl * 4 + ctx])) { b.new_base = (uint8_t)lvl; break; }
            bool frac_done = false;
            if (!cand_.empty()) {
                int hctx = (bx > 0 ? hits[i - 1] : 0) + (i >= bw ? hits[i - bw] : 0);
                hit = r.bit(ctx_[kHit + hctx]);
                if (hit) {
                    int lo = 0, hi = 16, depth = 0;
                    while (hi - lo > 1) { int mid = (lo + hi) / 2; if (r.bit(ctx_[kIdx + depth++])) lo = mid; else hi = mid; }
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
            if (!frac_done) { b.sx = r.bit(ctx_[kFrac]) ? 1 : 0; b.sy = r.bit(ctx_[kFrac + 1]) ? 1 : 0; }
        }
        hits.push_back(hit);
        modes.push_back(b.mode);
        res[i] = b;
    }
    refresh_history(res, true);
    return res;
}

}  // namespace rivh
