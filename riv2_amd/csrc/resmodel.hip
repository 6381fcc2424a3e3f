// Context modelling of a P-frame's residual symbols on the GPU: the frame's whole symbol sequence,
// each symbol as (start << 8 | freq), exactly as the host residual coder buffers it for its rANS32
// coder (entropy.cpp; reitero_residual/src/rans.rs:43-155 and residual.rs:670-809). The rANS32
// coding of the sequence stays on the host.
//
// The sequence is a function of the coefficients and of the adaptive binary contexts (BinProb,
// rans.rs:25-85), and every context is its own chain: its probability at a use depends only on the
// bits it observed before, in coding order. So the frame splits into
//   * a layout, free of contexts: per MB and plane the symbol count, and so every symbol's position
//     (the EOB class value, then per position up to the last non-zero its significance bit and,
//     after a non-zero, the sign and the magnitude class value, put_plane, rans.rs:189-327), and the
//     bypass symbols (offset bits and signs), which need no context;
//   * 418 chains, one per context: significance of each luma position (256) and chroma position
//     (64, U then V per MB), the EOB class trees (8 luma, 6 chroma contexts) and the magnitude class
//     trees per band (4 luma and 3 chroma bands, 12 contexts each). Each chain walks its uses in
//     coding order, 64 at a time, one per lane, from a word per use that the layout wrote where the
//     chain reads it in one coalesced load (the next round's load is issued before a round runs).
//
// Within 64 uses, a context's state before use j is its state before the first plus the zeros and
// ones observed in between, until an observed counter saturates (255); a saturating use halves both
// counters, or, when the other counter is 1, leaves the state as it is. So a chain's 64 uses take a
// prefix count, and the pass repeats from the first saturating use; a run of uses that leave the
// state as it is costs one more step.
//
// Kernels, in stream order: count (per MB: symbols per plane, per mask word the position of its
// first significance symbol, non-zeros per band), scan (each MB's first symbol and first event per
// band), walk (per MB and mask word: the chains' words, the bypass symbols, the band events), the
// chains, and push (the symbols to mapped host memory).
#include <hip/hip_runtime.h>
#include "riv_common.hpp"
#include "riv_kernels.hpp"

namespace {
using riv::PACK_STRIDE;

// chain words: bit 31 the context is used, bits 1..30 the symbol's position, bit 0 the bit observed
constexpr uint32_t kUsed = 0x80000000u;
__device__ __forceinline__ uint32_t chain_word(bool act, uint32_t pos, int bit) { return act ? (kUsed | (pos << 1) | (uint32_t)bit) : 0u; }

// band event lists: band b's region starts at band_region(b) * nb entries and holds at most
// (positions in the band) * nb events. Luma bands: position 0, 1-15, 16-63, 64-255; chroma (U
// and V in one list, U first per MB): 0, 1-7, 8-63.
__device__ __forceinline__ int band_region(int b) {
    switch (b) { case 0: return 0; case 1: return 1; case 2: return 16; case 3: return 64; case 4: return 256; case 5: return 258; default: return 272; }
}

// value class: 0, 1, then k for [2^(k-1), 2^k - 1] (rans.rs:43-50)
__device__ __forceinline__ int value_class(int v) { return v <= 1 ? v : 32 - __clz(v); }
// symbols of a class value coded over n contexts: k ones, a terminating zero when k < n - 1, and
// k - 1 offset bits (put_class_value, rans.rs:69-96)
__device__ __forceinline__ int class_symbols(int k, int n) { return k + (k < n - 1 ? 1 : 0) + (k > 1 ? k - 1 : 0); }
__device__ __forceinline__ uint16_t bypass_symbol(int bit) { return bit ? 0x8080u : 0x80u; }

// P(bit == 0) * 256 of a context with counts z, o (both 1..255): floor(256 z / (z + o)), read as
// 1 when 0 (rans.rs:30-44). rcp[d] = floor(2^32 / d) + 1 for d = z + o in 2..510: for x < 2^16,
// x * rcp[d] / 2^32 exceeds x / d by less than 2^-16, below the gap 1/d from a non-integer x / d
// to the next integer, so its floor is the quotient's.
__device__ __forceinline__ int pzero(int z, int o, const uint32_t* rcp) {
    const int q = (int)(__umulhi((uint32_t)(z << 8), rcp[z + o]) & 0xffu);
    return q ? q : 1;
}

__device__ __forceinline__ uint16_t bit_symbol(int p, int bit) { return bit ? (uint16_t)((p << 8) | (256 - p)) : (uint16_t)p; }

// One round of a chain: lane j (in coding order) uses the context when `act`, observing `bit`,
// and gets its symbol in `out`. (z, o): the context's counts before the round, then after it.
__device__ __forceinline__ void chain_round(int& z, int& o, bool act, int bit, uint16_t& out, int lane, const uint32_t* rcp) {
    const uint64_t A = __ballot(act), Om = __ballot(act && bit), Zm = A & ~Om;
    const uint64_t lt = (1ull << lane) - 1ull;
    uint64_t U = A;   // uses not yet coded: always a suffix of A
    while (U) {
        // counts before each use, if no counter saturates on the way
        const int zj = z + __popcll(Zm & U & lt), oj = o + __popcll(Om & U & lt);
        const bool mine = (U >> lane) & 1ull;
        const uint64_t S = __ballot(mine && (bit ? oj >= 255 : zj >= 255));
        if (!S) {
            if (mine) out = bit_symbol(pzero(zj, oj, rcp), bit);
            z += __popcll(Zm & U);
            o += __popcll(Om & U);
            return;
        }
        // uses up to the first saturating one are exact
        const int js = __ffsll((unsigned long long)S) - 1;
        const uint64_t C = U & (js == 63 ? ~0ull : ((2ull << js) - 1ull));
        if ((C >> lane) & 1ull) out = bit_symbol(pzero(zj, oj, rcp), bit);
        const int zs = __shfl(zj, js, 64), os = __shfl(oj, js, 64), bs = __shfl(bit, js, 64);
        // BinProb::observe on a saturated counter (rans.rs:73-84): both halved and the observed one
        // re-based, or nothing when the other counter is 1
        const int oth = bs ? zs : os;
        const bool keep = oth == 1;
        if (!keep) {
            const int h = (oth + 1) >> 1;
            z = bs ? h : 129;
            o = bs ? 129 : h;
        } else {
            z = zs;
            o = os;
        }
        U &= ~C;
        if (keep && U) {
            // every further use observing the same bit also keeps the state, up to the first that
            // observes the other one
            const uint64_t B = U & (bs ? Zm : Om);
            const uint64_t run = B ? (U & ((1ull << (__ffsll((unsigned long long)B) - 1)) - 1ull)) : U;
            if ((run >> lane) & 1ull) out = bit_symbol(pzero(z, o, rcp), bit);
            U &= ~run;
        }
    }
}

// A round's exit state from its entry state (z, o), from its use and bit masks, without coding
// the uses one by one: while some counter would reach 255 among the remaining uses R, find the
// first use that saturates one; if it halves both counters, the uses after it (fewer than 64)
// cannot saturate again and only add; if it leaves the state as it is (the other counter is 1),
// so do the following uses of the same bit, and the first use of the other bit steps that counter
// from 1 to 2. Each pass is a few scalar operations and one ballot; a round takes one to three.
__device__ __forceinline__ void round_exit(int& z, int& o, unsigned long long A, unsigned long long Om, int lane) {
    const unsigned long long Zm0 = A & ~Om;
    const unsigned long long lt = (1ull << lane) - 1ull;
    unsigned long long R = A;
    for (;;) {
        const unsigned long long Zm = Zm0 & R, Oq = Om & R;
        const int nz = __popcll(Zm), no = __popcll(Oq);
        if (z + nz <= 255 && o + no <= 255) {
            z += nz;
            o += no;
            return;
        }
        const bool sat = (((Zm >> lane) & 1ull) && z + __popcll(Zm & lt) == 255) || (((Oq >> lane) & 1ull) && o + __popcll(Oq & lt) == 255);
        const int js = __ffsll(__ballot(sat)) - 1;   // exists: a counter reaches 255 among R
        const unsigned long long before = (1ull << js) - 1ull, after = js == 63 ? 0ull : ~((2ull << js) - 1ull);
        const int bs = (int)((Om >> js) & 1ull);
        const int zs = z + __popcll(Zm & before), os = o + __popcll(Oq & before);
        const int oth = bs ? zs : os;
        if (oth != 1) {
            const int h = (oth + 1) >> 1;
            z = (bs ? h : 129) + __popcll(Zm & after);
            o = (bs ? 129 : h) + __popcll(Oq & after);
            return;
        }
        // (255, 1) or (1, 255): kept through the same-bit uses up to the first of the other bit
        z = zs;
        o = os;
        const unsigned long long B = (bs ? Zm : Oq) & after;
        if (!B) return;
        const int jb = __ffsll(B) - 1;
        if (bs) z += 1; else o += 1;
        R = jb == 63 ? 0ull : (R & ~((2ull << jb) - 1ull));
    }
}

// Per MB (one wave): per plane the last non-zero, the EOB class and the symbol count; per mask
// word the MB-relative position of its first significance symbol; per band the non-zeros.
// Non-coded MBs get zero counts. bad[b]: a magnitude beyond the class table (the host coder
// rejects such a frame).
__global__ __launch_bounds__(256) void k_rm_count(const uint8_t* __restrict__ pack, const uint8_t* __restrict__ coded, int nb,
                                                  uint16_t* __restrict__ cnt, uint8_t* __restrict__ last, uint8_t* __restrict__ ke,
                                                  uint16_t* __restrict__ bcnt, uint16_t* __restrict__ wst, uint8_t* __restrict__ bad) {
    const int b = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (b >= nb) return;
    if (!coded[b]) {
        if (lane < 3) cnt[lane * nb + b] = 0;
        if (lane < 7) bcnt[lane * nb + b] = 0;
        if (lane == 0) bad[b] = 0;
        return;
    }
    const uint8_t* slot = pack + (size_t)b * PACK_STRIDE;
    const int16_t* vals = (const int16_t*)(slot + 48);
    uint64_t m[6];
#pragma unroll
    for (int k = 0; k < 6; k++) m[k] = ((const uint64_t*)slot)[k];
    const uint64_t lt = (1ull << lane) - 1ull;
    int rank = 0;   // non-zeros of the MB before the current word
    int plstart = 0;
    bool err = false;
    uint16_t ws[6];
    for (int pl = 0; pl < 3; pl++) {
        const int w0 = pl == 0 ? 0 : 3 + pl, nw = pl == 0 ? 4 : 1;
        int lp = 0;
        for (int k = w0 + nw - 1; k >= w0; k--)
            if (m[k]) { lp = 64 * (k - w0) + 63 - __clzll((long long)m[k]); break; }
        const int kk = value_class(lp), neob = class_symbols(kk, pl == 0 ? 9 : 7);
        int run = 0;   // symbols of this plane's non-zeros so far
        for (int k = w0; k < w0 + nw; k++) {
            ws[k] = (uint16_t)(plstart + neob + 64 * (k - w0) + run);
            const bool nz = (m[k] >> lane) & 1ull;
            int ns = 0;
            if (nz) {
                const int v = vals[rank + __popcll(m[k] & lt)];
                const int km = value_class((v < 0 ? -v : v) - 1);
                err |= km > 12;
                ns = 1 + class_symbols(km, 13);   // sign, magnitude class value
            }
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int y = __shfl_xor(ns, d, 64);
                ns += y;
            }
            run += ns;   // the word's total, on every lane
            rank += __popcll(m[k]);
        }
        const int total = neob + lp + 1 + run;
        if (lane == 0) {
            last[pl * nb + b] = (uint8_t)lp;
            ke[pl * nb + b] = (uint8_t)kk;
            cnt[pl * nb + b] = (uint16_t)total;
        }
        plstart += total;
    }
    const bool any_err = __ballot(err) != 0ull;
    if (lane < 6)
        wst[6 * b + lane] = lane == 0 ? ws[0] : lane == 1 ? ws[1] : lane == 2 ? ws[2] : lane == 3 ? ws[3] : lane == 4 ? ws[4] : ws[5];
    if (lane == 0) {
        bad[b] = any_err ? 1 : 0;
        bcnt[0 * nb + b] = (uint16_t)(m[0] & 1ull);
        bcnt[1 * nb + b] = (uint16_t)__popcll(m[0] & 0xfffeull);
        bcnt[2 * nb + b] = (uint16_t)__popcll(m[0] & ~0xffffull);
        bcnt[3 * nb + b] = (uint16_t)(__popcll(m[1]) + __popcll(m[2]) + __popcll(m[3]));
        bcnt[4 * nb + b] = (uint16_t)((m[4] & 1ull) + (m[5] & 1ull));
        bcnt[5 * nb + b] = (uint16_t)(__popcll(m[4] & 0xfeull) + __popcll(m[5] & 0xfeull));
        bcnt[6 * nb + b] = (uint16_t)(__popcll(m[4] & ~0xffull) + __popcll(m[5] & ~0xffull));
    }
}

// Exclusive prefix sums over MBs, in tiles of 1024 (a workgroup of four waves per tile, four
// consecutive MBs per thread; every load before any store): each MB's first symbol and its first
// event in each band's list, relative to its tile, and the tile's totals (tt, 8 per tile: symbols,
// then the seven bands' events) and whether any MB of it has a magnitude beyond the class table.
__global__ __launch_bounds__(256) void k_rm_scan(int nb, const uint16_t* __restrict__ cnt, const uint16_t* __restrict__ bcnt,
                                                 const uint8_t* __restrict__ bad, uint32_t* __restrict__ off, uint32_t* __restrict__ boff,
                                                 uint32_t* __restrict__ tt) {
    __shared__ uint32_t wsum[4][8];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int b0 = blockIdx.x * 1024 + 4 * t;
    uint32_t v[4][8];
    int anybad = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int b = b0 + i;
        const bool in = b < nb;
        v[i][0] = in ? (uint32_t)cnt[b] + cnt[nb + b] + cnt[2 * nb + b] : 0u;
#pragma unroll
        for (int j = 0; j < 7; j++) v[i][1 + j] = in ? bcnt[j * nb + b] : 0u;
        anybad |= in ? bad[b] : 0;
    }
    uint32_t incl[8], mine[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        mine[j] = v[0][j] + v[1][j] + v[2][j] + v[3][j];
        incl[j] = mine[j];
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t y = __shfl_up(incl[j], d, 64);
            if (lane >= d) incl[j] += y;
        }
    }
    if (lane == 63)
#pragma unroll
        for (int j = 0; j < 8; j++) wsum[wv][j] = incl[j];
    anybad = __syncthreads_or(anybad);
    uint32_t run[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
        uint32_t s = incl[j] - mine[j];
        for (int w = 0; w < wv; w++) s += wsum[w][j];
        run[j] = s;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int b = b0 + i;
        if (b < nb) {
            off[b] = run[0];
#pragma unroll
            for (int j = 0; j < 7; j++) boff[j * nb + b] = run[1 + j];
        }
#pragma unroll
        for (int j = 0; j < 8; j++) run[j] += v[i][j];
    }
    if (t < 8) tt[blockIdx.x * 9 + t] = wsum[0][t] + wsum[1][t] + wsum[2][t] + wsum[3][t];
    if (t == 8) tt[blockIdx.x * 9 + 8] = anybad ? 1u : 0u;
}

// The tiles' exclusive prefix sums (tp, 8 per tile) and the frame's header: hdr[0] its symbols,
// hdr[1] its status (1: a magnitude beyond the class table, 2: more symbols than `cap`), hdr[2 +
// band] each band's events. One wave.
__global__ __launch_bounds__(64) void k_rm_tiles(int ntiles, const uint32_t* __restrict__ tt, uint32_t* __restrict__ tp,
                                                 uint32_t* __restrict__ hdr, uint32_t cap) {
    const int j = threadIdx.x;
    if (j >= 9) return;
    uint32_t s = 0;
    for (int k = 0; k < ntiles; k++) {
        const uint32_t x = tt[k * 9 + j];
        if (j < 8) tp[k * 8 + j] = s;
        s = j < 8 ? s + x : (s | x);
    }
    if (j == 0) hdr[0] = s;
    else if (j < 8) hdr[1 + j] = s;
    if (j == 8 || j == 0) {
        // status: from the sums of lanes 0 and 8
        const uint32_t total = __shfl(s, 0, 64), anybad = __shfl(s, 8, 64);
        if (j == 0) hdr[1] = (anybad ? 1u : 0u) | (total > cap ? 2u : 0u);
    }
}

// Per MB and mask word (a workgroup of six waves per 64 MBs: wave k takes word k, lane l MB l of
// the tile), walking the word's 64 positions in order: the significance chains' words (rows of nb
// words, so a chain's 64 MBs are one coalesced load), the EOB chains' words (with the plane's
// first word), the sign and offset-bit bypass symbols, and each non-zero's event in its band's
// list: the position of its first magnitude symbol (<< 4) and its magnitude class.
// Rows of `sigw`: luma position p (0..255), U 256 + p, V 320 + p. Rows of `eobw`: luma context j
// (0..7), U 8 + j, V 14 + j (j 0..5).
__global__ __launch_bounds__(384) void k_rm_walk(const uint8_t* __restrict__ pack, const uint8_t* __restrict__ coded, int nb,
                                                 const uint16_t* __restrict__ cnt, const uint8_t* __restrict__ last,
                                                 const uint8_t* __restrict__ ke, const uint16_t* __restrict__ wst,
                                                 const uint32_t* __restrict__ off, const uint32_t* __restrict__ boff,
                                                 const uint32_t* __restrict__ tp, const uint32_t* __restrict__ hdr,
                                                 uint32_t* __restrict__ sigw, uint32_t* __restrict__ eobw,
                                                 uint32_t* __restrict__ ev, uint16_t* __restrict__ sym) {
    __shared__ int16_t lv[64 * 384];
    const int k = threadIdx.x >> 6, lane = threadIdx.x & 63, b = blockIdx.x * 64 + lane;
    if (hdr[1] || b >= nb) return;
    const int pl = k < 4 ? 0 : k - 3, w0 = pl == 0 ? 0 : 3 + pl;
    const int p0 = pl == 0 ? 64 * k : 0;
    const int row0 = pl == 0 ? p0 : 256 + 64 * (pl - 1);
    const int erow0 = pl == 0 ? 0 : 8 + 6 * (pl - 1), ne = pl == 0 ? 8 : 6;
    if (!coded[b]) {
        for (int j = 0; j < 64; j++) sigw[(size_t)(row0 + j) * nb + b] = 0u;
        if (k == w0)
            for (int j = 0; j < ne; j++) eobw[(size_t)(erow0 + j) * nb + b] = 0u;
        return;
    }
    // every load first (a load issued after a store waits for it), then only stores
    const uint8_t* slot = pack + (size_t)b * PACK_STRIDE;
    const int16_t* vals = (const int16_t*)(slot + 48);
    uint64_t m[6];
#pragma unroll
    for (int i = 0; i < 6; i++) m[i] = ((const uint64_t*)slot)[i];
    const uint32_t* tb = tp + (b >> 10) * 8;   // the MB's tile's prefixes
    const uint32_t o = off[b] + tb[0];
    const uint32_t plstart = pl == 0 ? 0u : pl == 1 ? (uint32_t)cnt[b] : (uint32_t)cnt[b] + cnt[nb + b];
    const int lp = last[pl * nb + b], kk = ke[pl * nb + b];
    uint32_t pos = o + wst[6 * b + k];
    int rank = 0;   // the MB's non-zeros before this word
    for (int i = 0; i < k; i++) rank += __popcll(m[i]);
    const uint64_t mw = m[k];
    // the word's values, to this thread's column of the workgroup's LDS
    const int c = __popcll(mw);
    for (int r0 = 0; r0 < c; r0 += 8) {
        int16_t t8[8];
#pragma unroll
        for (int i = 0; i < 8; i++) t8[i] = r0 + i < c ? vals[rank + r0 + i] : (int16_t)0;
#pragma unroll
        for (int i = 0; i < 8; i++)
            if (r0 + i < c) lv[(r0 + i) * 384 + threadIdx.x] = t8[i];
    }
    // the word's next event in each of its bands (e0..e2): luma word 0 holds bands 0..2 (each
    // starting there), words 1..3 band 3 (after the earlier words' non-zeros); U and V hold bands
    // 4..6, V's after U's
    uint32_t e0, e1 = 0, e2 = 0;
    auto first = [&](int band) { return (uint32_t)band_region(band) * nb + boff[band * nb + b] + tb[1 + band]; };
    if (pl == 0 && k == 0) {
        e0 = first(0);
        e1 = first(1);
        e2 = first(2);
    } else if (pl == 0) {
        e0 = first(3) + (k >= 2 ? __popcll(m[1]) : 0) + (k >= 3 ? __popcll(m[2]) : 0);
    } else {
        const bool v2 = pl == 2;
        e0 = first(4) + (v2 ? (uint32_t)(m[4] & 1ull) : 0u);
        e1 = first(5) + (v2 ? (uint32_t)__popcll(m[4] & 0xfeull) : 0u);
        e2 = first(6) + (v2 ? (uint32_t)__popcll(m[4] & ~0xffull) : 0u);
    }
    if (k == w0) {   // the plane's EOB class value: its context uses and its offset bits
        const int n = pl == 0 ? 9 : 7;
        for (int j = 0; j < ne; j++) eobw[(size_t)(erow0 + j) * nb + b] = chain_word(kk >= j, o + plstart + j, kk > j);
        if (kk >= 2) {
            const uint32_t base = o + plstart + kk + (kk < n - 1 ? 1 : 0);
            const int ov = lp - (1 << (kk - 1));
            for (int i = 0; i <= kk - 2; i++) sym[base + i] = bypass_symbol((ov >> (kk - 2 - i)) & 1);
        }
    }
    int r = 0;
    for (int j = 0; j < 64; j++) {
        const int p = p0 + j;
        const bool act = p <= lp;
        const int bit = (int)((mw >> j) & 1ull);
        sigw[(size_t)(row0 + j) * nb + b] = chain_word(act, pos, bit);
        if (!act) continue;
        if (bit) {
            const int v = lv[r++ * 384 + threadIdx.x];
            const int a = (v < 0 ? -v : v) - 1, km = value_class(a);
            sym[pos + 1] = bypass_symbol(v < 0);
            if (km >= 2) {
                const uint32_t mb = pos + 2 + km + (km < 12 ? 1 : 0);
                const int ov = a - (1 << (km - 1));
                for (int i = 0; i <= km - 2; i++) sym[mb + i] = bypass_symbol((ov >> (km - 2 - i)) & 1);
            }
            // the band within the word: luma word 0 bands 0..2, other luma words band 3, chroma 4..6
            const int sub = pl == 0 ? (k == 0 ? (p == 0 ? 0 : p < 16 ? 1 : 2) : 0) : (p == 0 ? 0 : p < 8 ? 1 : 2);
            const uint32_t idx = sub == 0 ? e0++ : sub == 1 ? e1++ : e2++;
            ev[idx] = ((pos + 2u) << 4) | (uint32_t)km;
            pos += 2 + class_symbols(km, 13);
        } else {
            pos += 1;
        }
    }
}

// One workgroup of eight waves per context chain (418): luma significance (0..255), chroma
// significance (256..319), luma EOB (320..327), chroma EOB (328..333), magnitude per band and
// context (334..417). A chain's uses go 64 per round (row chains: MB l of each 64, or for U then V
// rows MB l / 2 of each 32, plane l & 1; magnitude chains: event l of each 64), in batches of 64
// rounds, through three phases:
//   A. the words are loaded (rounds spread over the waves, loads issued together) and each
//      round's use and bit masks kept;
//   B. one wave finds each round's entry state (see below);
//   C. the rounds' symbols are formed from their entry states and stored.
// Only B is serial, so the phases are pipelined over three LDS buffers: while wave 0 runs B on
// batch k, waves 1..7 run C on batch k - 1 and A on batch k + 1, one barrier per step.
constexpr int kChainWaves = 8, kChainBatch = 64, kChainBufs = 3;
__global__ __launch_bounds__(64 * kChainWaves) void k_rm_chains(int nb, const uint32_t* __restrict__ sigw, const uint32_t* __restrict__ eobw,
                                                                const uint32_t* __restrict__ ev, const uint32_t* __restrict__ hdr,
                                                                uint16_t* __restrict__ sym) {
    __shared__ uint32_t words[kChainBufs][kChainBatch * 64];
    __shared__ unsigned long long am[kChainBufs][kChainBatch][2];
    __shared__ int entry[kChainBufs][kChainBatch][2];
    __shared__ uint32_t rcp[512];
    if (hdr[1]) return;
    const int c = blockIdx.x, t = threadIdx.x, lane = t & 63;
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);   // uniform per wave: its branches and loops stay scalar
    for (int d = t; d < 512; d += 64 * kChainWaves) rcp[d] = d >= 2 ? (uint32_t)((1ull << 32) / (uint64_t)d) + 1u : 0u;
    // the chain's source
    const uint32_t* r0 = nullptr;
    const uint32_t* r1 = nullptr;
    const uint32_t* e = nullptr;
    int i = 0, total = 0, rounds;
    if (c < 256) r0 = sigw + (size_t)c * nb;
    else if (c < 320) { r0 = sigw + (size_t)c * nb; r1 = sigw + (size_t)(c + 64) * nb; }
    else if (c < 328) r0 = eobw + (size_t)(c - 320) * nb;
    else if (c < 334) { r0 = eobw + (size_t)(c - 320) * nb; r1 = eobw + (size_t)(c - 314) * nb; }
    if (c < 334) {
        rounds = r1 ? (nb + 31) / 32 : (nb + 63) / 64;
    } else {
        const int band = (c - 334) / 12;
        i = (c - 334) % 12;
        total = (int)hdr[2 + band];
        e = ev + (size_t)band_region(band) * nb;
        rounds = (total + 63) / 64;
    }
    auto word = [&](int r) -> uint32_t {
        if (e) {
            const int k = r * 64 + lane;
            if (k >= total) return 0u;
            const uint32_t x = e[k];
            const int km = (int)(x & 15u);
            return chain_word(km >= i, (x >> 4) + (uint32_t)i, km > i);
        }
        const int b = r1 ? r * 32 + (lane >> 1) : r * 64 + lane;
        if (b >= nb) return 0u;
        return (r1 && (lane & 1)) ? r1[b] : r0[b];
    };
    // A for batch k by waves w0, w0 + nw, ...: rounds w - w0, w - w0 + nw, ... of the batch
    auto phase_a = [&](int k, int w0, int nw) {
        const int bb = k % kChainBufs, rb = k * kChainBatch, nr = min(kChainBatch, rounds - rb);
        constexpr int kMax = (kChainBatch + kChainWaves - 2) / (kChainWaves - 1);   // rounds per wave with 7 waves
        uint32_t w[kMax];
#pragma unroll
        for (int j = 0; j < kMax; j++) {
            const int rr = wv - w0 + nw * j;
            w[j] = rr < nr ? word(rb + rr) : 0u;
        }
#pragma unroll
        for (int j = 0; j < kMax; j++) {
            const int rr = wv - w0 + nw * j;
            if (rr >= nr) break;
            words[bb][rr * 64 + lane] = w[j];
            const unsigned long long A = __ballot((w[j] & kUsed) != 0u), Om = __ballot((w[j] & (kUsed | 1u)) == (kUsed | 1u));
            if (lane == 0) { am[bb][rr][0] = A; am[bb][rr][1] = Om; }
        }
    };
    const int nbat = (rounds + kChainBatch - 1) / kChainBatch;
    if (nbat > 0) phase_a(0, 0, kChainWaves);
    __syncthreads();
    int z = 1, o = 1;   // wave 0: the context's counts (a fresh context: one zero, one one)
    for (int k = 0; k <= nbat; k++) {
        if (wv == 0) {
            if (k < nbat) {
                const int bb = k % kChainBufs, nr = min(kChainBatch, rounds - k * kChainBatch);
                {
                // B: lane r holds round r's masks and counts. From a known state before round `from`,
                // the entry states of the following rounds are that state plus the counts in between
                // (one prefix sum per batch), up to the first round in which a counter reaches 255;
                // that round's exit comes from round_exit, and the walk resumes after it. From a state
                // that a run of one bit leaves as it is ((255, 1) or (1, 255)), every round up to the
                // first with the other bit enters with it.
                const unsigned long long Al = lane < nr ? am[bb][lane][0] : 0ull, Ol = lane < nr ? am[bb][lane][1] : 0ull;
                const int nzl = __popcll(Al & ~Ol), nol = __popcll(Ol);
                int xz = nzl, xo = nol;
    #pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const int yz = __shfl_up(xz, d, 64), yo = __shfl_up(xo, d, 64);
                    if (lane >= d) { xz += yz; xo += yo; }
                }
                xz -= nzl;   // counts of the rounds before this lane's round
                xo -= nol;
                const int alo = (int)(uint32_t)Al, ahi = (int)(uint32_t)(Al >> 32), olo = (int)(uint32_t)Ol, ohi = (int)(uint32_t)(Ol >> 32);
                auto exit_of = [&](int f) {   // round f's exit from (z, o), its entry state
                    const unsigned long long A = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane(ahi, f) << 32) |
                                                 (uint32_t)__builtin_amdgcn_readlane(alo, f);
                    const unsigned long long Om = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane(ohi, f) << 32) |
                                                  (uint32_t)__builtin_amdgcn_readlane(olo, f);
                    round_exit(z, o, A, Om, lane);
                };
                int ez = 0, eo = 0;
                int from = 0;
                while (from < nr) {
                    const bool open = lane >= from && lane < nr;
                    if ((z == 255 && o == 1) || (z == 1 && o == 255)) {
                        const unsigned long long brk = __ballot(open && (z == 255 ? nol : nzl) > 0);
                        const int f = brk ? __ffsll(brk) - 1 : nr;
                        if (open && lane <= f) { ez = z; eo = o; }
                        if (f >= nr) break;
                        exit_of(f);
                        from = f + 1;
                        continue;
                    }
                    const int oz = __builtin_amdgcn_readlane(xz, from), oo = __builtin_amdgcn_readlane(xo, from);
                    const int zr = z + xz - oz, orr = o + xo - oo;
                    const unsigned long long brk = __ballot(open && (zr + nzl > 255 || orr + nol > 255));
                    const int f = brk ? __ffsll(brk) - 1 : nr;
                    if (open && lane <= f) { ez = zr; eo = orr; }
                    if (f >= nr) {
                        z = __builtin_amdgcn_readlane(zr + nzl, nr - 1);
                        o = __builtin_amdgcn_readlane(orr + nol, nr - 1);
                        break;
                    }
                    z = __builtin_amdgcn_readlane(zr, f);
                    o = __builtin_amdgcn_readlane(orr, f);
                    exit_of(f);
                    from = f + 1;
                }
                if (lane < nr) { entry[bb][lane][0] = ez; entry[bb][lane][1] = eo; }
            }
            }
        } else {
            if (k >= 1) {   // C for batch k - 1
                const int bb = (k - 1) % kChainBufs, nr = min(kChainBatch, rounds - (k - 1) * kChainBatch);
                for (int rr = wv - 1; rr < nr; rr += kChainWaves - 1) {
                    const uint32_t x = words[bb][rr * 64 + lane];
                    int zz = entry[bb][rr][0], oo = entry[bb][rr][1];
                    uint16_t out = 0;
                    chain_round(zz, oo, (x & kUsed) != 0u, (int)(x & 1u), out, lane, rcp);
                    if (x & kUsed) sym[(x >> 1) & 0x3fffffffu] = out;
                }
            }
            if (k + 1 < nbat) phase_a(k + 1, 1, kChainWaves - 1);
        }
        __syncthreads();
    }
}

// The frame's header words (symbol count and status) to mapped host memory; when the status is
// set, the MBs' records too (nb * PACK_STRIDE bytes, in 16-byte stores), which the host then codes
// itself. The symbols stay on the device: the host stage copies them with the copy engine.
__global__ __launch_bounds__(256) void k_rm_push(const uint32_t* __restrict__ hdr, const uint8_t* __restrict__ pack, int nb,
                                                 uint32_t* __restrict__ hhdr, uint8_t* __restrict__ hpack) {
    const uint32_t n = hdr[0], status = hdr[1];
    if (blockIdx.x == 0 && threadIdx.x == 0) { hhdr[0] = n; hhdr[1] = status; }
    if (!status) return;
    const size_t n16 = (size_t)nb * PACK_STRIDE / 16;
    const uint4* s = (const uint4*)pack;
    uint4* d = (uint4*)hpack;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (size_t)gridDim.x * blockDim.x) d[i] = s[i];
}
}  // namespace

namespace rivk {
constexpr int kChains = 256 + 64 + 8 + 6 + 7 * 12;

hipError_t res_model(hipStream_t st, const ResModelBufs& B, uint32_t* host_hdr, uint8_t* host_pack) {
    const int nb = B.nb;
    if (nb <= 0 || B.cap >= (1u << 28)) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_rm_count, dim3((nb + 3) / 4), dim3(256), 0, st, B.pack, B.coded, nb, B.cnt, B.last, B.ke, B.bcnt, B.wst, B.bad);
    const int ntiles = (nb + 1023) / 1024;
    hipLaunchKernelGGL(k_rm_scan, dim3(ntiles), dim3(256), 0, st, nb, B.cnt, B.bcnt, B.bad, B.off, B.boff, B.tt);
    hipLaunchKernelGGL(k_rm_tiles, dim3(1), dim3(64), 0, st, ntiles, B.tt, B.tp, B.hdr, B.cap);
    hipLaunchKernelGGL(k_rm_walk, dim3((nb + 63) / 64), dim3(384), 0, st, B.pack, B.coded, nb, B.cnt, B.last, B.ke, B.wst, B.off, B.boff,
                       B.tp, B.hdr, B.sigw, B.eobw, B.ev, B.sym);
    hipLaunchKernelGGL(k_rm_chains, dim3(kChains), dim3(64 * kChainWaves), 0, st, nb, B.sigw, B.eobw, B.ev, B.hdr, B.sym);
    hipLaunchKernelGGL(k_rm_push, dim3(256), dim3(256), 0, st, B.hdr, B.pack, nb, host_hdr, host_pack);
    return hipGetLastError();
}
}  // namespace rivk
