// gfx950 kernels of the RIV backend: colour conversion, motion search (relaxation passes) and
// mode decision, motion compensation, adaptive quantisation, residual transforms, deblock.
// Each kernel states the reference function it reproduces bit-exactly (paths relative
// to the reference tree, under reitero_video/).
#include <hip/hip_ext.h>
#include "riv_dev.hpp"
#include "riv_kernels.hpp"

using namespace rivdev;
using riv::BlockRec;
using riv::SearchParams;
using riv::SearchStream;

namespace rivdev {
__constant__ int32_t c_B16[16][16];
__constant__ int32_t c_B8[8][8];
}

// ============================================================================================
// Colour conversion
// ============================================================================================
// pad_rgb24 (reitero_encode/src/encoder.rs:1049-1096: replicate the last column, then the last
// row) fused with the project's BT.709 limited-range RGB24 -> YUV420 (14-bit fixed point, 2x2
// box chroma). One thread per chroma sample; writes the 2x2 luma quad and one U, one V.
__global__ void k_rgb_to_yuv420(const uint8_t* __restrict__ rgb, int dw, int dh, int sw, int sh,
                                uint8_t* __restrict__ Y, uint8_t* __restrict__ U, uint8_t* __restrict__ V) {
    int cw = sw >> 1;
    int cx = blockIdx.x * blockDim.x + threadIdx.x, cy = blockIdx.y;
    if (cx >= cw) return;
    int R = 0, G = 0, B = 0;
#pragma unroll
    for (int j = 0; j < 2; j++) {
        int y = min(2 * cy + j, dh - 1);
#pragma unroll
        for (int i = 0; i < 2; i++) {
            int x = min(2 * cx + i, dw - 1);
            const uint8_t* p = rgb + ((size_t)y * dw + x) * 3;
            int r = p[0], g = p[1], b = p[2];
            int yy = (2991 * r + 10064 * g + 1016 * b + (16 << 14) + (1 << 13)) >> 14;
            Y[(size_t)(2 * cy + j) * sw + 2 * cx + i] = (uint8_t)dclamp(yy, 0, 255);
            R += r; G += g; B += b;
        }
    }
    int u = (-1649 * R - 5547 * G + 7196 * B + (128 << 16) + (1 << 15)) >> 16;
    int v = (7196 * R - 6536 * G - 660 * B + (128 << 16) + (1 << 15)) >> 16;
    U[(size_t)cy * cw + cx] = (uint8_t)dclamp(u, 0, 255);
    V[(size_t)cy * cw + cx] = (uint8_t)dclamp(v, 0, 255);
}

// YUV420 -> RGB24 (BT.709 limited, nearest chroma) fused with crop_rgb24
// (reitero_decode/src/decoder.rs:18-32): only the display area is produced.
__global__ void k_yuv420_to_rgb(const uint8_t* __restrict__ Y, const uint8_t* __restrict__ U, const uint8_t* __restrict__ V,
                                int sw, int dw, int dh, uint8_t* __restrict__ rgb) {
    int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= dw || y >= dh) return;
    int cw = sw >> 1;
    int yy = Y[(size_t)y * sw + x] - 16, uu = U[(size_t)(y >> 1) * cw + (x >> 1)] - 128, vv = V[(size_t)(y >> 1) * cw + (x >> 1)] - 128;
    int r = (19077 * yy + 29372 * vv + (1 << 13)) >> 14;
    int g = (19077 * yy - 3494 * uu - 8731 * vv + (1 << 13)) >> 14;
    int b = (19077 * yy + 34610 * uu + (1 << 13)) >> 14;
    uint8_t* p = rgb + ((size_t)y * dw + x) * 3;
    p[0] = (uint8_t)dclamp(r, 0, 255); p[1] = (uint8_t)dclamp(g, 0, 255); p[2] = (uint8_t)dclamp(b, 0, 255);
}

// ============================================================================================
// Integer SAD table
// ============================================================================================
// An MB whose whole search area [x0-sr, x0+15+sr] x [y0-sr, y0+15+sr] lies inside the frame
// clamps no 4x4 origin and no pixel, so its integer SAD at (dx, dy) equals the half-pel SAD at
// (2dx, 2dy); only the other ("edge") MBs get an integer table of their own.
__host__ __device__ inline bool mb_interior(int x0, int y0, int W, int H, int sr) {
    return x0 >= sr && y0 >= sr && x0 + 16 + sr <= W && y0 + 16 + sr <= H;
}

// Every integer candidate a block's search can visit lies in [-sr, sr]^2, and its SAD does not
// depend on the search path, so all (2sr+1)^2 of them are computed up front, fully parallel:
// one workgroup per MB, one position per thread. SAD semantics of satd_scalar.rs:4-99: each 4x4
// sub-block's reference origin is clamped to [0,W-4] x [0,H-4] as a unit, no per-pixel clamp.
// The reference window (frame rows and columns the clamped origins can touch) is staged in LDS
// as four byte-shifted dword copies, so every 4-pixel reference row is one aligned LDS read.
// Entry (dy+sr)*(2sr+1) + (dx+sr) of block b sits at out[b*stride + ...], stride a multiple of 8.
__global__ __launch_bounds__(256) void k_sad_table(const uint8_t* __restrict__ cur, const uint8_t* __restrict__ ref, int W, int H, int bw,
                                                   int sr, int ndwt, uint16_t* __restrict__ out, int stride) {
    extern __shared__ __align__(16) uint32_t sm[];
    const int b = blockIdx.x, bx = b % bw, by = b / bw, x0 = bx * 16, y0 = by * 16, t = threadIdx.x;
    if (mb_interior(x0, y0, W, H, sr)) return;   // its SADs are the half-pel table's even entries
    const int D = 2 * sr + 1, P = D * D, MAXR = 16 + 2 * sr, CS = MAXR * ndwt;
    const int wx = max(0, x0 - sr) & ~3, wy = max(0, y0 - sr);
    const int NR = min(H, y0 + 16 + sr) - wy;
    const int wxe = min(W, (x0 + 16 + sr + 3) & ~3);   // exclusive, a multiple of 4 (W is a multiple of 16)
    const int nd = (wxe - wx) >> 2;
    uint32_t* curs = sm;        // 64 dwords: the current block
    uint32_t* cp = sm + 64;     // copy s, row r, dword i = bytes wx+4i+s .. +3 of reference row wy+r
    if (t < 64) curs[t] = *(const uint32_t*)(cur + (size_t)(y0 + (t >> 2)) * W + x0 + 4 * (t & 3));
    for (int e = t; e < NR * nd; e += 256) {
        const int r = e / nd, i = e - r * nd;
        const uint8_t* row = ref + (size_t)(wy + r) * W + wx;
        const uint32_t d0 = *(const uint32_t*)(row + 4 * i);
        const uint32_t d1 = i + 1 < nd ? *(const uint32_t*)(row + 4 * i + 4) : 0u;
        uint32_t* q = cp + r * ndwt + i;
        q[0] = d0;
        q[CS] = __builtin_amdgcn_alignbyte(d1, d0, 1);
        q[2 * CS] = __builtin_amdgcn_alignbyte(d1, d0, 2);
        q[3 * CS] = __builtin_amdgcn_alignbyte(d1, d0, 3);
    }
    __syncthreads();
    uint32_t c[64];
#pragma unroll
    for (int k = 0; k < 64; k++) c[k] = curs[k];
    for (int p = t; p < P; p += 256) {
        const int dy = p / D - sr, dx = p % D - sr;
        int base[4];
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int ox = dclamp(x0 + 4 * g + dx, 0, W - 4) - wx;
            base[g] = (ox & 3) * CS + (ox >> 2);
        }
        uint32_t acc = 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int oy = dclamp(y0 + 4 * j + dy, 0, H - 4) - wy;
#pragma unroll
            for (int i = 0; i < 4; i++) {
#pragma unroll
                for (int g = 0; g < 4; g++) acc = sad4(c[(4 * j + i) * 4 + g], cp[base[g] + (oy + i) * ndwt], acc);
            }
        }
        out[(size_t)b * stride + p] = (uint16_t)acc;
    }
}

// ============================================================================================
// Half-pel SAD table
// ============================================================================================
// refine_halfpel (motion/scalar.rs:759-839) scores half-pel positions (hx, hy) in
// [-2sr-1, 2sr+1]^2 around the integer result with floor interpolation and per-pixel clamping
// (satd_scalar.rs:113-166). Per-pixel clamping equals plain interpolation on the replicate-
// padded reference window, so the window is staged once per MB as four interpolated planes
// (integer, x-half, y-half, xy-half); position (hx, hy) is then a plain 16x16 SAD against plane
// (hx&1, hy&1) at integer offset (hx>>1, hy>>1). A thread scores 4 x-offsets x K y-offsets of one
// plane: v_qsad_pk_u16_u8 gives the SADs of 4 consecutive byte offsets per instruction, and each
// reference row is read once for the K y-offsets it serves. Entry (hy+2sr+1)*(4sr+3) + (hx+2sr+1)
// of block b sits at out[b*stride + ...], stride a multiple of 8.
template <int K>
__global__ __launch_bounds__(128) void k_hp_table(const uint8_t* __restrict__ cur, const uint8_t* __restrict__ ref, int W, int H, int bw,
                                                  int sr, int ndp, uint16_t* __restrict__ out, int stride) {
    extern __shared__ __align__(16) uint32_t sm[];
    const int b = blockIdx.x, bx = b % bw, by = b / bw, x0 = bx * 16, y0 = by * 16, t = threadIdx.x;
    const int WS = 16 + 2 * sr + 2, NRP = WS + K, NRA = NRP + 1, NDA = ndp + 1, PS = NRP * ndp, D2 = 4 * sr + 3, HR = 2 * sr + 1;
    const int wx0 = x0 - sr - 1, wxa = wx0 & ~3, xo = wx0 - wxa, wy0 = y0 - sr - 1;
    uint32_t* curs = sm;            // 64 dwords: the current block
    uint32_t* A = sm + 64;          // NRA x NDA: replicate-padded window, dword-aligned span from wxa
    uint32_t* P = A + NRA * NDA;    // 4 planes of NRP x ndp dwords
    if (t < 64) curs[t] = *(const uint32_t*)(cur + (size_t)(y0 + (t >> 2)) * W + x0 + 4 * (t & 3));
    for (int e = t; e < NRA * NDA; e += blockDim.x) {
        const int r = e / NDA, k = e - r * NDA, gx = wxa + 4 * k;
        const uint8_t* row = ref + (size_t)dclamp(wy0 + r, 0, H - 1) * W;
        // W is a multiple of 16: an aligned dword is wholly inside or wholly outside the row
        A[e] = (gx >= 0 && gx < W) ? *(const uint32_t*)(row + gx) : 0x01010101u * (gx < 0 ? row[0] : row[W - 1]);
    }
    __syncthreads();
    for (int e = t; e < NRP * ndp; e += blockDim.x) {
        const int r = e / ndp, k = e - r * ndp;
        const uint32_t a0 = A[r * NDA + k], a1 = __builtin_amdgcn_alignbyte(A[r * NDA + k + 1], a0, 1);
        const uint32_t b0 = A[(r + 1) * NDA + k], b1 = __builtin_amdgcn_alignbyte(A[(r + 1) * NDA + k + 1], b0, 1);
        P[e] = a0;
        P[PS + e] = avg2f(a0, a1);
        P[2 * PS + e] = avg2f(a0, b0);
        P[3 * PS + e] = avg4f(a0, a1, b0, b1);
    }
    __syncthreads();
    uint32_t c[64];
#pragma unroll
    for (int q = 0; q < 64; q++) c[q] = curs[q];
    // x-offset groups: parity 0 has qx in [-sr, sr], parity 1 has qx in [-sr-1, sr] (hx = 2qx + parity)
    const int gx0 = (2 * sr + 1 + 3) / 4, gx1 = (2 * sr + 2 + 3) / 4, cy0 = (2 * sr + 1 + K - 1) / K, cy1 = (2 * sr + 2 + K - 1) / K;
    const int ntask = (gx0 + gx1) * (cy0 + cy1);
    for (int task = t; task < ntask; task += blockDim.x) {
        const int xi = task % (gx0 + gx1), yi = task / (gx0 + gx1);
        const int px = xi >= gx0, py = yi >= cy0;
        const int qx0 = (px ? -sr - 1 + 4 * (xi - gx0) : -sr + 4 * xi);
        const int qy0 = (py ? -sr - 1 + K * (yi - cy0) : -sr + K * yi);
        const uint32_t* pl = P + (px + 2 * py) * PS;
        const int u0 = xo + qx0 + sr + 1, m = u0 >> 2, sh = u0 & 3, r0 = qy0 + sr + 1;
        unsigned long long acc[K];
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = 0;
#pragma unroll
        for (int tt = 0; tt < K + 15; tt++) {
            const uint32_t* rw = pl + (r0 + tt) * ndp + m;
            uint32_t d[6];
#pragma unroll
            for (int i = 0; i < 6; i++) d[i] = rw[i];
            // group g's 8 reference bytes, built in place as a 64-bit operand
            unsigned long long pr[4];
#pragma unroll
            for (int g = 0; g < 4; g++)
                pr[g] = (unsigned long long)__builtin_amdgcn_alignbyte(d[g + 1], d[g], sh) |
                        ((unsigned long long)__builtin_amdgcn_alignbyte(d[g + 2], d[g + 1], sh) << 32);
            // consecutive instructions feed different accumulators, so they do not wait on each other
#pragma unroll
            for (int g = 0; g < 4; g++) {
#pragma unroll
                for (int k = 0; k < K; k++) {
                    const int j = tt - k;
                    if (j >= 0 && j < 16) acc[k] = __builtin_amdgcn_qsad_pk_u16_u8(pr[g], c[j * 4 + g], acc[k]);
                }
            }
        }
        uint16_t* ob = out + (size_t)b * stride;
#pragma unroll
        for (int k = 0; k < K; k++) {
            const int hy = 2 * (qy0 + k) + py;
            if (hy <= HR) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int hx = 2 * (qx0 + q) + px;
                    if (hx <= HR) ob[(hy + HR) * D2 + hx + HR] = (uint16_t)(acc[k] >> (16 * q));
                }
            }
        }
    }
}

// ============================================================================================
// Motion search: shared helpers (MV-cost model, lane argmin, lane selection)
// ============================================================================================
namespace {

__device__ __forceinline__ int median3(int a, int b, int c) {   // motion/scalar.rs:24-43
    return max(min(a, b), min(max(a, b), c));
}

// The NEW-vector term of estimate_mv_cost for component distances (a, b) to a predictor:
// (lam2 * ((6 + bits[a]) + bits[b])) as i64, evaluated in f64 in the reference's order
// (motion/scalar.rs:453-544). A search tabulates it once per launch for a, b < 2*sr + 4, which
// covers every candidate (|d| <= sr + 2) against every predictor (|d| <= sr + 1).
__device__ __forceinline__ int new_mv_cost(double lam2, const double* bits, int a, int b) {
    return (int)__dmul_rn(lam2, __dadd_rn(__dadd_rn(6.0, bits[a]), bits[b]));
}

struct CostCtx {
    int nx, ny, rx, ry, mx, my;
    const int* tab;   // tab[a * td + b] = new_mv_cost(lam2, bits, a, b)
    int td;
    int c_zero, c_pred, lam_zero;
    // estimate_mv_cost (motion/scalar.rs:453-544): the min over the applicable terms. The host
    // keeps lam2 * 42 below 2^26, so every term truncates exactly in 32 bits and SAD + cost
    // leaves room for a lane index in the low bits of a search key.
    __device__ __forceinline__ int cost(int dx, int dy) const {
        const bool pm = ((dx == nx) & (dy == ny)) | ((dx == rx) & (dy == ry)) | ((dx == mx) & (dy == my));
        int best = pm ? c_pred : 0x3ffffff;
        best = ((dx == 0) & (dy == 0)) ? min(best, c_zero) : best;
        const int t0 = tab[__mul24(abs(dx - nx), td) + abs(dy - ny)];
        const int t1 = tab[__mul24(abs(dx - rx), td) + abs(dy - ry)];
        return lam_zero ? 0 : min(best, min(t0, t1));
    }
};

// Minimum of v over lanes 0..7, returned wave-uniform; the other lanes of each row of 16
// must hold values that do not win. Used as an argmin: keys carry the lane index in their low bits.
__device__ __forceinline__ int lane_min8(int v) {
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true));    // quad_perm [1,0,3,2]
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true));    // quad_perm [2,3,0,1]
    v = min(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, true));   // row_half_mirror
    return __builtin_amdgcn_readlane(v, 0);
}

__device__ __forceinline__ uint32_t sgpr(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
// per-lane selection among wave-uniform values: lane k of the result is vals[k] for k < N, else
// vals[0] (a chain of lane compares and selects, no branches)
template <int N>
__device__ __forceinline__ uint32_t lane_pick(int lane, const uint32_t (&vals)[N]) {
    uint32_t r = vals[0] & (0u - (uint32_t)(lane == 0 || lane >= N));
#pragma unroll
    for (int k = 1; k < N; k++) r |= vals[k] & (0u - (uint32_t)(lane == k));   // masks, not a table lookup
    return r;
}
__device__ __forceinline__ int mvx(uint32_t w) { return (int)(int8_t)(w & 0xff); }
__device__ __forceinline__ int mvy(uint32_t w) { return (int)(int8_t)((w >> 8) & 0xff); }
__device__ __forceinline__ int mv_class(int mag) {   // mv_class_from_magnitude (mv_rans.rs:81-87)
    if (mag == 0) return 0;
    return min(32 - __clz(mag), 10);
}

}  // namespace

// ============================================================================================
// MV mode decision
// ============================================================================================
// Mode decision on the reconstructed prefix (encoder.rs:447-733), per MB, run by the motion
// search's epilogue on the final field. It needs the reconstructed MVs of the left, top,
// top-right and top-left MBs and of the co-located MB of the previous P-frame (derive_mv_predictors_with_stats / gather_mv_candidates,
// mv_predictor.rs:135-272). A reconstructed MV equals the search MV: for a matched mode it is
// the matched candidate, which equals the search MV; for NEW it is base + clamp(search - base),
// and with every MV inside [-sr-1, sr] (sr <= 63) the difference never reaches the clamp. So
// the whole frame's decisions follow from the search field in parallel. The identity is still
// checked for every NEW block: a violation sets the stream's error word *serr to 2.
__device__ __forceinline__ uint4 mode_record(uint32_t wS, int best_sad, bool has_l, bool has_t, bool has_tr, bool has_tl, bool has_tp,
                                             uint32_t wL, uint32_t wT, uint32_t wTR, uint32_t wTL, uint32_t wP,
                                             const uint32_t* rdo_bits, int rdo_max, int* serr) {
    // deduplicated (dx,dy,sx,sy) of left, top, top-right, top-left, temporal; the first two are
    // Nearest / Near, padded with zero. Every MV word is canonical (sx, sy in {0,1}), so equal
    // half-pel positions are equal words.
    uint32_t c0 = 0, c1 = 0; int ncand = 0;
#define RIV_ADDC(present, q)                                                         \
    {                                                                        \
        const uint32_t q_ = (q);                                             \
        const bool add_ = (present) && !(ncand >= 1 && c0 == q_) && !(ncand >= 2 && c1 == q_); \
        c1 = (add_ && ncand == 1) ? q_ : c1;                                 \
        c0 = (add_ && ncand == 0) ? q_ : c0;                                 \
        ncand += add_ ? 1 : 0;                                               \
    }
    RIV_ADDC(has_l, wL) RIV_ADDC(has_t, wT) RIV_ADDC(has_tr, wTR) RIV_ADDC(has_tl, wTL) RIV_ADDC(has_tp, wP)
#undef RIV_ADDC
    const uint32_t nearest = ncand > 0 ? c0 : 0u, nearp = ncand > 1 ? c1 : 0u, temporal = has_tp ? wP : 0u;
    const int S0 = mvx(wS), S1 = mvy(wS), S2 = (wS >> 16) & 0xff, S3 = wS >> 24;
    uint8_t mode;
    if (wS == 0u) mode = riv::M_ZERO;
    else if (wS == nearest) mode = riv::M_NEAREST;
    else if (wS == nearp) mode = riv::M_NEAR;
    else if (has_tr && wS == wTR) mode = riv::M_TOPRIGHT;
    else if (has_tl && wS == wTL) mode = riv::M_TOPLEFT;
    else if (wS == temporal) mode = riv::M_TEMPORAL;
    else mode = riv::M_NEW;
    uint8_t chosen = 0; int rdx = 0, rdy = 0;
    if (mode == riv::M_NEW) {   // NEW base by estimated bits (encoder.rs:567-661, :128-159): the first cheapest
        const uint32_t bases[5] = {nearest, nearp, wTR, wTL, temporal};
        const bool have[5] = {true, true, has_tr, has_tl, true};
        unsigned best = 0xffffffffu;
        int bbx = mvx(nearest), bby = mvy(nearest);
#pragma unroll
        for (int k = 0; k < 5; k++) {
            if (have[k]) {
                const int kx = mvx(bases[k]), ky = mvy(bases[k]);
                const int cxs = mv_class(abs(dclamp(S0 - kx, -128, 127))), cys = mv_class(abs(dclamp(S1 - ky, -128, 127)));
                const unsigned cost = (k < 4 ? k + 1 : 4) + (cxs == 0 ? 4u : 4u + cxs) + (cys == 0 ? 4u : 4u + cys);
                if (cost < best) { best = cost; chosen = (uint8_t)k; bbx = kx; bby = ky; }
            }
        }
        rdx = dclamp(S0 - bbx, -128, 127);
        rdy = dclamp(S1 - bby, -128, 127);
        if (dclamp(bbx + rdx, -128, 127) != S0 || dclamp(bby + rdy, -128, 127) != S1)
            *serr = 2;   // the stream's error word (mapped host memory: a plain store)
    }
    // RdoContext::decide via the host-built table (rdo.rs:140-178)
    uint32_t rdo = 0;
    if (S2 == 0 && S3 == 0 && best_sad <= rdo_max) rdo = (rdo_bits[best_sad >> 5] >> (best_sad & 31)) & 1u;
    const uint32_t w2 = (uint32_t)mode | ((uint32_t)chosen << 8) | ((uint32_t)(uint8_t)rdx << 16) | ((uint32_t)(uint8_t)rdy << 24);
    return make_uint4(wS, wS, w2, (uint32_t)best_sad | (rdo << 16));
}

// ============================================================================================
// Motion search by relaxation passes
// ============================================================================================
// A block's search result is a function of constant data and of the integer (dx, dy) of its
// left, top-left, top and top-right neighbours' results (ME seeds and MV-cost predictors,
// motion/scalar.rs:56-156). Every one of those neighbours comes earlier in raster order, so the
// dependencies form a DAG and the reference's raster-order result is the unique fixed point of
// x = F(x). Passes of F, each computed from the previous pass's field, reach it exactly from any
// start field: a block on anti-diagonal t = bx + 2*by is final after pass t + 1, so 2*(bh-1) + bw
// passes always suffice, and a pass that changes no integer MV proves that the fixed point has
// been reached. A block is re-evaluated only when the integer part of a neighbour's word differs
// from the one its last evaluation used. On camera-like frames the field settles in 10-20 passes,
// against the 2*(bh-1) + bw dependent steps of a wavefront.
//
// One persistent launch, every workgroup resident at once (the host sizes the grid to one
// workgroup per CU). Each workgroup owns a contiguous range of blocks (in XCD-band order); wave w
// of it owns blocks w + k * waves of the range, lane k holding block k's current word, so that a
// pass's few changed blocks land on different waves. Passes end at a grid barrier on a per-pass
// arrival counter that also counts the workgroups whose blocks changed. The field words and the
// counters are read and written with relaxed agent-scope atomics (coherent across the chip
// without cache maintenance); each wave waits for its own stores before arriving.
//
// Each wave keeps its blocks' replicate-padded reference windows, and the blocks' own 16 rows, in
// LDS slots for the whole launch (the reference frame does not change during the search), so a
// re-evaluation reads no global memory; when a wave owns more blocks than it has slots, the
// blocks past the resident ones share one slot that is reloaded for each evaluation.
namespace {

// Candidate SADs are computed one candidate per 16-lane row: lane j of a row holds row j of the
// block (16 pixels, 4 dwords) and sums its 16 absolute differences; a 4-step DPP reduction
// inside the row gives the candidate's SAD in all 16 lanes. A wave thus scores 4 candidates per
// round with no cross-row traffic.
struct JWin {
    const uint32_t* wp;   // the window: NR rows of ND dwords, dword i of a row = bytes 4i .. 4i+3
    int ND;
    int W, H, x0, y0, wxa, wy0;
    uint4 cq;             // row j = lane & 15 of the current block
    int j;
    // Block row j at offset (d, e) reads the window from byte cb + d of row j + sr + 1 + e, i.e.
    // from dword ((cb+d) >> 2) + e * ND + rowc, shifted by (cb+d) & 3 bytes.
    int cb;               // x0 - wxa (>= sr + 1)
    int rowc;             // (j + sr + 1) * ND
};

// 16 bytes from byte s (0..3) of q[0]: five dword reads, four byte-aligns.
__device__ __forceinline__ void jw_read16(const uint32_t* q, int s, uint32_t r[4]) {
    uint32_t d[5];
#pragma unroll
    for (int k = 0; k < 5; k++) d[k] = q[k];
#pragma unroll
    for (int k = 0; k < 4; k++) r[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], (uint32_t)s);
}
// The same 16 bytes (a) and the 16 one byte further on (b), from one set of five dword reads.
__device__ __forceinline__ void jw_read16x2(const uint32_t* q, int s, uint32_t a[4], uint32_t b[4]) {
    uint32_t d[5];
#pragma unroll
    for (int k = 0; k < 5; k++) d[k] = q[k];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        a[k] = __builtin_amdgcn_alignbyte(d[k + 1], d[k], (uint32_t)s);
        b[k] = s == 3 ? d[k + 1] : __builtin_amdgcn_alignbyte(d[k + 1], d[k], (uint32_t)(s + 1));
    }
}

__device__ __forceinline__ uint32_t row_sum(uint32_t v) {   // sum over the lane's 16-lane row, in every lane of it
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);    // quad_perm [1,0,3,2]
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);    // quad_perm [2,3,0,1]
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, true);   // row_half_mirror
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, true);   // row_mirror
    return v;
}

// Row SAD at integer MV (dx, dy): each 4x4 sub-block's reference origin is clamped to
// [0,W-4] x [0,H-4] as a whole (sad_block_int_limit, satd_scalar.rs:4-99); inside the frame's
// interior no origin is clamped and the 4 dwords of a row are consecutive.
template <bool INTERIOR>
__device__ __forceinline__ uint32_t jw_row_int(const JWin& w, int dx, int dy) {
    uint32_t r[4];
    if (INTERIOR) {
        const int c = w.cb + dx;
        jw_read16(w.wp + (c >> 2) + __mul24(dy, w.ND) + w.rowc, c & 3, r);
    } else {
        const int oy = dclamp(w.y0 + 4 * (w.j >> 2) + dy, 0, w.H - 4) + (w.j & 3) - w.wy0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int ox = dclamp(w.x0 + 4 * k + dx, 0, w.W - 4) - w.wxa;
            const uint32_t* q = w.wp + __mul24(oy, w.ND) + (ox >> 2);
            r[k] = __builtin_amdgcn_alignbyte(q[1], q[0], (uint32_t)(ox & 3));
        }
    }
    uint32_t acc = __builtin_amdgcn_sad_u8(w.cq.x, r[0], 0u);
    acc = __builtin_amdgcn_sad_u8(w.cq.y, r[1], acc);
    acc = __builtin_amdgcn_sad_u8(w.cq.z, r[2], acc);
    acc = __builtin_amdgcn_sad_u8(w.cq.w, r[3], acc);
    return row_sum(acc);
}

// Row SAD at half-pel MV (hx, hy): floor interpolation with every sample position clamped to the
// frame (sample_luma_halfpel, satd_scalar.rs:113-166), which is plain interpolation on the
// replicate-padded window. TYPE: 0 integer position, 1 one half-pel component (x when hx is odd,
// else y), 2 both components half-pel.
template <int TYPE>
__device__ __forceinline__ uint32_t jw_row_hp(const JWin& w, int hx, int hy) {
    const int c = w.cb + (hx >> 1);
    const uint32_t* q = w.wp + (c >> 2) + __mul24(hy >> 1, w.ND) + w.rowc;
    uint32_t v[4];
    if (TYPE == 0) {
        jw_read16(q, c & 3, v);
    } else if (TYPE == 1) {
        // the second sample: one byte right (hx odd) or one row below
        const int c2 = c + (hx & 1);
        const uint32_t* q2 = w.wp + (c2 >> 2) + __mul24((hy >> 1) + ((hx & 1) ^ 1), w.ND) + w.rowc;
        uint32_t a[4], b[4];
        jw_read16(q, c & 3, a);
        jw_read16(q2, c2 & 3, b);
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = avg2f(a[k], b[k]);
    } else {
        uint32_t a[4], b[4], e[4], f[4];
        jw_read16x2(q, c & 3, a, b);
        jw_read16x2(q + w.ND, c & 3, e, f);
#pragma unroll
        for (int k = 0; k < 4; k++) v[k] = avg4f(a[k], b[k], e[k], f[k]);
    }
    uint32_t acc = __builtin_amdgcn_sad_u8(w.cq.x, v[0], 0u);
    acc = __builtin_amdgcn_sad_u8(w.cq.y, v[1], acc);
    acc = __builtin_amdgcn_sad_u8(w.cq.z, v[2], acc);
    acc = __builtin_amdgcn_sad_u8(w.cq.w, v[3], acc);
    return row_sum(acc);
}

__device__ __forceinline__ int lane_bcast(int v, int src) { return __builtin_amdgcn_ds_bpermute(src << 2, v); }

// SADs of the integer candidates held by lanes 0..N-1 (cx, cy); lane k receives candidate k's SAD.
// Round R puts candidate 4R + r on row r (rows past N repeat candidate 0, which is in range).
template <int N, bool INTERIOR>
__device__ __forceinline__ int jw_set_int(const JWin& w, int lane, int cx, int cy) {
    const int r = lane >> 4;
    int sv = 0;
#pragma unroll
    for (int R = 0; R < (N + 3) / 4; R++) {
        const int k = 4 * R + r < N ? 4 * R + r : 0;
        const int s = (int)jw_row_int<INTERIOR>(w, lane_bcast(cx, k), lane_bcast(cy, k));
        const int got = lane_bcast(s, (lane - 4 * R) << 4);   // row lane - 4R holds candidate `lane`
        sv = (lane >= 4 * R && lane < 4 * R + 4) ? got : sv;
    }
    return sv;
}

// The window of MB (x0, y0): rows y0-sr-1 .. y0+16+sr and columns from the dword-aligned
// wxa <= x0-sr-1, replicate-padded outside the frame (W is a multiple of 16, so an aligned
// dword is wholly inside or wholly outside a row), written to wp as NR rows of ND dwords; the
// block's own 16 rows follow at wp + riv::search_window_dwords(sr).
// A lane covers rows jr, jr + rstep, ... of dword column ji; the loads of JU rows are all issued
// before any of them is stored, so a window costs about one memory latency.
__device__ __forceinline__ void jw_load(const SearchStream& S, int sr, uint32_t* wp, int ND, int NR, int x0, int y0, int lane) {
    constexpr int JU = 9;
    if (lane < 16) {
        const uint4 v = *(const uint4*)(S.cur + (size_t)(y0 + lane) * S.W + x0);
        *(uint4*)(wp + riv::search_window_dwords(sr) + 4 * lane) = v;
    }
    const int wxa = (x0 - sr - 1) & ~3, wy0 = y0 - sr - 1;
    const int rstep = 64 / ND, jr = lane / ND, ji = lane - jr * ND;
    if (jr >= rstep) return;
    const int g0 = wxa + 4 * ji;
    const bool in0 = g0 >= 0 && g0 < S.W;
    for (int rb = jr; rb < NR; rb += JU * rstep) {
        uint32_t d0[JU];
#pragma unroll
        for (int u = 0; u < JU; u++) {
            const int r = min(rb + u * rstep, NR - 1);
            const uint8_t* row = S.ref + (size_t)dclamp(wy0 + r, 0, S.H - 1) * S.W;
            d0[u] = in0 ? *(const uint32_t*)(row + g0) : 0x01010101u * (g0 < 0 ? row[0] : row[S.W - 1]);
        }
#pragma unroll
        for (int u = 0; u < JU; u++) {
            const int r = rb + u * rstep;
            if (r < NR) wp[r * ND + ji] = d0[u];
        }
    }
}

// One block's search: process_block_hex (motion/scalar.rs:546-636) with the seeds and cost
// predictors of gather_local_motion_candidates (:56-156) taken from the neighbour words lS, tS,
// trS, tlS and the temporal word pT. Returns the canonical search MV word; *sad_out = its
// half-pel SAD.
template <bool INTERIOR>
__device__ uint32_t jsearch(const SearchParams& p, const SearchStream& S, const int* ctab, uint32_t* wp, bool load, int ND, int NR, int bx,
                            int by, uint32_t lS, uint32_t tS, uint32_t trS, uint32_t tlS, uint32_t pT, uint4 pr, int lane, int* sad_out,
                            bool* took, uint32_t prev_int, uint32_t prev_word, int prev_sad, uint32_t* int_out) {
    const int sr = p.sr, x0 = bx * 16, y0 = by * 16;
    if (load) jw_load(S, sr, wp, ND, NR, x0, y0, lane);
    JWin w;
    w.wp = wp; w.ND = ND; w.W = S.W; w.H = S.H; w.x0 = x0; w.y0 = y0;
    w.wxa = (x0 - sr - 1) & ~3; w.wy0 = y0 - sr - 1;
    w.j = lane & 15;
    w.cq = *(const uint4*)(wp + riv::search_window_dwords(sr) + 4 * w.j);
    w.cb = x0 - w.wxa; w.rowc = (w.j + sr + 1) * ND;

    const bool has_l = bx > 0, has_t = by > 0, has_tr = has_t && bx + 1 < S.bw, has_tl = has_t && bx > 0, has_tp = S.prev_mv != nullptr;
    // predictor i in lane i+1, in predictor order: left, top, top-right, top-left, temporal;
    // (dx, dy) packed in 16 bits. Lane 0 holds the (0,0) candidate.
    const uint32_t q0 = lS & 0xffffu, q1 = tS & 0xffffu, q2 = trS & 0xffffu, q3 = tlS & 0xffffu, q4 = pT & 0xffffu;
    const uint32_t presm = (has_l ? 2u : 0u) | (has_t ? 4u : 0u) | (has_tr ? 8u : 0u) | (has_tl ? 16u : 0u) | (has_tp ? 32u : 0u);
    const uint32_t cand[6] = {0u, q0, q1, q2, q3, q4};
    const uint32_t myq = lane_pick(lane, cand);
    // a predictor is a duplicate when an earlier present predictor has the same (dx, dy)
    const bool dup = ((lane > 1) & ((presm >> 1) & 1u) & (myq == q0)) | ((lane > 2) & ((presm >> 2) & 1u) & (myq == q1)) |
                     ((lane > 3) & ((presm >> 3) & 1u) & (myq == q2)) | ((lane > 4) & ((presm >> 4) & 1u) & (myq == q3));
    const bool first = (lane < 6) & ((presm >> lane) & 1u) & !dup;
    CostCtx cc;
    {
        // nearest / near of the cost model: the first two distinct present predictors
        const unsigned long long um = __ballot(first), um2 = um & (um - 1);
        const uint32_t n0 = um ? sgpr((uint32_t)__builtin_amdgcn_readlane((int)myq, __builtin_ctzll(um))) : 0u;
        const uint32_t n1 = um2 ? sgpr((uint32_t)__builtin_amdgcn_readlane((int)myq, __builtin_ctzll(um2))) : 0u;
        cc.nx = mvx(n0); cc.ny = mvy(n0); cc.rx = mvx(n1); cc.ry = mvy(n1);
        const uint32_t qm = has_tr ? q2 : (has_tl ? q3 : 0u);
        cc.mx = median3(has_l ? mvx(q0) : 0, has_t ? mvx(q1) : 0, mvx(qm));
        cc.my = median3(has_l ? mvy(q0) : 0, has_t ? mvy(q1) : 0, mvy(qm));
    }
    cc.tab = ctab; cc.td = 2 * sr + 4; cc.c_zero = p.c_zero; cc.c_pred = p.c_pred; cc.lam_zero = p.lam_zero;
    int sdx = 0, sdy = 0;
    int scost;   // SAD + MV cost of the current integer best
    bool skip_int;
    {
        // push_candidate (:46-53) + evaluate_seed_candidates (:638-678): (0,0), then each in-range
        // non-zero predictor not already listed. The sequential "first strictly better" scan is an
        // argmin with ties going to the lower lane, i.e. the earlier candidate.
        const int qx = mvx(myq), qy = mvy(myq);
        const bool myv = lane == 0 || (first & (myq != 0u) & (abs(qx) <= sr) & (abs(qy) <= sr));
        const int dx = myv ? qx : 0, dy = myv ? qy : 0;
        const int sv = jw_set_int<6, INTERIOR>(w, lane, dx, dy);
        const int key = myv ? ((sv + cc.cost(dx, dy)) << 3) | lane : 0x7fffffff;
        skip_int = p.zthr > 0 && __builtin_amdgcn_readlane(sv, 0) <= p.zthr;
        if (skip_int) {
            scost = __builtin_amdgcn_readlane(key, 0) >> 3;
        } else {
            const int m = lane_min8(key), win = m & 7;
            scost = m >> 3;
            const uint32_t wq = (uint32_t)__builtin_amdgcn_readlane((int)myq, win);
            sdx = mvx(wq); sdy = mvy(wq);
            if (p.pthr > 0 && __builtin_amdgcn_readlane(sv, win) <= p.pthr) skip_int = true;
        }
    }
    // learned form: a confident block scores the network's four proposals and walks a small
    // diamond (at most 8 steps) from the best; when that walk's cost is no worse than the best
    // seed's, it replaces the hexagon search
    bool learned = false;
    if (!skip_int && (pr.z & 1u)) {
        const uint32_t props[4] = {pr.x & 0xffffu, pr.x >> 16, pr.y & 0xffffu, pr.y >> 16};
        const uint32_t mq = lane_pick(lane, props);
        const int cx = mvx(mq), cy = mvy(mq);
        const bool v = lane < 4 && abs(cx) <= sr && abs(cy) <= sr;
        const int sv = jw_set_int<4, INTERIOR>(w, lane, v ? cx : 0, v ? cy : 0);
        const int key = v ? ((sv + cc.cost(cx, cy)) << 3) | lane : 0x7fffffff;
        const int m = lane_min8(key);
        int lcost = m >> 3;
        int lx = __builtin_amdgcn_readlane(cx, m & 7), ly = __builtin_amdgcn_readlane(cy, m & 7);
        const int dox = lane < 4 ? (int)((0x2213u >> (4 * lane)) & 15u) - 2 : 0;
        const int doy = lane < 4 ? (int)((0x1322u >> (4 * lane)) & 15u) - 2 : 0;
        for (int it = 0; it < 8; it++) {
            const int dx = lx + dox, dy = ly + doy;
            const bool vd = lane < 4 && abs(dx) <= sr && abs(dy) <= sr;
            const int sd = jw_set_int<4, INTERIOR>(w, lane, vd ? dx : lx, vd ? dy : ly);
            const int kd = vd ? ((sd + cc.cost(dx, dy)) << 3) | lane : 0x7fffffff;
            const int md = lane_min8(kd);
            if ((md >> 3) >= lcost) break;
            lcost = md >> 3;
            lx = __builtin_amdgcn_readlane(dx, md & 7); ly = __builtin_amdgcn_readlane(dy, md & 7);
        }
        if (lcost <= scost) { learned = true; scost = lcost; sdx = lx; sdy = ly; }
    }
    *took = learned;
    if (!skip_int && !learned) {   // refine_integer_motion (:680-757): large hexagon until no improvement, then small diamond
        const int hox = lane < 6 ? (int)((0x310134u >> (4 * lane)) & 15u) - 2 : 0;
        const int hoy = lane < 6 ? (int)((0x002442u >> (4 * lane)) & 15u) - 2 : 0;
        for (int it = 0; it < 4096; it++) {
            const int dx = sdx + hox, dy = sdy + hoy;
            const bool v = lane < 6 && abs(dx) <= sr && abs(dy) <= sr;
            const int ex = v ? dx : sdx, ey = v ? dy : sdy;
            const int sv = jw_set_int<6, INTERIOR>(w, lane, ex, ey);
            const int key = v ? ((sv + cc.cost(dx, dy)) << 3) | lane : 0x7fffffff;
            const int m = lane_min8(key);
            if ((m >> 3) >= scost) break;
            scost = m >> 3;
            sdx = __builtin_amdgcn_readlane(dx, m & 7); sdy = __builtin_amdgcn_readlane(dy, m & 7);
        }
        const int dox = lane < 4 ? (int)((0x2213u >> (4 * lane)) & 15u) - 2 : 0;
        const int doy = lane < 4 ? (int)((0x1322u >> (4 * lane)) & 15u) - 2 : 0;
        const int dx = sdx + dox, dy = sdy + doy;
        const bool v = lane < 4 && abs(dx) <= sr && abs(dy) <= sr;
        const int ex = v ? dx : sdx, ey = v ? dy : sdy;
        const int sv = jw_set_int<4, INTERIOR>(w, lane, ex, ey);
        const int key = v ? ((sv + cc.cost(dx, dy)) << 3) | lane : 0x7fffffff;
        const int m = lane_min8(key);
        if ((m >> 3) < scost) { sdx = __builtin_amdgcn_readlane(dx, m & 7); sdy = __builtin_amdgcn_readlane(dy, m & 7); }
    }
    // The half-pel refine and the canonical vector depend only on the integer best and the window:
    // a re-evaluation that lands on the integer vector of the block's previous one (prev_int, with
    // bit 16 set once there is one) has their result already.
    {
        const uint32_t ib = ((uint32_t)(uint8_t)sdx | ((uint32_t)(uint8_t)sdy << 8)) | 0x10000u;
        *int_out = ib;
        if (ib == prev_int) { *sad_out = prev_sad; return prev_word; }
    }
    // refine_halfpel (:759-839): the base re-scored with the half-pel kernel, then the 8
    // neighbours in (sy outer, sx inner) order at penalty c_pred (every neighbour of an even base
    // is fractional). Candidate 0 is the base, 1..8 the neighbours; the rounds group them by
    // interpolation type: the diagonal ones (1, 3, 6, 8), the axial ones (2, 4, 5, 7), the base.
    int best_hx, best_hy, best_sad;
    {
        // one candidate per 16-lane row (row r = lane >> 4), the set's argmin read from the rows' keys
        const int r = lane >> 4;
        auto rows_min = [](int v) {
            return min(min(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
                       min(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
        };
        auto sel = [](int k) { return k == 0 ? 4 : (k <= 4 ? k - 1 : k); };   // 3x3 raster index
        const int kd = (0x8631 >> (4 * r)) & 15, ka = (0x7542 >> (4 * r)) & 15;
        const int sd = (int)jw_row_hp<2>(w, 2 * sdx + sel(kd) % 3 - 1, 2 * sdy + sel(kd) / 3 - 1);
        const int sa = (int)jw_row_hp<1>(w, 2 * sdx + sel(ka) % 3 - 1, 2 * sdy + sel(ka) / 3 - 1);
        const int s0 = __builtin_amdgcn_readlane((int)jw_row_hp<0>(w, 2 * sdx, 2 * sdy), 0);
        const int m1 = rows_min(min(((sd + p.c_pred) << 4) | kd, ((sa + p.c_pred) << 4) | ka));
        const int m = min(m1, s0 << 4), k = m & 15;
        best_sad = (m >> 4) - (k == 0 ? 0 : p.c_pred);
        best_hx = 2 * sdx + sel(k) % 3 - 1; best_hy = 2 * sdy + sel(k) / 3 - 1;
    }
    // canonicalised search MV (motion/scalar.rs:619-634, motion_vector.rs:103-116)
    int hx = dclamp(best_hx, -255, 255), hy = dclamp(best_hy, -255, 255);
    int dx = dclamp(hx / 2, -128, 127), dy = dclamp(hy / 2, -128, 127), fx = hx % 2, fy = hy % 2;
    if (dx == -128 && fx == -1) fx = 0;
    if (dy == -128 && fy == -1) fy = 0;
    if (fx == -1) { fx = 1; dx = max(-128, dx - 1); }
    if (fy == -1) { fy = 1; dy = max(-128, dy - 1); }
    *sad_out = best_sad;
    return sgpr((uint32_t)(uint8_t)dx | ((uint32_t)(uint8_t)dy << 8) | ((uint32_t)fx << 16) | ((uint32_t)fy << 24));
}

__device__ __forceinline__ uint32_t field_load(const uint32_t* a) { return __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ __forceinline__ void field_store(uint32_t* a, uint32_t v) { __hip_atomic_store(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// A start-field word as a search input: its integer vector when that lies within the search range,
// else (0,0). The start field is only a guess (the relaxation reaches the same fixed point from any
// start), so words left by another frame or size of batch are safe to use.
__device__ __forceinline__ uint32_t seed_word(uint32_t w, int sr) {
    return (abs(mvx(w)) <= sr && abs(mvy(w)) <= sr) ? (w & 0xffffu) : 0u;
}

}  // namespace

// Natural log of a 16x16 block's variance from its sum and sum of squares, as the adaptive
// quantiser defines it (reitero_residual/src/residual.rs:116-167): ((sum_sq*256 - sum^2) as f32)
// / 65536, floored at 4, then glibc's logf.
__device__ __forceinline__ float aq_log_var(uint32_t s, uint32_t s2) {
    float var = (float)((uint64_t)s2 * 256ull - (uint64_t)s * s) / 65536.0f;
    return glibc_logf(fmaxf(var, 4.0f));
}

// Sums (Σv, Σv²) of a block's 16x16 luma prediction at half-pel vector (hx, hy), rounded as the
// encoder's motion compensation rounds it (sample_hp_round below: (a+b+1)>>1, (a+b+c+d+2)>>2;
// clamping to the frame equals interpolating the replicate-padded window), from the block's
// window slot wp; in every lane. Lane l forms pixels 4(l&3) .. 4(l&3)+3 of row l>>2.
__device__ __forceinline__ void window_pred_sums(const uint32_t* wp, int ND, int sr, int cb, int hx, int hy, int lane, uint32_t* s1, uint32_t* s2) {
    const int j = lane >> 2, c = cb + (hx >> 1) + 4 * (lane & 3);
    const uint32_t* q = wp + (j + sr + 1 + (hy >> 1)) * ND + (c >> 2);
    const uint32_t sh = (uint32_t)(c & 3);
    const uint32_t a = __builtin_amdgcn_alignbyte(q[1], q[0], sh);
    const uint32_t b = sh == 3u ? q[1] : __builtin_amdgcn_alignbyte(q[1], q[0], sh + 1u);
    const uint32_t cc = __builtin_amdgcn_alignbyte(q[ND + 1], q[ND], sh);
    const uint32_t d = sh == 3u ? q[ND + 1] : __builtin_amdgcn_alignbyte(q[ND + 1], q[ND], sh + 1u);
    const bool xo = (hx & 1) != 0, yo = (hy & 1) != 0;
    uint32_t t1 = 0u, t2 = 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t pa = (a >> (8 * i)) & 0xffu, pb = (b >> (8 * i)) & 0xffu, pc = (cc >> (8 * i)) & 0xffu, pd = (d >> (8 * i)) & 0xffu;
        const uint32_t v = !xo ? (!yo ? pa : (pa + pc + 1u) >> 1) : (!yo ? (pa + pb + 1u) >> 1 : (pa + pb + pc + pd + 2u) >> 2);
        t1 += v; t2 += v * v;
    }
    *s1 = wave_sum(t1);
    *s2 = wave_sum(t2);
}

__global__ __launch_bounds__(1024) void k_me_search(SearchParams p) {
    extern __shared__ __align__(16) uint32_t lds32[];
    const int lane = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int gw = blockIdx.x * (blockDim.x >> 6) + wv;
    const int nb = p.nb, sr = p.sr, NS = p.nstreams;
    const int NR = 2 * sr + 18, ND = riv::search_pitch(sr), SLOT = riv::search_slot_dwords(sr);
    const int TD = 2 * sr + 4;
    int* ctab = (int*)lds32;                           // TD x TD NEW-vector cost terms
    int* wflag = ctab + ((TD * TD + 3) & ~3);          // [0] a wave of this workgroup changed a block, [1] pass outcome
    SearchStream* sd = (SearchStream*)(wflag + 4);     // the launch's streams
    // this wave's slots: lane k's block in slot k for k < nres, the other lanes' blocks in slot nres
    uint32_t* slots = lds32 + riv::search_slots_offset(sr, NS) / 4 + (size_t)wv * p.nslot * SLOT;
    uint64_t resident = 0;   // slots that hold their block's window
    for (int i = threadIdx.x; i < TD * TD; i += blockDim.x) ctab[i] = new_mv_cost(p.lam2, p.bits, i / TD, i % TD);
    if (!p.streams) { if (threadIdx.x == 0) sd[0] = p.s0; }
    else for (int i = threadIdx.x; i < NS * (int)(sizeof(SearchStream) / 4); i += blockDim.x) ((uint32_t*)sd)[i] = ((const uint32_t*)p.streams)[i];
    const int BW = riv::barrier_words(p.maxp);
    if (blockIdx.x == 0)   // the other bank was the previous launch's; it is next launch's, so clear it
        for (int i = threadIdx.x; i < BW; i += blockDim.x) p.arrive[(size_t)(p.bank ^ 1) * BW + i] = 0u;
    unsigned* arrive = p.arrive + (size_t)p.bank * BW;
    unsigned* release = arrive + BW - 1;   // pass + 1 once every workgroup has finished the pass; bit 31: a block changed
    constexpr int L = riv::kBarrierLeaves;
    const int leaf = blockIdx.x % L;
    const unsigned nleaf = gridDim.x / L + (leaf < (int)(gridDim.x % L) ? 1u : 0u), nlive = min(gridDim.x, (unsigned)L);
    if (threadIdx.x == 0) wflag[0] = 0;
    __syncthreads();

    // Each workgroup owns a contiguous range of blocks, nb / groups of them rounded either way, so
    // the CUs' shares differ by at most one block. The ranges go to the workgroups in XCD-band
    // order: workgroup i runs on XCD i % 8, and XCD x's workgroups take the x-th eighth of the
    // blocks, so blocks next to each other, and their overlapping reference windows, share an L2.
    // Lane k of wave w owns block w + k * waves of the range (the host keeps a range within 64
    // blocks per wave), so a pass's few changed blocks sit on different waves. The block is of
    // stream ms: the last with blk0 <= it.
    int myb;
    bool valid;
    {
        const int G = (int)gridDim.x, x = (int)(blockIdx.x & 7);
        int ord = (int)(blockIdx.x >> 3);
        for (int y = 0; y < x; y++) ord += (G - y + 7) >> 3;   // workgroups on XCDs before x
        const int b0 = (int)(((long long)ord * nb) / G), b1 = (int)(((long long)(ord + 1) * nb) / G);
        myb = b0 + wv + lane * (int)(blockDim.x >> 6);
        valid = myb < b1;
    }
    int ms = 0;
    for (int lo = 0, hi = NS - 1; lo < hi;) {
        const int mid = (lo + hi + 1) >> 1;
        if (sd[mid].blk0 <= myb) lo = mid; else hi = mid - 1;
        ms = lo;
    }
    const int mbw = sd[ms].bw, mloc = valid ? myb - sd[ms].blk0 : 0;
    const int mbx = mloc % mbw, mby = mloc / mbw;
    const uint32_t mypT = (valid && sd[ms].prev_mv) ? *(const uint32_t*)(sd[ms].prev_mv + (size_t)sd[ms].prev_stride * mloc) : 0u;
    // start field: bank 0 as the previous launch over this field left it (each block's final word),
    // a guess that spares the second pass every block whose guess its first pass confirms
    uint32_t myw = valid ? seed_word(field_load(p.field + myb), sr) : 0u;
    // learned form: the block's proposals, read once (an evaluation takes them by lane read)
    const uint4 myprop = (valid && p.prop) ? p.prop[myb] : make_uint4(0u, 0u, 0u, 0u);
    int mysad = 0;
    uint32_t myint = 0u;   // the integer vector the block's last evaluation found (bit 16: there was one)
    bool mytook = false;   // the block's last evaluation took the network's proposal
    int outcome = 0;     // 1 converged, -1 barrier timeout, 0 pass limit
    int pass = 0;
    // The field is double-buffered: pass p reads bank p & 1 (the words as pass p - 1 left them)
    // and writes bank (p + 1) & 1. A block writes a bank only when the bank does not hold its
    // current word (bit b of `fresh`: bank b does), so after two passes without a change it writes
    // nothing. A block is searched again when the integer part of a neighbour word differs from
    // the one its last search used (uL .. uTL).
    uint32_t uL = 0u, uT = 0u, uTR = 0u, uTL = 0u, fresh = 0u;
    for (; pass < p.plimit; pass++) {
        const uint32_t* src = p.field + (size_t)(pass & 1) * nb;
        uint32_t* dst = p.field + (size_t)((pass + 1) & 1) * nb;
        const uint32_t dbit = 1u << ((pass + 1) & 1);
        uint32_t nL = 0u, nT = 0u, nTR = 0u, nTL = 0u;
        bool dirty = valid;
        if (valid) {
            if (mbx > 0) nL = field_load(src + myb - 1) & 0xffffu;
            if (mby > 0) {
                nT = field_load(src + myb - mbw) & 0xffffu;
                if (mbx + 1 < mbw) nTR = field_load(src + myb - mbw + 1) & 0xffffu;
                if (mbx > 0) nTL = field_load(src + myb - mbw - 1) & 0xffffu;
            }
            if (pass == 0) {   // every block searches once, from its neighbours' guesses
                nL = seed_word(nL, sr); nT = seed_word(nT, sr); nTR = seed_word(nTR, sr); nTL = seed_word(nTL, sr);
            } else {
                dirty = ((nL ^ uL) | (nT ^ uT) | (nTR ^ uTR) | (nTL ^ uTL)) != 0u;
            }
            uL = nL; uT = nT; uTR = nTR; uTL = nTL;
        }
        // blocks whose inputs did not change keep their word, written where the bank lacks it
        if (valid && !dirty && !(fresh & dbit)) { field_store(dst + myb, myw); fresh |= dbit; }
        bool chg = false;
        for (unsigned long long dm = __ballot(dirty); dm; dm &= dm - 1) {
            const int k = __builtin_ctzll(dm);
            const int b = __builtin_amdgcn_readlane(myb, k);
            const int bx = __builtin_amdgcn_readlane(mbx, k), by = __builtin_amdgcn_readlane(mby, k);
            const SearchStream& S = sd[__builtin_amdgcn_readlane(ms, k)];
            const uint32_t lS = sgpr((uint32_t)__builtin_amdgcn_readlane((int)nL, k)) & 0x7fffffffu;
            const uint32_t tS = sgpr((uint32_t)__builtin_amdgcn_readlane((int)nT, k)) & 0x7fffffffu;
            const uint32_t trS = sgpr((uint32_t)__builtin_amdgcn_readlane((int)nTR, k)) & 0x7fffffffu;
            const uint32_t tlS = sgpr((uint32_t)__builtin_amdgcn_readlane((int)nTL, k)) & 0x7fffffffu;
            const uint32_t pT = sgpr((uint32_t)__builtin_amdgcn_readlane((int)mypT, k));
            int sad;
            const bool own = k < p.nres;
            uint32_t* wp = slots + (own ? k : p.nres) * SLOT;
            const bool load = !own || !((resident >> k) & 1u);
            if (own) resident |= 1ull << k;
            const uint4 pr = make_uint4(sgpr((uint32_t)__builtin_amdgcn_readlane((int)myprop.x, k)),
                                        sgpr((uint32_t)__builtin_amdgcn_readlane((int)myprop.y, k)),
                                        sgpr((uint32_t)__builtin_amdgcn_readlane((int)myprop.z, k)), 0u);
            bool took;
            const uint32_t old = sgpr((uint32_t)__builtin_amdgcn_readlane((int)myw, k));
            const uint32_t pint = sgpr((uint32_t)__builtin_amdgcn_readlane((int)myint, k));
            const int psad = __builtin_amdgcn_readlane(mysad, k);
            uint32_t nint;
            const uint32_t wS = mb_interior(bx * 16, by * 16, S.W, S.H, sr)
                                    ? jsearch<true>(p, S, ctab, wp, load, ND, NR, bx, by, lS, tS, trS, tlS, pT, pr, lane, &sad, &took,
                                                    pint, old, psad, &nint)
                                    : jsearch<false>(p, S, ctab, wp, load, ND, NR, bx, by, lS, tS, trS, tlS, pT, pr, lane, &sad, &took,
                                                     pint, old, psad, &nint);
            const bool c = ((wS ^ old) & 0xffffu) != 0u;
            chg |= c;
            const bool wr = wS != old || !(sgpr((uint32_t)__builtin_amdgcn_readlane((int)fresh, k)) & dbit);
            if (lane == k) { myw = wS; mysad = sad; mytook = took; myint = nint; fresh = wS != old ? dbit : fresh | dbit; }
            if (lane == 0 && wr) field_store(dst + b, wS);
        }
        // ---- pass barrier: a two-level tree of arrival counters (arrivals in bits 0-15, arrivals
        //      that changed a block in bits 16-31); the last arrival at the root publishes the pass
        //      on the release word, which every workgroup polls ----
        if (chg && lane == 0) __hip_atomic_store(&wflag[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        asm volatile("s_waitcnt vmcnt(0)" ::: "memory");   // this wave's field stores are performed
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned mine = wflag[0] ? 1u : 0u;
            wflag[0] = 0;
            unsigned* cnt = arrive + (size_t)pass * (L + 1);
            const unsigned old = __hip_atomic_fetch_add(cnt + leaf, 1u + (mine << 16), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((old & 0xffffu) + 1u == nleaf) {   // last at this leaf: carry the leaf to the root
                const unsigned lchg = (old >> 16) + mine;
                const unsigned rold = __hip_atomic_fetch_add(cnt + L, 1u + ((lchg ? 1u : 0u) << 16), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((rold & 0xffffu) + 1u == nlive)   // last overall
                    __hip_atomic_store(release, (unsigned)(pass + 1) | (((rold >> 16) + (lchg ? 1u : 0u)) ? 0x80000000u : 0u),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            int res = -1;
            for (unsigned spins = 0; spins < (1u << 20); spins++) {
                const unsigned v = __hip_atomic_load(release, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((v & 0x7fffffffu) == (unsigned)(pass + 1)) { res = (int)(v >> 31); break; }
                __builtin_amdgcn_s_sleep(1);
            }
            wflag[1] = res;
        }
        __syncthreads();
        const int res = wflag[1];
        if (res < 0) { outcome = -1; break; }
        if (res == 0) { outcome = 1; break; }
    }
    if (outcome == 0 && p.truncate) {   // keep the field of the last pass run
        outcome = 1;
        pass = p.plimit - 1;
    }
    if (outcome != 1) {
        const int code = outcome < 0 ? 1 : 3;
        // every stream of the launch failed: its error word tells the host not to code the frame
        for (int i = threadIdx.x; i < NS; i += blockDim.x) sd[i].passes[1] = code;
        return;
    }
    // the mode decision, on the final field (the last pass's bank holds every block's word: the
    // writes completed before the last barrier)
    if (valid) {
        const SearchStream& S = sd[ms];
        const uint32_t* fin = p.field + (size_t)((pass + 1) & 1) * nb;
        const bool has_l = mbx > 0, has_t = mby > 0, has_tr = has_t && mbx + 1 < mbw, has_tl = has_t && mbx > 0;
        const uint32_t wL = has_l ? field_load(fin + myb - 1) : 0u;
        const uint32_t wT = has_t ? field_load(fin + myb - mbw) : 0u;
        const uint32_t wTR = has_tr ? field_load(fin + myb - mbw + 1) : 0u;
        const uint32_t wTL = has_tl ? field_load(fin + myb - mbw - 1) : 0u;
        const uint4 r = mode_record(myw, mysad, has_l, has_t, has_tr, has_tl, S.prev_mv != nullptr, wL, wT, wTR, wTL, mypT,
                                    S.rdo_bits, S.rdo_max, S.passes + 1);
        *(uint4*)&S.rec[mloc] = r;
        *(uint4*)&S.host_rec[mloc] = r;
        // the next launch's start field: bank 0, where it does not hold myw already
        if (!(fresh & 1u)) field_store(p.field + myb, myw);
    }
    if (p.lstat) {   // this wave's own three counters (same-address atomics from every wave would serialise)
        const unsigned long long vb = __ballot(valid), cb = __ballot(valid && (myprop.z & 1u)), tb = __ballot(valid && mytook);
        if (lane == 0 && vb) {
            unsigned long long* q = p.lstat + 3 * (size_t)gw;
            q[0] += (unsigned long long)__popcll(vb);
            q[1] += (unsigned long long)__popcll(cb);
            q[2] += (unsigned long long)__popcll(tb);
        }
    }
    // the adaptive quantiser's input: per block, the log variance of its luma prediction at its
    // final vector, formed from the block's window (reloaded where the block's slot is shared)
    for (unsigned long long vm = __ballot(valid); vm; vm &= vm - 1) {
        const int k = __builtin_ctzll(vm);
        const SearchStream& S = sd[__builtin_amdgcn_readlane(ms, k)];
        if (!S.lnv) continue;
        const int bx = __builtin_amdgcn_readlane(mbx, k), by = __builtin_amdgcn_readlane(mby, k);
        const uint32_t w = sgpr((uint32_t)__builtin_amdgcn_readlane((int)myw, k));
        const int hx = 2 * mvx(w) + (int)((w >> 16) & 0xffu), hy = 2 * mvy(w) + (int)(w >> 24);
        const bool own = k < p.nres;
        uint32_t* wp = slots + (own ? k : p.nres) * SLOT;
        if (!own) jw_load(S, sr, wp, ND, NR, bx * 16, by * 16, lane);
        uint32_t s1, s2;
        window_pred_sums(wp, ND, sr, bx * 16 - ((bx * 16 - sr - 1) & ~3), hx, hy, lane, &s1, &s2);
        if (lane == 0) S.lnv[__builtin_amdgcn_readlane(mloc, k)] = aq_log_var(s1, s2);
    }
    if (blockIdx.x == 0)
        for (int i = threadIdx.x; i < NS; i += blockDim.x)
            *sd[i].passes = pass + 1;
}


// ============================================================================================
// Motion compensation + prediction-variance sums
// ============================================================================================
// build_predicted (reitero_video_common/src/fast_motion.rs:103-352; samplers motion.rs:28-117):
// rounding half-pel bilinear with per-pixel clamping; integer MVs whose luma block lies inside
// the frame take the fast path, which copies chroma at floor((x0+dx)/2) and so drops the
// chroma half-pel phase of odd integer MVs. One 256-thread workgroup per MB; also emits
// (sum, sum of squares) of the predicted luma block for the adaptive quantiser.
__device__ __forceinline__ uint8_t sample_hp_round(const uint8_t* p, int W, int H, int xh, int yh) {
    xh = dclamp(xh, 0, (W - 1) * 2); yh = dclamp(yh, 0, (H - 1) * 2);
    int x0 = xh >> 1, y0 = yh >> 1;
    bool xo = xh & 1, yo = yh & 1;
    int x1 = x0 + 1 < W ? x0 + 1 : x0, y1 = y0 + 1 < H ? y0 + 1 : y0;
    int a = p[(size_t)y0 * W + x0];
    if (!xo && !yo) return (uint8_t)a;
    int b = p[(size_t)y0 * W + x1], d = p[(size_t)y1 * W + x0], e = p[(size_t)y1 * W + x1];
    if (xo && !yo) return (uint8_t)((a + b + 1) >> 1);
    if (!xo && yo) return (uint8_t)((a + d + 1) >> 1);
    return (uint8_t)((a + b + d + e + 2) >> 2);
}

__global__ __launch_bounds__(256) void k_mc(const uint8_t* __restrict__ rY, const uint8_t* __restrict__ rU, const uint8_t* __restrict__ rV,
                                            int W, int H, const int8_t* __restrict__ mvs, int mv_stride,
                                            uint8_t* __restrict__ pY, uint8_t* __restrict__ pU, uint8_t* __restrict__ pV,
                                            uint32_t* __restrict__ var_sums, float* __restrict__ lnv) {
    __shared__ uint32_t red[8];
    const int b = blockIdx.x, bw = W >> 4, bx = b % bw, by = b / bw, t = threadIdx.x;
    const int cw = W >> 1, ch = H >> 1;
    const int8_t* m = mvs + (size_t)b * mv_stride;
    const int mvx = m[0] * 2 + m[2], mvy = m[1] * 2 + m[3];
    const int x0 = bx * 16, y0 = by * 16;
    const int x = x0 + (t & 15), y = y0 + (t >> 4);
    bool fast = false;
    int sx0 = 0, sy0 = 0;
    if (((mvx | mvy) & 1) == 0) {
        sx0 = x0 + mvx / 2; sy0 = y0 + mvy / 2;
        fast = sx0 >= 0 && sy0 >= 0 && sx0 + 15 < W && sy0 + 15 < H;
    }
    uint8_t v;
    if (fast) v = rY[(size_t)(sy0 + (t >> 4)) * W + sx0 + (t & 15)];
    else v = sample_hp_round(rY, W, H, 2 * x + mvx, 2 * y + mvy);
    pY[(size_t)y * W + x] = v;
    if (t < 64) {
        const int ci = t & 7, cj = t >> 3, cx = (x0 >> 1) + ci, cy = (y0 >> 1) + cj;
        uint8_t u, vv;
        if (fast) {
            size_t o = (size_t)((sy0 >> 1) + cj) * cw + (sx0 >> 1) + ci;
            u = rU[o]; vv = rV[o];
        } else {
            int cxh = (4 * cx + mvx) >> 1, cyh = (4 * cy + mvy) >> 1;
            u = sample_hp_round(rU, cw, ch, cxh, cyh);
            vv = sample_hp_round(rV, cw, ch, cxh, cyh);
        }
        pU[(size_t)cy * cw + cx] = u; pV[(size_t)cy * cw + cx] = vv;
    }
    if (var_sums) {
        uint32_t s1 = wave_sum(v), s2 = wave_sum((uint32_t)v * v);
        if ((t & 63) == 0) { red[(t >> 6) * 2] = s1; red[(t >> 6) * 2 + 1] = s2; }
        __syncthreads();
        if (t == 0) {
            const uint32_t s = red[0] + red[2] + red[4] + red[6], s2 = red[1] + red[3] + red[5] + red[7];
            var_sums[2 * b] = s;
            var_sums[2 * b + 1] = s2;
            if (lnv) lnv[b] = aq_log_var(s, s2);
        }
    }
}

// ============================================================================================
// Adaptive quantisation (compute_aq_quant_steps, reitero_residual/src/residual.rs:116-167)
// ============================================================================================
// variance = ((sum_sq*256 - sum^2) as f32) / 65536, floored at 4, and its natural log with
// glibc's logf (computed per MB where the sums are formed, in k_mc); the mean of the logs is a
// sequential f32 sum in raster order (the rounding sequence is part of the result);
// step = base * exp2f((0.8 * (ln - avg)) / 6).
//
// The sequential sum, exactly, in parallel. Every log is >= ln 4 > 0, so the running sum only
// grows. While it stays within one binade [2^e, 2^(e+1)) every partial sum is a multiple of that
// binade's ulp u, and adding x rounds to s + u * (x rounded to a multiple of u), independent of s
// except when x lies exactly halfway between multiples of u, where ties-to-even makes it depend
// on the parity of s / u (the parity that follows is determined too). So a run of values adds a
// fixed K(e, parity) ulps to any start in binade e with that parity, provided the result stays in
// binade e (the partial sums in between are then in it as well); and adding K ulps to a float of
// binade e is adding K to its bit pattern, the exponent field unchanged.
// Runs of 32 values (sub-chunks) get K for the binade their (f64) prefix sum predicts and the
// adjacent one, both parities, by running them in f32 from 2^e (+u); runs of 8 sub-chunks
// (chunks) compose their sub-chunks' K the same way in integers. One wave then chains the chunks,
// adding K to the sum's bits where the binade and parity match and the exponent field stays; a
// chunk that does not (it crosses a binade, or is among the first) is chained by its sub-chunks,
// and a sub-chunk that does not by its values one by one, in f32. The result is the sequential
// sum bit for bit whatever the estimates; they only decide how often the slower levels run.
constexpr int kAqSub = 32;          // values per sub-chunk
constexpr int kAqGroup = 8;         // sub-chunks per chunk
constexpr int kAqStage = 32768;     // most logs the parallel path takes (staged in LDS)
constexpr int kAqNoK = 0x40000000;  // no K: adding it changes the exponent field
__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
// log j's LDS word: an XOR swizzle inside each group of 32, so that the lanes of a wave, each
// reading value i of its own sub-chunk, hit 32 different banks
__device__ __forceinline__ int aq_pos(int j) { return j ^ ((j >> 5) & 31); }
// the binade of P and its neighbour on the side P is near
__device__ __forceinline__ int2 aq_binades(double P) {
    const float pf = (float)P;
    const int e = (int)(fbits(pf) >> 23);
    const uint32_t man = fbits(pf) & 0x7fffffu;
    return make_int2(e, man < (1u << 17) ? e - 1 : (man >= 0x7fffffu - (1u << 17) ? e + 1 : e));
}
// K for binade eb.x and eb.y (even, odd start) at sum bits b (exponent field e), or kAqNoK
__device__ __forceinline__ int aq_pick(uint32_t b, int4 k, int2 eb) {
    const int e = (int)(b >> 23);
    const bool odd = (b & 1u) != 0u;
    return e == eb.x ? (odd ? k.y : k.x) : (e == eb.y ? (odd ? k.w : k.z) : kAqNoK);
}

// An MB's quantiser step: base * exp2(0.8 * (ln v - mean) / 6) in f32 (residual.rs:150-166).
__device__ __forceinline__ float aq_step(float lnv, float mean, float base) {
    const float off = __fmul_rn(0.8f, __fsub_rn(lnv, mean));
    return __fmul_rn(base, glibc_exp2f(__fdiv_rn(off, 6.0f)));
}

size_t aq_lds_bytes() {
    const int nsub = kAqStage / kAqSub, nch = nsub / kAqGroup;
    return (size_t)kAqStage * 4 + (size_t)nsub * (8 + 16) + (size_t)nch * (8 + 16 + 8) + 16;
}

__global__ __launch_bounds__(1024) void k_aq(float* __restrict__ lnv, int nb, float base, float* __restrict__ steps) {
    extern __shared__ __align__(16) uint32_t aq_lds[];
    const int t = threadIdx.x, nt = blockDim.x, lane = t & 63;
    const int nsub = (nb + kAqSub - 1) / kAqSub, nch = (nsub + kAqGroup - 1) / kAqGroup;
    const int NSUB = kAqStage / kAqSub, NCH = NSUB / kAqGroup;
    float* xs = (float*)aq_lds;
    double* pre = (double*)(xs + kAqStage);      // sub-chunk sums, then their exclusive prefix; then binades
    int2* sbin = (int2*)pre;
    int4* sk = (int4*)(pre + NSUB);
    double* cpre = (double*)(sk + NSUB);
    int4* ck = (int4*)(cpre + NCH);
    int2* cbin = (int2*)(ck + NCH);
    float* s_mean = (float*)(cbin + NCH);
    if (nb > kAqStage) {   // beyond the staged size: the plain sequential sum
        if (t == 0) {
            float s = 0.0f;
            for (int i = 0; i < nb; i++) s = __fadd_rn(s, lnv[i]);
            *s_mean = __fdiv_rn(s, (float)nb);
            lnv[nb] = *s_mean;
        }
        __syncthreads();
    } else {
        for (int j0 = t; j0 < nb; j0 += 8 * nt) {   // 8 loads in flight per thread
            float v[8];
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = j0 + u * nt < nb ? lnv[j0 + u * nt] : 0.0f;
#pragma unroll
            for (int u = 0; u < 8; u++) if (j0 + u * nt < nb) xs[aq_pos(j0 + u * nt)] = v[u];
        }
        __syncthreads();
        // sub-chunk t (nsub <= 1024 threads): its values stay in registers from its sum to its K
        const bool own = t < nsub;
        const int lo = t * kAqSub, nv = own ? min(kAqSub, nb - lo) : 0;
        float v[kAqSub];
#pragma unroll
        for (int i = 0; i < kAqSub; i++) v[i] = i < nv ? xs[aq_pos(lo + i)] : 0.0f;
        {   // sub-chunk sums (f64: only an estimate, any order will do)
            double cs = 0.0;
#pragma unroll
            for (int i = 0; i < kAqSub; i++) cs += (double)v[i];
            if (own) pre[t] = cs;
        }
        __syncthreads();
        {   // exclusive prefix of the sub-chunk sums (an estimate as well): each wave scans 64, the
            // waves' totals are added after (cpre holds them until it is written)
            const int w = t >> 6;
            const int c = t;
            const double x = c < nsub ? pre[c] : 0.0;
            double inc = x;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const double o = __shfl_up(inc, d);
                if (lane >= d) inc += o;
            }
            if (lane == 63) cpre[w] = inc;
            __syncthreads();
            double carry = 0.0;
            for (int g = 0; g < w; g++) carry += cpre[g];
            __syncthreads();
            if (c < nsub) pre[c] = carry + inc - x;
        }
        __syncthreads();
        for (int k = t; k < nch; k += nt) cpre[k] = pre[k * kAqGroup];
        if (own) {   // sub-chunk K: the run in f32 from 2^e (+u)
            const int2 eb = aq_binades(pre[t]);
            const uint32_t b0 = (uint32_t)eb.x << 23, b2 = (uint32_t)eb.y << 23;
            float s0 = __uint_as_float(b0), s1 = __uint_as_float(b0 | 1u), s2 = __uint_as_float(b2), s3 = __uint_as_float(b2 | 1u);
            bool pos = true;
#pragma unroll
            for (int i = 0; i < kAqSub; i++) {   // values past nb are +0, which leaves a positive sum as it is
                pos &= i >= nv || v[i] > 0.0f;
                s0 = __fadd_rn(s0, v[i]); s1 = __fadd_rn(s1, v[i]); s2 = __fadd_rn(s2, v[i]); s3 = __fadd_rn(s3, v[i]);
            }
            const bool ve = pos && eb.x >= 1 && eb.x <= 253, va = pos && eb.y >= 1 && eb.y <= 253;
            int4 k;
            k.x = ve && (fbits(s0) >> 23) == (uint32_t)eb.x ? (int)(fbits(s0) - b0) : kAqNoK;
            k.y = ve && (fbits(s1) >> 23) == (uint32_t)eb.x ? (int)(fbits(s1) - (b0 | 1u)) : kAqNoK;
            k.z = va && (fbits(s2) >> 23) == (uint32_t)eb.y ? (int)(fbits(s2) - b2) : kAqNoK;
            k.w = va && (fbits(s3) >> 23) == (uint32_t)eb.y ? (int)(fbits(s3) - (b2 | 1u)) : kAqNoK;
            sk[t] = k;
            sbin[t] = eb;
        }
        __syncthreads();
        for (int q = t; q < 4 * nch; q += nt) {   // chunk K, one (chunk, binade, parity) per thread
            const int k = q >> 2, which = q & 3;
            const int2 eb = aq_binades(cpre[k]);
            const int e = which < 2 ? eb.x : eb.y;
            const uint32_t b0 = ((uint32_t)e << 23) | (uint32_t)(which & 1);
            const int c0 = k * kAqGroup, nsc = min(kAqGroup, nsub - c0);
            int4 kv[kAqGroup];
            int2 ev[kAqGroup];
#pragma unroll
            for (int j = 0; j < kAqGroup; j++)
                if (j < nsc) { kv[j] = sk[c0 + j]; ev[j] = sbin[c0 + j]; }
            uint32_t b = b0;
            bool ok = e >= 1 && e <= 253;
#pragma unroll
            for (int j = 0; j < kAqGroup; j++)
                if (ok && j < nsc) {
                    b += (uint32_t)aq_pick(b, kv[j], ev[j]);
                    ok = (b >> 23) == (uint32_t)e;
                }
            ((int*)&ck[k])[which] = ok ? (int)(b - b0) : kAqNoK;
            if (which == 0) cbin[k] = eb;
        }
        __syncthreads();
        if (t < 64) {
            // one wave, every lane holding the same sum: lanes fetch together what the chain reads
            // in turn (the chunks' K, a chunk's 8 sub-chunk K, a sub-chunk's 32 values, one LDS
            // instruction each) and the chain takes them by lane reads
            uint32_t b = 0u;   // the sum's bits
            // every chunk's K and binades in lanes (chunk k in lane k & 63 of set k >> 6), taken by
            // lane reads: the chain waits for no memory
            const int4 z4 = make_int4(0, 0, 0, 0);
            const int2 z2 = make_int2(-9, -9);
            const int4 ka = lane < nch ? ck[lane] : z4, kb = lane + 64 < nch ? ck[lane + 64] : z4;
            const int2 ea = lane < nch ? cbin[lane] : z2, eb2 = lane + 64 < nch ? cbin[lane + 64] : z2;
            for (int k = 0; k < nch; k++) {
                const int kl = k & 63;
                const int4 ks = k < 64 ? ka : kb;
                const int2 es = k < 64 ? ea : eb2;
                const int4 kk = make_int4(__builtin_amdgcn_readlane(ks.x, kl), __builtin_amdgcn_readlane(ks.y, kl),
                                          __builtin_amdgcn_readlane(ks.z, kl), __builtin_amdgcn_readlane(ks.w, kl));
                const int2 ee = make_int2(__builtin_amdgcn_readlane(es.x, kl), __builtin_amdgcn_readlane(es.y, kl));
                const uint32_t nx = b + (uint32_t)aq_pick(b, kk, ee);
                if ((nx >> 23) == (b >> 23)) { b = nx; continue; }
                const int c0 = k * kAqGroup, nsc = min(kAqGroup, nsub - c0);
                const int4 skl = lane < nsc ? sk[c0 + lane] : make_int4(0, 0, 0, 0);
                const int2 sbl = lane < nsc ? sbin[c0 + lane] : make_int2(-9, -9);
                for (int j = 0; j < nsc; j++) {
                    const int4 kj = make_int4(__builtin_amdgcn_readlane(skl.x, j), __builtin_amdgcn_readlane(skl.y, j),
                                              __builtin_amdgcn_readlane(skl.z, j), __builtin_amdgcn_readlane(skl.w, j));
                    const int2 ej = make_int2(__builtin_amdgcn_readlane(sbl.x, j), __builtin_amdgcn_readlane(sbl.y, j));
                    const uint32_t ns = b + (uint32_t)aq_pick(b, kj, ej);
                    if ((ns >> 23) == (b >> 23)) { b = ns; continue; }
                    const int lo = (c0 + j) * kAqSub;
                    // past nb: +0, which leaves a positive sum as it is
                    const float vl = (lane < kAqSub && lo + lane < nb) ? xs[aq_pos(lo + lane)] : 0.0f;
                    float sf = __uint_as_float(b);
#pragma unroll
                    for (int i = 0; i < kAqSub; i++) sf = __fadd_rn(sf, __uint_as_float((uint32_t)__builtin_amdgcn_readlane((int)fbits(vl), i)));
                    b = fbits(sf);
                }
            }
            if (t == 0) {
                *s_mean = __fdiv_rn(__uint_as_float(b), (float)nb);
                lnv[nb] = *s_mean;
            }
        }
        __syncthreads();
    }
    // per-MB steps (when asked for: the encoder's residual kernel forms each MB's step itself), the
    // logs from LDS where they are staged
    if (!steps) return;
    const float mean = *s_mean;
    const bool staged = nb <= kAqStage;
    for (int b = t; b < nb; b += nt) steps[b] = aq_step(staged ? xs[aq_pos(b)] : lnv[b], mean, base);
}

// The adaptive quantiser's per-MB input straight from a luma plane (the plane taken as its own
// prediction): the log variance of each 16x16 block, one wave per block, lane l summing the four
// pixels 4(l&3) .. 4(l&3)+3 of row l>>2 (one aligned dword; W is a multiple of 16).
__global__ __launch_bounds__(256) void k_aq_logs(const uint8_t* __restrict__ Y, int W, int nb, float* __restrict__ lnv) {
    const int b = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (b >= nb) return;   // whole waves
    const int bw = W >> 4, x0 = (b % bw) * 16, y0 = (b / bw) * 16;
    const uint32_t px = *(const uint32_t*)(Y + (size_t)(y0 + (lane >> 2)) * W + x0 + 4 * (lane & 3));
    uint32_t t1 = 0u, t2 = 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t v = (px >> (8 * i)) & 0xffu;
        t1 += v; t2 += v * v;
    }
    const uint32_t s1 = wave_sum(t1), s2 = wave_sum(t2);
    if (lane == 0) lnv[b] = aq_log_var(s1, s2);
}

// elementwise logf/exp2f for parity tests of the transcendental restatements
__global__ void k_logf_exp2f(const float* in, size_t n, int which, float* out) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = which == 0 ? glibc_logf(in[i]) : glibc_exp2f(in[i]);
}

// ============================================================================================
// Residual transforms (reitero_dct/src/dct16.rs, dct8.rs, fast_dct8.rs)
// ============================================================================================
namespace {
// FDCT: rows then columns, i64 accumulation, +2^17 >> 18 (dct16.rs:151-194, dct8.rs:32-54).
// `res` is the 16x16 residual (row-major, int16 in LDS); coefficient (u,v) -> out[v*16+u].
__device__ __forceinline__ void fdct16(const int16_t* res, int32_t* tmp, int32_t* out, int t) {
    {
        int y = t >> 4, u = t & 15; long long acc = 0;
#pragma unroll
        for (int n = 0; n < 16; n++) acc += (long long)res[y * 16 + n] * c_B16[n][u];
        tmp[y * 16 + u] = (int32_t)((acc + (1ll << 17)) >> 18);
    }
    __syncthreads();
    {
        int v = t >> 4, u = t & 15; long long acc = 0;
#pragma unroll
        for (int y = 0; y < 16; y++) acc += (long long)tmp[y * 16 + u] * c_B16[y][v];
        out[v * 16 + u] = (int32_t)((acc + (1ll << 17)) >> 18);
    }
}
// 8x8 FDCT for one plane by 64 threads (tt = 0..63)
__device__ __forceinline__ void fdct8_rows(const int16_t* res, int32_t* tmp, int tt) {
    int y = tt >> 3, u = tt & 7; long long acc = 0;
#pragma unroll
    for (int n = 0; n < 8; n++) acc += (long long)res[y * 8 + n] * c_B8[n][u];
    tmp[y * 8 + u] = (int32_t)((acc + (1ll << 17)) >> 18);
}
__device__ __forceinline__ int32_t fdct8_cols(const int32_t* tmp, int tt) {
    int v = tt >> 3, u = tt & 7; long long acc = 0;
#pragma unroll
    for (int y = 0; y < 8; y++) acc += (long long)tmp[y * 8 + u] * c_B8[y][v];
    return (int32_t)((acc + (1ll << 17)) >> 18);
}
// IDCT16: rows then columns, wrapping i32 (dct16.rs:201-233); d = dequantised coefficients.
__device__ __forceinline__ int16_t idct16(const int32_t* d, int32_t* tmp, int t) {
    {
        int v = t >> 4, x = t & 15; int32_t acc = 0;
#pragma unroll
        for (int u = 0; u < 16; u++) acc = wadd(acc, wmul(d[v * 16 + u], c_B16[x][u]));
        tmp[v * 16 + x] = wadd(acc, 1 << 17) >> 18;
    }
    __syncthreads();
    int y = t >> 4, x = t & 15; int32_t acc = 0;
#pragma unroll
    for (int v = 0; v < 16; v++) acc = wadd(acc, wmul(tmp[v * 16 + x], c_B16[y][v]));
    return (int16_t)dclamp(wadd(acc, 1 << 17) >> 18, -32768, 32767);
}
// AAN/stb 8x8 IDCT with the reference's F2F constants (fast_dct8.rs:123-174, :254-330)
__device__ __forceinline__ void aan1d(const int32_t* s, int32_t xsc, int32_t* x, int32_t* tt) {
    const int32_t F0541 = 2218, F1848 = -7572, F0765 = 3136, F1176 = 4817, F0299 = 1223, F2053 = 8411, F3073 = 12595,
                  F1501 = 6150, F0900 = -3685, F2563 = -10496, F1962 = -8035, F0390 = -1598;
    int32_t p2 = s[2], p3 = s[6];
    int32_t p1 = wmul(wadd(p2, p3), F0541);
    int32_t t2 = wadd(p1, wmul(p3, F1848)), t3 = wadd(p1, wmul(p2, F0765));
    p2 = s[0]; p3 = s[4];
    int32_t t0 = wshl(wadd(p2, p3), 12), t1 = wshl(wsub(p2, p3), 12);
    x[0] = wadd(wadd(t0, t3), xsc); x[3] = wadd(wsub(t0, t3), xsc);
    x[1] = wadd(wadd(t1, t2), xsc); x[2] = wadd(wsub(t1, t2), xsc);
    int32_t a0 = s[7], a1 = s[5], a2 = s[3], a3 = s[1];
    int32_t q3 = wadd(a0, a2), q4 = wadd(a1, a3), q1 = wadd(a0, a3), q2 = wadd(a1, a2);
    int32_t p5 = wmul(wadd(q3, q4), F1176);
    a0 = wmul(a0, F0299); a1 = wmul(a1, F2053); a2 = wmul(a2, F3073); a3 = wmul(a3, F1501);
    q1 = wadd(p5, wmul(q1, F0900)); q2 = wadd(p5, wmul(q2, F2563));
    q3 = wmul(q3, F1962); q4 = wmul(q4, F0390);
    tt[3] = wadd(wadd(a3, q1), q4); tt[2] = wadd(wadd(a2, q2), q3); tt[1] = wadd(wadd(a1, q2), q4); tt[0] = wadd(wadd(a0, q1), q3);
}
// first pass over row r (horizontal frequencies) -> tmp[x*8 + r]
__device__ __forceinline__ void aan_pass1(const int16_t* coef, const uint16_t* q, int r, int32_t* tmp) {
    int32_t s[8], x[4], t[4];
#pragma unroll
    for (int u = 0; u < 8; u++) s[u] = wmul((int32_t)coef[r * 8 + u], (int32_t)q[r * 8 + u]);
    aan1d(s, 512, x, t);
    tmp[0 * 8 + r] = wadd(x[0], t[3]) >> 10; tmp[7 * 8 + r] = wsub(x[0], t[3]) >> 10;
    tmp[1 * 8 + r] = wadd(x[1], t[2]) >> 10; tmp[6 * 8 + r] = wsub(x[1], t[2]) >> 10;
    tmp[2 * 8 + r] = wadd(x[2], t[1]) >> 10; tmp[5 * 8 + r] = wsub(x[2], t[1]) >> 10;
    tmp[3 * 8 + r] = wadd(x[3], t[0]) >> 10; tmp[4 * 8 + r] = wsub(x[3], t[0]) >> 10;
}
// second pass over column xc -> out[y*8 + xc]
__device__ __forceinline__ void aan_pass2(const int32_t* tmp, int xc, int16_t* out) {
    int32_t s[8], x[4], t[4];
#pragma unroll
    for (int r = 0; r < 8; r++) s[r] = tmp[xc * 8 + r];
    aan1d(s, 65536, x, t);
    int32_t o[8] = {wadd(x[0], t[3]) >> 17, wadd(x[1], t[2]) >> 17, wadd(x[2], t[1]) >> 17, wadd(x[3], t[0]) >> 17,
                    wsub(x[3], t[0]) >> 17, wsub(x[2], t[1]) >> 17, wsub(x[1], t[2]) >> 17, wsub(x[0], t[3]) >> 17};
#pragma unroll
    for (int y = 0; y < 8; y++) out[y * 8 + xc] = (int16_t)dclamp(o[y], -32768, 32767);
}
__device__ __forceinline__ int16_t quant_aq(int32_t c, float qs, float dz) {   // dct16.rs:321-333
    float r = __fdiv_rn((float)c, qs);
    return fabsf(r) < dz ? 0 : f32_to_i16_sat(roundf(r));
}
__device__ __forceinline__ int16_t quant_matrix(int32_t c, uint16_t q) {   // dct16.rs:467-480
    float qs = (float)(q < 1 ? 1 : q);
    return f32_to_i16_sat(roundf(__fdiv_rn((float)c, qs)));
}

struct ResidLds {
    int16_t res[384];       // Y 256 | U 64 | V 64 residual (or pixel-128 for intra)
    int32_t tmp[256];
    int32_t tmpc[128];
    int32_t coef32[384];
    int16_t coef[384];
    int32_t deq[256];
    int16_t rc[128];        // chroma residual reconstruction
    uint16_t qc[64];
};

// dequant + IDCT16 + add/clamp on luma, AAN on chroma, for one MB whose int16 coefficients
// are in L.coef; writes the reconstruction (pred + residual, clamped) to out planes.
// Luma: (c*round(qs*16384) + 8192) >> 14 (dct16.rs:371-381); chroma: c*max(1, round(qs))
// (fast_dct8.rs:95). With `intra`, the tables yt/ct replace the AQ step and +128 replaces pred.
// predY / predC >= 0: thread t's luma / chroma prediction sample, used instead of reading pY / pU / pV.
__device__ void recon_mb(ResidLds& L, int t, float qs, bool intra, const uint16_t* yt, const uint16_t* ct,
                         const uint8_t* pY, const uint8_t* pU, const uint8_t* pV, uint8_t* oY, uint8_t* oU, uint8_t* oV,
                         int W, int x0, int y0, int predY = -1, int predC = -1) {
    const int cw = W >> 1;
    if (intra) {
        L.deq[t] = wmul((int32_t)L.coef[t], (int32_t)(yt[t] < 1 ? 1 : yt[t]));
        if (t < 64) L.qc[t] = ct[t];
    } else {
        int32_t qfp = f32_to_i32_sat(roundf(__fmul_rn(qs, 16384.0f)));
        L.deq[t] = wadd(wmul((int32_t)L.coef[t], qfp), 1 << 13) >> 14;
        if (t < 64) L.qc[t] = f32_to_u16_sat(fmaxf(roundf(qs), 1.0f));
    }
    __syncthreads();
    int16_t r = idct16(L.deq, L.tmp, t);
    {
        int yy = t >> 4, xx = t & 15;
        size_t o = (size_t)(y0 + yy) * W + x0 + xx;
        int base = intra ? 128 : (predY >= 0 ? predY : pY[o]);
        oY[o] = (uint8_t)dclamp(base + (int)r, 0, 255);
    }
    if (t < 16) aan_pass1(L.coef + 256 + (t >> 3) * 64, L.qc, t & 7, L.tmpc + (t >> 3) * 64);
    __syncthreads();
    if (t < 16) aan_pass2(L.tmpc + (t >> 3) * 64, t & 7, L.rc + (t >> 3) * 64);
    __syncthreads();
    if (t < 128) {
        int pl = t >> 6, tt = t & 63, yy = tt >> 3, xx = tt & 7;
        size_t o = (size_t)((y0 >> 1) + yy) * cw + (x0 >> 1) + xx;
        const uint8_t* ps = pl ? pV : pU;
        uint8_t* od = pl ? oV : oU;
        int base = intra ? 128 : (predC >= 0 ? predC : ps[o]);
        od[o] = (uint8_t)dclamp(base + (int)L.rc[pl * 64 + tt], 0, 255);
    }
}

// FDCT of L.res (luma 16x16 + two chroma 8x8) into L.coef32
__device__ void fdct_mb(ResidLds& L, int t) {
    fdct16(L.res, L.tmp, L.coef32, t);
    if (t < 128) fdct8_rows(L.res + 256 + (t >> 6) * 64, L.tmpc + (t >> 6) * 64, t & 63);
    __syncthreads();
    if (t < 128) L.coef32[256 + t] = fdct8_cols(L.tmpc + (t >> 6) * 64, t & 63);
    __syncthreads();
}
}  // namespace

// Writes the MB's quantised coefficients (LDS) to its packed slot in mapped host memory: wave 0
// forms the six non-zero masks with ballots, and each non-zero lane stores its value at its rank.
__device__ __forceinline__ void write_packed(const int16_t* coef, uint8_t* slot, int t) {
    if (t >= 64) return;
    int base = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int16_t v = coef[64 * k + t];
        const unsigned long long m = __ballot(v != 0);
        if (t == 0) ((unsigned long long*)slot)[k] = m;
        if (v != 0) ((int16_t*)(slot + 48))[base + __popcll(m & ((1ull << t) - 1ull))] = v;
        base += __popcll(m);
    }
}

// The same record split for the host's reads: the six masks, and the values when there are at
// most PACK_MAIN_VALUES of them, in the MB's compact slot (main, PACK_MAIN bytes per MB, so a
// frame's records are one short contiguous run); more values go to the MB's full-size slot
// (over). Mapped memory the device has just written is read from DRAM, so the bytes and pages the
// host touches are what it costs.
__device__ __forceinline__ void write_packed_split(const int16_t* coef, uint8_t* main, uint8_t* over, int t) {
    if (t >= 64) return;
    unsigned long long m[6];
    int nnz = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) { m[k] = __ballot(coef[64 * k + t] != 0); nnz += __popcll(m[k]); }
#pragma unroll
    for (int k = 0; k < 6; k++) if (t == k) ((unsigned long long*)main)[k] = m[k];
    int16_t* vals = (int16_t*)((nnz <= riv::PACK_MAIN_VALUES ? main : over) + 48);
    int base = 0;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        const int16_t v = coef[64 * k + t];
        if (v != 0) vals[base + __popcll(m[k] & ((1ull << t) - 1ull))] = v;
        base += __popcll(m[k]);
    }
}

// P-frame residual: ResidualEncoder::encode_inter (reitero_residual/src/residual.rs:396-809)
// minus the rANS: residual = cur - pred for MBs the RDO did not skip, FDCT, dead-zone AQ
// quantisation, all-zero -> optimised skip, else dequant + IDCT reconstruction. Writes the
// pre-deblock reconstruction, the coefficients of coded MBs and the optimised skip flags.
__global__ __launch_bounds__(256) void k_inter_residual(const uint8_t* cY, const uint8_t* cU, const uint8_t* cV,
                                                        const uint8_t* rY, const uint8_t* rU, const uint8_t* rV,
                                                        uint8_t* pY, uint8_t* pU, uint8_t* pV,
                                                        int W, int H, const BlockRec* rec, const float* lnv, float base,
                                                        float* steps, float dz,
                                                        uint8_t* oY, uint8_t* oU, uint8_t* oV, uint8_t* main, uint8_t* over,
                                                        uint8_t* mask, uint8_t* host_opt_skip, uint8_t* dpack, uint8_t* dcoded) {
    __shared__ ResidLds L;
    const int b = blockIdx.x, bw = W >> 4, bx = b % bw, by = b / bw, t = threadIdx.x, cw = W >> 1, ch = H >> 1;
    const int x0 = bx * 16, y0 = by * 16;
    // the MB's adaptive quantiser step is formed where it is used, once, from its log variance and
    // the frame's mean (lnv[nb]); kept in steps (skipped MBs too) for the host's inspection
    __shared__ float s_qs;
    if (t == 0) { const float q = aq_step(lnv[b], lnv[gridDim.x], base); s_qs = q; steps[b] = q; }
    // the prediction (motion compensation with the search MV, which the mode decision keeps;
    // fast_motion.rs:103-352): thread t forms luma sample t and, below 128, chroma sample t & 63 of
    // plane t >> 6, the samples it later uses; they also go to pY / pU / pV
    const int8_t* m = rec[b].s;
    const int mvx = m[0] * 2 + m[2], mvy = m[1] * 2 + m[3];
    int sx0 = 0, sy0 = 0;
    bool fast = false;
    if (((mvx | mvy) & 1) == 0) {
        sx0 = x0 + mvx / 2; sy0 = y0 + mvy / 2;
        fast = sx0 >= 0 && sy0 >= 0 && sx0 + 15 < W && sy0 + 15 < H;
    }
    const size_t o = (size_t)(y0 + (t >> 4)) * W + x0 + (t & 15);
    const int predY = fast ? rY[(size_t)(sy0 + (t >> 4)) * W + sx0 + (t & 15)]
                           : sample_hp_round(rY, W, H, 2 * (x0 + (t & 15)) + mvx, 2 * (y0 + (t >> 4)) + mvy);
    pY[o] = (uint8_t)predY;
    const int pl = t >> 6, tt = t & 63;
    size_t oc = 0;
    int predC = -1;
    if (t < 128) {
        const int ci = tt & 7, cj = tt >> 3, cx = (x0 >> 1) + ci, cy = (y0 >> 1) + cj;
        oc = (size_t)cy * cw + cx;
        const uint8_t* rc = pl ? rV : rU;
        predC = fast ? rc[(size_t)((sy0 >> 1) + cj) * cw + (sx0 >> 1) + ci] : sample_hp_round(rc, cw, ch, (4 * cx + mvx) >> 1, (4 * cy + mvy) >> 1);
        (pl ? pV : pU)[oc] = (uint8_t)predC;
    }
    if (rec[b].rdo_skip) {   // RDO skip: prediction only
        oY[o] = (uint8_t)predY;
        if (t < 128) (pl ? oV : oU)[oc] = (uint8_t)predC;
        if (t == 0) { mask[b] = *(const uint32_t*)rec[b].r != 0u; host_opt_skip[b] = 1; if (dcoded) dcoded[b] = 0; }
        return;
    }
    L.res[t] = (int16_t)((int)cY[o] - predY);
    if (t < 128) L.res[256 + t] = (int16_t)((int)(pl ? cV : cU)[oc] - predC);
    __syncthreads();
    fdct_mb(L, t);
    const float qs = s_qs;   // written before the barriers above
    int nz = 0;
    for (int i = t; i < 384; i += 256) { int16_t q = quant_aq(L.coef32[i], qs, dz); L.coef[i] = q; nz |= q != 0; }
    nz = __syncthreads_or(nz);
    if (!nz) {   // optimised skip (residual.rs:618-674)
        oY[o] = (uint8_t)predY;
        if (t < 128) (pl ? oV : oU)[oc] = (uint8_t)predC;
        if (t == 0) { mask[b] = *(const uint32_t*)rec[b].r != 0u; host_opt_skip[b] = 1; if (dcoded) dcoded[b] = 0; }
        return;
    }
    if (t == 0) { mask[b] = 1; host_opt_skip[b] = 0; if (dcoded) dcoded[b] = 1; }
    if (main) write_packed_split(L.coef, main + (size_t)b * riv::PACK_MAIN, over + (size_t)b * riv::PACK_STRIDE, t);
    if (dpack) write_packed(L.coef, dpack + (size_t)b * riv::PACK_STRIDE, t);
    recon_mb(L, t, qs, false, nullptr, nullptr, pY, pU, pV, oY, oU, oV, W, x0, y0, predY, predC);
}

// I-frame: ResidualEncoder::encode_intra (residual.rs:267-388) minus the rANS: pixel - 128,
// FDCT, per-coefficient matrix quantisation (no dead zone), integer-table dequant + IDCT, +128.
__global__ __launch_bounds__(256) void k_intra(const uint8_t* cY, const uint8_t* cU, const uint8_t* cV, int W,
                                               const uint16_t* yt, const uint16_t* ct, uint8_t* oY, uint8_t* oU, uint8_t* oV,
                                               uint8_t* pack) {
    __shared__ ResidLds L;
    __shared__ uint16_t syt[256], sct[64];
    const int b = blockIdx.x, bw = W >> 4, bx = b % bw, by = b / bw, t = threadIdx.x, cw = W >> 1;
    const int x0 = bx * 16, y0 = by * 16;
    syt[t] = yt[t];
    if (t < 64) sct[t] = ct[t];
    {
        size_t o = (size_t)(y0 + (t >> 4)) * W + x0 + (t & 15);
        L.res[t] = (int16_t)((int)cY[o] - 128);
        if (t < 128) {
            int pl = t >> 6, tt = t & 63;
            size_t oc = (size_t)((y0 >> 1) + (tt >> 3)) * cw + (x0 >> 1) + (tt & 7);
            L.res[256 + t] = (int16_t)((pl ? (int)cV[oc] : (int)cU[oc]) - 128);
        }
    }
    __syncthreads();
    fdct_mb(L, t);
    for (int i = t; i < 384; i += 256) {
        int16_t q = quant_matrix(L.coef32[i], i < 256 ? syt[i] : sct[(i - 256) & 63]);
        L.coef[i] = q;
    }
    __syncthreads();
    write_packed(L.coef, pack + (size_t)b * riv::PACK_STRIDE, t);
    recon_mb(L, t, 0.0f, true, syt, sct, nullptr, nullptr, nullptr, oY, oU, oV, W, x0, y0);
}

// Decoder P-frame reconstruction: ResidualDecoder::decode_inter (residual.rs:889-1079) after
// the rANS: skipped MBs keep the prediction, coded MBs add dequant + IDCT of their coefficients.
__global__ __launch_bounds__(256) void k_inter_recon(const uint8_t* pY, const uint8_t* pU, const uint8_t* pV, int W,
                                                     const uint8_t* skip, const float* steps, const int16_t* coeffs,
                                                     uint8_t* oY, uint8_t* oU, uint8_t* oV) {
    __shared__ ResidLds L;
    const int b = blockIdx.x, bw = W >> 4, bx = b % bw, by = b / bw, t = threadIdx.x, cw = W >> 1;
    const int x0 = bx * 16, y0 = by * 16;
    if (skip[b]) {
        size_t o = (size_t)(y0 + (t >> 4)) * W + x0 + (t & 15);
        oY[o] = pY[o];
        if (t < 128) {
            int pl = t >> 6, tt = t & 63;
            size_t oc = (size_t)((y0 >> 1) + (tt >> 3)) * cw + (x0 >> 1) + (tt & 7);
            if (pl) oV[oc] = pV[oc]; else oU[oc] = pU[oc];
        }
        return;
    }
    for (int i = t; i < 384; i += 256) L.coef[i] = coeffs[(size_t)b * 384 + i];
    __syncthreads();
    recon_mb(L, t, steps[b], false, nullptr, nullptr, pY, pU, pV, oY, oU, oV, W, x0, y0);
}

// Decoder I-frame reconstruction (residual.rs:815-887 after the rANS)
__global__ __launch_bounds__(256) void k_intra_recon(int W, const uint16_t* yt, const uint16_t* ct, const int16_t* coeffs,
                                                     uint8_t* oY, uint8_t* oU, uint8_t* oV) {
    __shared__ ResidLds L;
    __shared__ uint16_t syt[256], sct[64];
    const int b = blockIdx.x, bw = W >> 4, bx = b % bw, by = b / bw, t = threadIdx.x;
    syt[t] = yt[t];
    if (t < 64) sct[t] = ct[t];
    for (int i = t; i < 384; i += 256) L.coef[i] = coeffs[(size_t)b * 384 + i];
    __syncthreads();
    recon_mb(L, t, 0.0f, true, syt, sct, nullptr, nullptr, nullptr, oY, oU, oV, W, bx * 16, by * 16);
}

// Plane-level FDCT/quant/IDCT for the reitero_dct seams (one MB-sized block per workgroup;
// `n8` selects the 8x8 path, which processes one 8x8 block per workgroup).
__global__ __launch_bounds__(256) void k_dct_plane(const int16_t* plane, int w, int n8, const float* qs, float dz, int16_t* coeffs, int16_t* recon) {
    __shared__ ResidLds L;
    const int b = blockIdx.x, t = threadIdx.x;
    if (!n8) {
        const int bw = w / 16, bx = b % bw, by = b / bw;
        L.res[t] = plane[(size_t)(by * 16 + (t >> 4)) * w + bx * 16 + (t & 15)];
        __syncthreads();
        fdct16(L.res, L.tmp, L.coef32, t);
        __syncthreads();
        int16_t q = quant_aq(L.coef32[t], qs[b], dz);
        L.coef[t] = q; coeffs[(size_t)b * 256 + t] = q;
        int32_t qfp = f32_to_i32_sat(roundf(__fmul_rn(qs[b], 16384.0f)));
        L.deq[t] = wadd(wmul((int32_t)q, qfp), 1 << 13) >> 14;
        __syncthreads();
        int16_t r = idct16(L.deq, L.tmp, t);
        recon[(size_t)(by * 16 + (t >> 4)) * w + bx * 16 + (t & 15)] = r;
    } else {
        const int bw = w / 8, bx = b % bw, by = b / bw;
        if (t < 64) L.res[t] = plane[(size_t)(by * 8 + (t >> 3)) * w + bx * 8 + (t & 7)];
        __syncthreads();
        if (t < 64) fdct8_rows(L.res, L.tmpc, t);
        __syncthreads();
        if (t < 64) {
            int16_t q = quant_aq(fdct8_cols(L.tmpc, t), qs[b], dz);
            L.coef[t] = q; coeffs[(size_t)b * 64 + t] = q;
            L.qc[t] = f32_to_u16_sat(fmaxf(roundf(qs[b]), 1.0f));
        }
        __syncthreads();
        if (t < 8) aan_pass1(L.coef, L.qc, t, L.tmpc);
        __syncthreads();
        if (t < 8) aan_pass2(L.tmpc, t, L.rc);
        __syncthreads();
        if (t < 64) recon[(size_t)(by * 8 + (t >> 3)) * w + bx * 8 + (t & 7)] = L.rc[t];
    }
}

// ============================================================================================
// In-loop deblock (reitero_video_common/src/deblock.rs): VP8 normal macroblock-edge filter,
// all vertical edges of a plane, then all horizontal ones. Edges are 16 (8 for chroma) pixels
// apart and a filter touches 4 pixels on each side, so every edge line of one pass is
// independent: one thread per (edge, line). Y, U and V run in the same launch.
// ============================================================================================
__device__ __forceinline__ int c8(int x) { return dclamp(x, -128, 127); }
__device__ __forceinline__ void mb_filter_line(int* px, int E, int I, int HEV) {   // deblock.rs:68-115
    int p3 = px[0], p2 = px[1], p1 = px[2], p0 = px[3], q0 = px[4], q1 = px[5], q2 = px[6], q3 = px[7];
    bool mask = 2 * abs(p0 - q0) + abs(q1 - p1) / 2 <= E && abs(p3 - p2) <= I && abs(p2 - p1) <= I && abs(p1 - p0) <= I &&
                abs(q1 - q0) <= I && abs(q2 - q1) <= I && abs(q3 - q2) <= I;
    if (!mask) return;
    int ps2 = p2 - 128, ps1 = p1 - 128, ps0 = p0 - 128, qs0 = q0 - 128, qs1 = q1 - 128, qs2 = q2 - 128;
    bool hev = abs(p1 - p0) > HEV || abs(q1 - q0) > HEV;
    int w = c8(c8(ps1 - qs1) + 3 * (qs0 - ps0));
    if (hev) {
        int f1 = c8(w + 4) >> 3, f2 = c8(w + 3) >> 3;
        px[4] = dclamp(c8(qs0 - f1) + 128, 0, 255);
        px[3] = dclamp(c8(ps0 + f2) + 128, 0, 255);
    } else {
        int a = c8((27 * w + 63) >> 7);
        int nq0 = c8(qs0 - a), np0 = c8(ps0 + a);
        a = c8((18 * w + 63) >> 7);
        int nq1 = c8(qs1 - a), np1 = c8(ps1 + a);
        a = c8((9 * w + 63) >> 7);
        int nq2 = c8(qs2 - a), np2 = c8(ps2 + a);
        px[1] = dclamp(np2 + 128, 0, 255); px[2] = dclamp(np1 + 128, 0, 255); px[3] = dclamp(np0 + 128, 0, 255);
        px[4] = dclamp(nq0 + 128, 0, 255); px[5] = dclamp(nq1 + 128, 0, 255); px[6] = dclamp(nq2 + 128, 0, 255);
    }
}
// mask: one byte per MB (1 = filter edges touching it), nullptr = every edge (intra)
__global__ void k_deblock(uint8_t* Y, uint8_t* U, uint8_t* V, int W, int H, int E, int I, int HEV, const uint8_t* mask, int horizontal) {
    const int bw = W >> 4, bh = H >> 4;
    const long long nY = horizontal ? (long long)(bh - 1) * W : (long long)(bw - 1) * H;
    const long long nC = horizontal ? (long long)(bh - 1) * (W >> 1) : (long long)(bw - 1) * (H >> 1);
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    uint8_t* P; int w, grid;
    if (i < nY) { P = Y; w = W; grid = 16; }
    else if (i < nY + nC) { P = U; w = W >> 1; grid = 8; i -= nY; }
    else if (i < nY + 2 * nC) { P = V; w = W >> 1; grid = 8; i -= nY + nC; }
    else return;
    int px[8];
    if (!horizontal) {
        // consecutive threads take consecutive edges of one row (bw - 1 per row in every plane),
        // so a wave's segments share a few cache lines; edges 16 (8) pixels apart never overlap
        int e = (int)(i % (bw - 1)) + 1, y = (int)(i / (bw - 1));
        int x = e * grid, by = y / grid;
        if (mask && !(mask[by * bw + e - 1] || mask[by * bw + e])) return;
        uint32_t* p = (uint32_t*)(P + (size_t)y * w + x - 4);   // x - 4 is a multiple of 4
        const uint32_t a = p[0], b = p[1];
#pragma unroll
        for (int k = 0; k < 4; k++) { px[k] = (int)((a >> (8 * k)) & 0xffu); px[4 + k] = (int)((b >> (8 * k)) & 0xffu); }
        mb_filter_line(px, E, I, HEV);
        uint32_t na = 0u, nb2 = 0u;
#pragma unroll
        for (int k = 0; k < 4; k++) { na |= (uint32_t)(uint8_t)px[k] << (8 * k); nb2 |= (uint32_t)(uint8_t)px[4 + k] << (8 * k); }
        p[0] = na; p[1] = nb2;
    } else {
        int e = (int)(i / w) + 1, x = (int)(i % w);
        int y = e * grid, bxx = x / grid;
        if (mask && !(mask[(e - 1) * bw + bxx] || mask[e * bw + bxx])) return;
        uint8_t* p = P + (size_t)(y - 4) * w + x;
#pragma unroll
        for (int k = 0; k < 8; k++) px[k] = p[(size_t)k * w];
        mb_filter_line(px, E, I, HEV);
#pragma unroll
        for (int k = 0; k < 8; k++) p[(size_t)k * w] = (uint8_t)px[k];
    }
}

// Deblock mask of a P-frame (encoder.rs:831-845): !optimised_skip || reconstructed MV != 0

// ============================================================================================
// Host-side launch wrappers
// ============================================================================================
namespace rivk {

hipError_t upload_bases(const int32_t* b16, const int32_t* b8) {
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_B16), b16, sizeof(int32_t) * 256);
    if (e != hipSuccess) return e;
    return hipMemcpyToSymbol(HIP_SYMBOL(c_B8), b8, sizeof(int32_t) * 64);
}

void rgb_to_yuv420(hipStream_t st, const uint8_t* rgb, int dw, int dh, int sw, int sh, uint8_t* Y, uint8_t* U, uint8_t* V) {
    dim3 g((sw / 2 + 127) / 128, sh / 2);
    hipLaunchKernelGGL(k_rgb_to_yuv420, g, dim3(128), 0, st, rgb, dw, dh, sw, sh, Y, U, V);
}
void yuv420_to_rgb(hipStream_t st, const uint8_t* Y, const uint8_t* U, const uint8_t* V, int sw, int dw, int dh, uint8_t* rgb) {
    dim3 g((dw + 255) / 256, dh);
    hipLaunchKernelGGL(k_yuv420_to_rgb, g, dim3(256), 0, st, Y, U, V, sw, dw, dh, rgb);
}

int sad_table_stride(int sr) { return ((2 * sr + 1) * (2 * sr + 1) + 7) & ~7; }
int hp_table_stride(int sr) { return ((4 * sr + 3) * (4 * sr + 3) + 7) & ~7; }
int table_region_bytes(int sr) { return 2 * (sad_table_stride(sr) + hp_table_stride(sr)); }
bool mb_is_interior(int x0, int y0, int W, int H, int sr) { return mb_interior(x0, y0, W, H, sr); }
void sad_table(hipStream_t st, const uint8_t* cur, const uint8_t* ref, int W, int H, int sr, uint16_t* out, int stride) {
    const int ndwt = (16 + 2 * sr + 6) / 4 + 1;
    const size_t lds = (size_t)(64 + 4 * (16 + 2 * sr) * ndwt) * 4;
    hipLaunchKernelGGL(k_sad_table, dim3((W / 16) * (H / 16)), dim3(256), lds, st, cur, ref, W, H, W / 16, sr, ndwt, out, stride);
}
static constexpr int HPK = 8;   // y-offsets per half-pel-table thread
void hp_table(hipStream_t st, const uint8_t* cur, const uint8_t* ref, int W, int H, int sr, uint16_t* out, int stride) {
    const int ndp = (2 * sr + 4) / 4 + 6, WS = 16 + 2 * sr + 2;
    const size_t lds = (size_t)(64 + (WS + HPK + 1) * (ndp + 1) + 4 * (WS + HPK) * ndp) * 4;
    hipLaunchKernelGGL(k_hp_table<HPK>, dim3((W / 16) * (H / 16)), dim3(128), lds, st, cur, ref, W, H, W / 16, sr, ndp, out, stride);
}
size_t search_lds_bytes(int sr, int waves, int nstreams, int nslot) {
    return riv::search_slots_offset(sr, nstreams) + (size_t)waves * nslot * riv::search_slot_dwords(sr) * sizeof(uint32_t);
}
bool search_shape(int sr, int nb, int nstreams, int cus, size_t budget, SearchShape* out, int max_waves) {
    // one resident workgroup per CU at most, as many waves per workgroup (up to max_waves) as hold
    // one slot each, at least one wave per 64 blocks; then as many slots per wave as fit, up to the
    // blocks a wave owns
    int waves = std::max(1, std::min(16, max_waves));
    while (waves > 1 && search_lds_bytes(sr, waves, nstreams, 1) > budget) waves--;
    if (search_lds_bytes(sr, waves, nstreams, 1) > budget) return false;
    const int groups = std::max(1, std::min(cus, (nb + waves - 1) / waves));
    const long long nw = (long long)groups * waves;
    if (nw * 64 < nb) return false;
    const int owned = (int)((nb + nw - 1) / nw);
    const size_t per = (size_t)waves * riv::search_slot_dwords(sr) * sizeof(uint32_t);
    const int fit = (int)std::min<size_t>(64, (budget - riv::search_slots_offset(sr, nstreams)) / per);
    out->waves = waves; out->groups = groups;
    out->nslot = std::min(owned, fit);
    out->nres = owned <= fit ? owned : fit - 1;
    out->lds = search_lds_bytes(sr, waves, nstreams, out->nslot);
    return true;
}
std::vector<SearchShape> search_shapes(int sr, int nb, int nstreams, int cus, size_t budget) {
    // the occupancy-derived shape first, then every other feasible (waves, groups) of a small grid,
    // each with as many window slots per wave as fit
    std::vector<SearchShape> out;
    SearchShape d;
    if (search_shape(sr, nb, nstreams, cus, budget, &d)) out.push_back(d);
    const int wv[] = {16, 14, 12, 10, 8, 6, 4};
    const int gq[] = {4, 3, 2};   // groups = cus * gq / 4
    for (int waves : wv) {
        if (search_lds_bytes(sr, waves, nstreams, 1) > budget) continue;
        for (int q : gq) {
            const int groups = std::max(1, std::min(cus * q / 4, (nb + waves - 1) / waves));
            const long long nw = (long long)groups * waves;
            if (nw * 64 < nb) continue;
            const int owned = (int)((nb + nw - 1) / nw);
            const size_t per = (size_t)waves * riv::search_slot_dwords(sr) * sizeof(uint32_t);
            const int fit = (int)std::min<size_t>(64, (budget - riv::search_slots_offset(sr, nstreams)) / per);
            if (fit < 1) continue;
            SearchShape sh;
            sh.waves = waves; sh.groups = groups;
            sh.nslot = std::min(owned, fit);
            sh.nres = owned <= fit ? owned : fit - 1;
            sh.lds = search_lds_bytes(sr, waves, nstreams, sh.nslot);
            bool dup = false;
            for (const SearchShape& o : out) dup |= o.waves == sh.waves && o.groups == sh.groups;
            if (!dup) out.push_back(sh);
        }
    }
    return out;
}
hipError_t me_search(hipStream_t st, const SearchParams& p, int groups, int waves, hipEvent_t ev0, hipEvent_t ev1) {
    const size_t lds = search_lds_bytes(p.sr, waves, p.nstreams, p.nslot);
    if (ev0 || ev1)   // the launch itself records the events at the kernel's start and end (no packets of their own)
        hipExtLaunchKernelGGL(k_me_search, dim3(groups), dim3(64 * waves), lds, st, ev0, ev1, 0, p);
    else
        hipLaunchKernelGGL(k_me_search, dim3(groups), dim3(64 * waves), lds, st, p);
    return hipGetLastError();
}
void mc(hipStream_t st, const uint8_t* rY, const uint8_t* rU, const uint8_t* rV, int W, int H, const int8_t* mvs, int mv_stride,
        uint8_t* pY, uint8_t* pU, uint8_t* pV, uint32_t* var_sums, float* lnv) {
    hipLaunchKernelGGL(k_mc, dim3((W / 16) * (H / 16)), dim3(256), 0, st, rY, rU, rV, W, H, mvs, mv_stride, pY, pU, pV, var_sums, lnv);
}
void aq(hipStream_t st, int nb, float base, float* lnv, float* steps) {
    hipLaunchKernelGGL(k_aq, dim3(1), dim3(1024), aq_lds_bytes(), st, lnv, nb, base, steps);
}
void aq_logs(hipStream_t st, const uint8_t* Y, int W, int H, float* lnv) {
    const int nb = (W / 16) * (H / 16);
    hipLaunchKernelGGL(k_aq_logs, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, st, Y, W, nb, lnv);
}
void logf_exp2f(hipStream_t st, const float* in, size_t n, int which, float* out) {
    hipLaunchKernelGGL(k_logf_exp2f, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, in, n, which, out);
}
void inter_residual(hipStream_t st, const uint8_t* cY, const uint8_t* cU, const uint8_t* cV, const uint8_t* rY, const uint8_t* rU,
                    const uint8_t* rV, uint8_t* pY, uint8_t* pU, uint8_t* pV, int W, int H, const BlockRec* rec, const float* lnv, float base,
                    float* steps, float dz, uint8_t* oY, uint8_t* oU, uint8_t* oV, uint8_t* main, uint8_t* over, uint8_t* mask,
                    uint8_t* host_opt_skip, uint8_t* dpack, uint8_t* dcoded) {
    hipLaunchKernelGGL(k_inter_residual, dim3((W / 16) * (H / 16)), dim3(256), 0, st, cY, cU, cV, rY, rU, rV, pY, pU, pV, W, H, rec, lnv, base,
                       steps, dz, oY, oU, oV, main, over, mask, host_opt_skip, dpack, dcoded);
}
void intra(hipStream_t st, const uint8_t* cY, const uint8_t* cU, const uint8_t* cV, int W, int H, const uint16_t* yt, const uint16_t* ct,
           uint8_t* oY, uint8_t* oU, uint8_t* oV, uint8_t* pack) {
    hipLaunchKernelGGL(k_intra, dim3((W / 16) * (H / 16)), dim3(256), 0, st, cY, cU, cV, W, yt, ct, oY, oU, oV, pack);
}
void inter_recon(hipStream_t st, const uint8_t* pY, const uint8_t* pU, const uint8_t* pV, int W, int H, const uint8_t* skip,
                 const float* steps, const int16_t* coeffs, uint8_t* oY, uint8_t* oU, uint8_t* oV) {
    hipLaunchKernelGGL(k_inter_recon, dim3((W / 16) * (H / 16)), dim3(256), 0, st, pY, pU, pV, W, skip, steps, coeffs, oY, oU, oV);
}
void intra_recon(hipStream_t st, int W, int H, const uint16_t* yt, const uint16_t* ct, const int16_t* coeffs, uint8_t* oY, uint8_t* oU,
                 uint8_t* oV) {
    hipLaunchKernelGGL(k_intra_recon, dim3((W / 16) * (H / 16)), dim3(256), 0, st, W, yt, ct, coeffs, oY, oU, oV);
}
void dct_plane(hipStream_t st, const int16_t* plane, int w, int h, int n8, const float* qs, float dz, int16_t* coeffs, int16_t* recon) {
    int nb = n8 ? (w / 8) * (h / 8) : (w / 16) * (h / 16);
    hipLaunchKernelGGL(k_dct_plane, dim3(nb), dim3(256), 0, st, plane, w, n8, qs, dz, coeffs, recon);
}
void deblock(hipStream_t st, uint8_t* Y, uint8_t* U, uint8_t* V, int W, int H, int level, bool intra, const uint8_t* mask) {
    if (level <= 0) return;
    // thresholds (deblock.rs:40-58)
    int I = level > 1 ? level : 1, E = (level + 2) * 2 + I, HEV;
    if (intra) HEV = level >= 40 ? 2 : (level >= 15 ? 1 : 0);
    else HEV = level >= 40 ? 3 : (level >= 20 ? 2 : (level >= 15 ? 1 : 0));
    int bw = W / 16, bh = H / 16;
    long long nv = (long long)(bw - 1) * H + 2ll * (bw - 1) * (H / 2);
    long long nh = (long long)(bh - 1) * W + 2ll * (bh - 1) * (W / 2);
    if (nv > 0) hipLaunchKernelGGL(k_deblock, dim3((unsigned)((nv + 255) / 256)), dim3(256), 0, st, Y, U, V, W, H, E, I, HEV, mask, 0);
    if (nh > 0) hipLaunchKernelGGL(k_deblock, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, st, Y, U, V, W, H, E, I, HEV, mask, 1);
}

}  // namespace rivk
