// Learned MV-proposal network on gfx950: BF16 MFMA (v_mfma_f32_16x16x16_bf16) with FP32
// accumulation, for inference inside the encoder and for training on (block, MV) pairs from the
// plain search.
//
// The network, per 16x16 luma block:
//   input   32x32x4 (NHWC): the current block with an 8-pixel border, the co-located reference
//           patch, their difference and a zero channel ((v - 128) / 64, (c - r) / 32);
//   conv1   3x3 stride 2, 4 -> 16 channels, 32x32 -> 16x16, ReLU;
//   conv2   3x3 stride 2, 16 -> 16, 16x16 -> 8x8, ReLU;
//   conv3   3x3 stride 2, 16 -> 16, 8x8 -> 4x4, ReLU (256 features);
//   fc1     the 256 features and 16 predictor inputs (seven vectors of the previous P-frame around
//           the block, /8) -> 128, ReLU;
//   head    128 -> 9: four candidate vectors (x, y, in units of 8 pixels) and a confidence logit.
//
// Every layer is a GEMM over K in steps of 16: a conv's K index is tap * C + channel (tap =
// 3 * ky + kx), zero past 9 * C. Weights are [K][N] row-major; every activation between layers is
// BF16 (ReLU applied to the FP32 sum plus bias), so the fused inference kernel and the layer-wise
// training path round at the same points and sum K in the same order.
#include <hip/hip_runtime.h>
#include <cstdint>
#include <algorithm>
#include "mvnet.hpp"

namespace {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {   // round to nearest even (v_cvt_pk_bf16_f32)
    return __builtin_bit_cast(uint16_t, (__bf16)f);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ v4f mfma16(v4s a, v4s b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0); }

// s_waitcnt lgkmcnt(0) with vmcnt and expcnt at their maxima: LDS (and scalar) accesses drain,
// vector memory loads stay in flight
__device__ __forceinline__ void lds_wait() { __builtin_amdgcn_s_waitcnt(0xC07F); }

}  // namespace

// ============================================================================================
// Inputs
// ============================================================================================
// One workgroup (256 threads) per sample: the 32x32x4 input patch (NHWC, bf16) and the 16
// predictor inputs, written to columns 256..271 of the sample's fc1 input row.
__device__ __forceinline__ void sample_of(const int* samples, int i, int* src, int* blk) {
    if (samples) { *src = samples[2 * i]; *blk = samples[2 * i + 1]; }
    else { *src = 0; *blk = i; }
}

__device__ __forceinline__ uint2 patch_px(const rivnet::Source& S, int bx, int by, int px, int py) {
    const int x = min(max(bx * 16 - 8 + px, 0), (int)S.W - 1), y = min(max(by * 16 - 8 + py, 0), (int)S.H - 1);
    const int c = S.cur[(size_t)y * S.W + x], r = S.ref[(size_t)y * S.W + x];
    const uint16_t a = f2bf((float)(c - 128) * (1.0f / 64.0f)), b = f2bf((float)(r - 128) * (1.0f / 64.0f));
    const uint16_t d = f2bf((float)(c - r) * (1.0f / 32.0f));
    return make_uint2((uint32_t)a | ((uint32_t)b << 16), (uint32_t)d);
}

// The seven vectors of the previous P-frame at the block and around it (centre, left, right,
// top, bottom, top-left, top-right; clamped to the frame), as (x/8, y/8), then two zeros.
__device__ __forceinline__ float pred_input(const rivnet::Source& S, int bx, int by, int k) {
    if (k >= 14 || !S.prev_mv) return 0.0f;
    const int v = k >> 1;
    // offsets + 1, two bits per vector: C (0,0), L (-1,0), R (1,0), T (0,-1), B (0,1), TL (-1,-1), TR (1,-1)
    const int ox = ((0x2161 >> (2 * v)) & 3) - 1, oy = ((0x215 >> (2 * v)) & 3) - 1;
    const int x = min(max(bx + ox, 0), (int)S.bw - 1), y = min(max(by + oy, 0), (int)S.bh - 1);
    return (float)S.prev_mv[(size_t)(y * S.bw + x) * S.prev_stride + (k & 1)] * 0.125f;
}

__global__ __launch_bounds__(256) void k_net_gather(const rivnet::Source* __restrict__ srcs, const int* __restrict__ samples, int n,
                                                    uint16_t* __restrict__ x0, uint16_t* __restrict__ a3) {
    const int i = blockIdx.x;
    if (i >= n) return;
    int si, blk;
    sample_of(samples, i, &si, &blk);
    const rivnet::Source& S = srcs[si];
    const int bx = blk % S.bw, by = blk / S.bw;
    uint2* out = (uint2*)(x0 + (size_t)i * 4096);
    for (int p = threadIdx.x; p < 1024; p += 256) out[p] = patch_px(S, bx, by, p & 31, p >> 5);
    if (threadIdx.x < 16) a3[(size_t)i * rivnet::FC_IN + 256 + threadIdx.x] = f2bf(pred_input(S, bx, by, threadIdx.x));
}

// ============================================================================================
// Training path: im2col, a strided GEMM, col2im, reductions, loss, Adam
// ============================================================================================
// out[(b, oy, ox)][k] for k < KP: input pixel (2oy + ky - 1, 2ox + kx - 1), channel ci, where
// k = (3ky + kx) * C + ci; zero past 9C and outside the image.
__global__ __launch_bounds__(256) void k_im2col(const uint16_t* __restrict__ x, int B, int Hi, int C, int KP, uint16_t* __restrict__ out) {
    const int Ho = Hi / 2;
    const size_t total = (size_t)B * Ho * Ho * KP;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (size_t)gridDim.x * 256) {
        const int k = (int)(t % KP);
        const size_t m = t / KP;
        const int ox = (int)(m % Ho), oy = (int)((m / Ho) % Ho);
        const size_t b = m / ((size_t)Ho * Ho);
        uint16_t v = 0;
        const int tap = k / C, ci = k - tap * C;
        if (tap < 9) {
            const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
            if (iy >= 0 && iy < Hi && ix >= 0 && ix < Hi) v = x[((b * Hi + iy) * Hi + ix) * C + ci];
        }
        out[t] = v;
    }
}

// C = op(A) op(B) with bf16 operands and f32 accumulation; A(m, k) = A[m*sam + k*sak],
// B(k, n) = B[k*sbk + n*sbn]. One wave per 16x16 tile of C; blockIdx.y is the K split.
// splits == 1: out offset (m / mg) * ms + (m % mg) * ldc + n; Z = sum + bias (if Z), Cb = bf16 of
// it, ReLU'd when relu. splits > 1: split s writes its partial sums to C + s * M * N (row-major).
__global__ __launch_bounds__(256) void k_gemm(rivnet::Gemm g) {
    const int lane = threadIdx.x & 63;
    const int tiles_m = (g.M + 15) / 16, tiles_n = (g.N + 15) / 16;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= tiles_m * tiles_n) return;
    const int tm = t % tiles_m, tn = t / tiles_m;
    const int kb = blockIdx.y * g.kc, ke = min(g.K, kb + g.kc);
    const int row = tm * 16 + (lane & 15), col = tn * 16 + (lane & 15), kq = 4 * (lane >> 4);
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    for (int k0 = kb; k0 < ke; k0 += 16) {
        v4s a, b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int k = k0 + kq + j;
            a[j] = (short)((row < g.M && k < ke) ? g.A[(long long)row * g.sam + (long long)k * g.sak] : 0);
            b[j] = (short)((col < g.N && k < ke) ? g.B[(long long)k * g.sbk + (long long)col * g.sbn] : 0);
        }
        acc = mfma16(a, b, acc);
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int m = tm * 16 + kq + j, n = col;
        if (m >= g.M || n >= g.N) continue;
        if (g.splits > 1) {
            g.C[(size_t)blockIdx.y * g.M * g.N + (size_t)m * g.N + n] = acc[j];
        } else {
            const long long off = (long long)(m / g.mg) * g.ms + (long long)(m % g.mg) * g.ldc + n;
            const float z = acc[j] + (g.bias ? g.bias[n] : 0.0f);
            if (g.C) g.C[off] = z;
            if (g.Cb) g.Cb[off] = f2bf(g.relu ? fmaxf(z, 0.0f) : z);
        }
    }
}

// out[i] = sum over s of part[s * count + i], in split order.
__global__ void k_sum_splits(const float* __restrict__ part, int splits, int count, float* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    float s = 0.0f;
    for (int k = 0; k < splits; k++) s += part[(size_t)k * count + i];
    out[i] = s;
}

// dZ (bf16, [B*Ho*Ho][C] = NHWC of the layer input) from the column gradient dcol
// ([B*Ho*Ho][KP], f32) and the layer input's pre-activation Z (f32, same NHWC layout): the sum
// over the taps that read the pixel, times ReLU'(Z).
__global__ __launch_bounds__(256) void k_col2im_relu(const float* __restrict__ dcol, int B, int Hi, int C, int KP, const float* __restrict__ z,
                                                     uint16_t* __restrict__ dz) {
    const int Ho = Hi / 2;
    const size_t total = (size_t)B * Hi * Hi * C;
    for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < total; t += (size_t)gridDim.x * 256) {
        const int ci = (int)(t % C);
        const size_t p = t / C;
        const int ix = (int)(p % Hi), iy = (int)((p / Hi) % Hi);
        const size_t b = p / ((size_t)Hi * Hi);
        float s = 0.0f;
        for (int ky = 0; ky < 3; ky++) {
            const int ny = iy + 1 - ky;
            if (ny < 0 || (ny & 1) || ny / 2 >= Ho) continue;
            for (int kx = 0; kx < 3; kx++) {
                const int nx = ix + 1 - kx;
                if (nx < 0 || (nx & 1) || nx / 2 >= Ho) continue;
                s += dcol[((b * Ho + ny / 2) * Ho + nx / 2) * KP + (3 * ky + kx) * C + ci];
            }
        }
        dz[t] = f2bf(z[t] > 0.0f ? s : 0.0f);
    }
}

// dz[m][n] = bf16(da * ReLU'(z)) where da and z sit at (m / mg) * ms + (m % mg) * ldc + n.
__global__ void k_relu_bwd(const float* __restrict__ da, const float* __restrict__ z, int M, int N, int mg, long long ms, int ldc,
                           uint16_t* __restrict__ dz) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)M * N) return;
    const int m = (int)(t / N), n = (int)(t % N);
    const long long off = (long long)(m / mg) * ms + (long long)(m % mg) * ldc + n;
    dz[t] = f2bf(z[off] > 0.0f ? da[off] : 0.0f);
}

// out[n] = sum over m of dz[m][n] (bias gradients); one workgroup per column, fixed order.
__global__ __launch_bounds__(256) void k_colsum(const uint16_t* __restrict__ dz, int M, int N, float* __restrict__ out) {
    __shared__ float part[256];
    const int n = blockIdx.x;
    float s = 0.0f;
    for (int m = threadIdx.x; m < M; m += 256) s += bf2f(dz[(size_t)m * N + n]);
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[n] = part[0];
}

__device__ __forceinline__ float smooth_l1(float d, float beta, float* g) {
    const float a = fabsf(d);
    if (a < beta) { *g = d / beta; return 0.5f * d * d / beta; }
    *g = d > 0.0f ? 1.0f : -1.0f;
    return a - 0.5f * beta;
}

// Per sample: the candidates' smooth-L1 distances to the target vector (units of 8 pixels, beta
// one pixel); the nearest candidate takes the full loss, every candidate a small share; the
// confidence is trained (logistic loss) to say whether the nearest candidate lies within one
// pixel of the target on both axes. dy = gradient of the batch mean; loss[i] per sample.
__global__ void k_net_loss(const float* __restrict__ y, const int8_t* __restrict__ tgt, int n, float* __restrict__ dy, float* __restrict__ loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* o = y + (size_t)i * rivnet::OUT;
    float* d = dy + (size_t)i * rivnet::OUT;
    const float tx = (float)tgt[2 * i] * 0.125f, ty = (float)tgt[2 * i + 1] * 0.125f;
    const float beta = 0.125f, share = 0.02f, inv = 1.0f / (float)n;
    float dist[4], gx[4], gy[4];
    int best = 0;
    for (int c = 0; c < 4; c++) {
        dist[c] = smooth_l1(o[2 * c] - tx, beta, &gx[c]) + smooth_l1(o[2 * c + 1] - ty, beta, &gy[c]);
        if (dist[c] < dist[best]) best = c;
    }
    float l = dist[best];
    for (int c = 0; c < 4; c++) {
        const float w = (c == best ? 1.0f : 0.0f) + share;
        l += share * dist[c];
        d[2 * c] = w * gx[c] * inv;
        d[2 * c + 1] = w * gy[c] * inv;
    }
    const float z = o[8];
    const float label = (fabsf(o[2 * best] - tx) <= 0.125f && fabsf(o[2 * best + 1] - ty) <= 0.125f) ? 1.0f : 0.0f;
    const float sp = z > 0.0f ? z + log1pf(expf(-z)) : log1pf(expf(z));   // softplus
    l += sp - label * z;
    d[8] = (1.0f / (1.0f + expf(-z)) - label) * inv;
    for (int c = 9; c < rivnet::OUT; c++) d[c] = 0.0f;
    loss[i] = l;
}

__global__ void k_to_bf16(const float* __restrict__ in, size_t n, uint16_t* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// Adam on the f32 master parameters; the bf16 copy follows. mask[i] == 0 keeps entry i fixed
// (the zero padding of conv1's K and of the head's unused outputs).
__global__ void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v, const float* __restrict__ g,
                       const uint8_t* __restrict__ mask, size_t n, float lr, float b1, float b2, float eps, float c1, float c2,
                       uint16_t* __restrict__ wb) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !mask[i]) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi; v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    wb[i] = f2bf(wi);
}

// ============================================================================================
// Fused inference: 16 blocks per workgroup, every activation in LDS
// ============================================================================================
// Wave w runs the convolutions of blocks 2w and 2w+1 of its group, one at a time, in its own LDS
// (the 32x32x4 input in bf16, conv1 out 16x16x16, conv2 out 8x8x16); conv3 writes the block's
// row of the group's fc1 input. Then the eight waves split fc1's 128 outputs (one 16-column tile
// each) and wave 0 runs the head and writes the proposals. 163,328 B (159.5 KiB) of LDS, 512 B
// under a CU's 160 KiB: one workgroup of 8 waves per CU, and no room for a wider tile or pitch. Each pixel is converted to bf16 once; a conv fragment is one 8-byte LDS read.
namespace {
constexpr int kInferWaves = 8;
// Row pitch of a wave's input tile in uint16 entries: 32 pixels x 4 channels and one pixel of
// padding, so that the eight rows one patch store covers start in different LDS banks.
constexpr int kX0Pitch = 32 * 4 + 4;
constexpr int kWaveLds = 32 * kX0Pitch + 4096 + 1024;   // uint16 entries: input, conv1 out, conv2 out
constexpr int kInferLds = kInferWaves * kWaveLds * 2 + 16 * rivnet::FC_IN * 2 + 16 * rivnet::FC_H * 2 + 16 * 16 * 4;
static_assert(kInferLds <= 160 * 1024, "the inference workgroup's LDS exceeds a CU's 160 KiB");

__device__ __forceinline__ bool patch_interior(const rivnet::Source& S, int bx, int by) {
    const int x0 = bx * 16 - 8, y0 = by * 16 - 8;
    return x0 >= 0 && y0 >= 0 && x0 + 32 <= S.W && y0 + 32 <= S.H;
}
// An interior block's patches as aligned dwords: lane's dword u < 4 holds current-frame pixels
// (row, c4 .. c4+3) with q = lane + 64u, row = q >> 3, c4 = 4 (q & 7); dword u + 4 the same
// pixels of the reference.
__device__ __forceinline__ void patch_fetch(const rivnet::Source& S, int bx, int by, uint32_t (&d)[8], int lane) {
    const int x0 = bx * 16 - 8, y0 = by * 16 - 8;
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int q = lane + 64 * (u & 3), row = q >> 3, c4 = (q & 7) * 4;
        d[u] = *(const uint32_t*)((u >= 4 ? S.ref : S.cur) + (size_t)(y0 + row) * S.W + x0 + c4);
    }
}
__device__ __forceinline__ uint2 px4(int c, int r) {   // the input's 4 channels of one pixel
    const uint32_t a = f2bf((float)(c - 128) * (1.0f / 64.0f)), b = f2bf((float)(r - 128) * (1.0f / 64.0f));
    const uint32_t d = f2bf((float)(c - r) * (1.0f / 32.0f));
    return make_uint2(a | (b << 16), d);
}
__device__ __forceinline__ void patch_store(const uint32_t (&d)[8], uint16_t* x, int lane) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int q = lane + 64 * u, row = q >> 3, c4 = (q & 7) * 4;
#pragma unroll
        for (int i = 0; i < 4; i++)
            *(uint2*)(x + row * kX0Pitch + (c4 + i) * 4) = px4((int)((d[u] >> (8 * i)) & 0xffu), (int)((d[u + 4] >> (8 * i)) & 0xffu));
    }
}
__device__ __forceinline__ void patch_edge(const rivnet::Source& S, int bx, int by, uint16_t* x, int lane) {
    const int x0 = bx * 16 - 8, y0 = by * 16 - 8;
    for (int p = lane; p < 1024; p += 64) {
        const int yy = min(max(y0 + (p >> 5), 0), S.H - 1), xx = min(max(x0 + (p & 31), 0), S.W - 1);
        const size_t o = (size_t)yy * S.W + xx;
        *(uint2*)(x + (p >> 5) * kX0Pitch + (p & 31) * 4) = px4(S.cur[o], S.ref[o]);
    }
}

// The 16-channel activations (conv1's 16x16 output, conv2's 8x8 output) keep each pixel's 16
// channels together, with channel group g (channels 4g .. 4g+3) of pixel (x, y) in slot
// g ^ act_swz(x, y): the 16 pixels one conv fragment reads (stride 2 in x, over two or four rows,
// 64 bytes apart unswizzled) then fall in different LDS banks.
template <int HI>
__device__ __forceinline__ int act_swz(int x, int y) {
    return HI == 16 ? (((x >> 3) & 1) | (((y >> 1) & 1) << 1)) : ((y >> 1) & 3);
}
// A fragment of a conv layer from its NHWC bf16 input x (C = 4: the input tile, row pitch
// kX0Pitch, tap 4 ks + (lane >> 4), all 4 channels; C = 16: swizzled activations, tap ks, channels
// 4 (lane >> 4) .. +3), output row m: one 8-byte LDS read, zero past tap 8 and outside the input.
template <int C, int HI>
__device__ __forceinline__ v4s conv_a(const uint16_t* x, int m, int ks, int lane) {
    constexpr int HO = HI / 2;
    const int oy = m / HO, ox = m % HO;
    const int tap = C == 4 ? 4 * ks + (lane >> 4) : ks;
    const int c0 = C == 4 ? 0 : 4 * (lane >> 4);
    const int iy = 2 * oy + tap / 3 - 1, ix = 2 * ox + tap % 3 - 1;
    // the read is unconditional (from entry 0 when outside), so that a tile's reads issue together
    // instead of each waiting behind a branch; the select then zeroes what lies outside
    const bool ok = tap < 9 && (unsigned)iy < (unsigned)HI && (unsigned)ix < (unsigned)HI;
    const int off = C == 4 ? iy * kX0Pitch + ix * 4 : (iy * HI + ix) * 16 + 4 * ((c0 >> 2) ^ act_swz<HI>(ix, iy));
    uint2 v = *(const uint2*)(x + (ok ? off : 0));
    v.x = ok ? v.x : 0u;
    v.y = ok ? v.y : 0u;
    return __builtin_bit_cast(v4s, v);
}
// B fragment: W[k][n] for k = 16 ks + 4 (lane >> 4) + j, n = n0 + (lane & 15).
__device__ __forceinline__ v4s w_frag(const uint16_t* w, int N, int ks, int n0, int lane) {
    v4s b;
    const int k = 16 * ks + 4 * (lane >> 4), n = n0 + (lane & 15);
#pragma unroll
    for (int j = 0; j < 4; j++) b[j] = (short)w[(k + j) * N + n];
    return b;
}
// D rows 4 (lane >> 4) + j (pixels m0 + row of an HO x HO map), column lane & 15 (channel):
// bf16(ReLU(sum + bias)) to the swizzled 16-channel activations out
template <int HO>
__device__ __forceinline__ void store_act(v4f acc, const float* bias, int m0, uint16_t* out, int lane) {
    const int n = lane & 15;
    const float bn = bias[n];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const int m = m0 + 4 * (lane >> 4) + j;
        out[m * 16 + 4 * ((n >> 2) ^ act_swz<HO>(m % HO, m / HO)) + (n & 3)] = f2bf(fmaxf(acc[j] + bn, 0.0f));
    }
}
// D rows 4 (lane >> 4) + j, column lane & 15: bf16(ReLU(sum + bias)) to out[(m0 + row) * ldo + col]
__device__ __forceinline__ void store_relu(v4f acc, const float* bias, int n0, int m0, uint16_t* out, int ldo, int lane) {
    const int n = n0 + (lane & 15);
    const float bn = bias[n];
#pragma unroll
    for (int j = 0; j < 4; j++) out[(m0 + 4 * (lane >> 4) + j) * ldo + n] = f2bf(fmaxf(acc[j] + bn, 0.0f));
}
}  // namespace

// One workgroup of 64 per fragment.
__global__ __launch_bounds__(64) void k_pack_frags(rivnet::Weights w, uint2* __restrict__ pk) {
    const int f = blockIdx.x, lane = threadIdx.x;
    const uint16_t* src = w.w1;
    int N = 16, ks = f, n0 = 0;
    if (f >= rivnet::FRAG_W5) { src = w.w5; N = rivnet::OUT; ks = f - rivnet::FRAG_W5; }
    else if (f >= rivnet::FRAG_W4) { src = w.w4; N = rivnet::FC_H; ks = (f - rivnet::FRAG_W4) % 17; n0 = 16 * ((f - rivnet::FRAG_W4) / 17); }
    else if (f >= rivnet::FRAG_W3) { src = w.w3; ks = f - rivnet::FRAG_W3; }
    else if (f >= rivnet::FRAG_W2) { src = w.w2; ks = f - rivnet::FRAG_W2; }
    pk[f * 64 + lane] = __builtin_bit_cast(uint2, w_frag(src, N, ks, n0, lane));
}

__global__ __launch_bounds__(512) void k_net_infer(const rivnet::Source* __restrict__ srcs, rivnet::Source s0, const int* __restrict__ samples, int n,
                                                   rivnet::Weights wt, float conf_logit, int sr, uint4* __restrict__ prop,
                                                   float* __restrict__ yout) {
    extern __shared__ __align__(16) uint16_t lds16[];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint16_t* x0 = lds16 + wv * kWaveLds;
    uint16_t* a1 = x0 + 32 * kX0Pitch;
    uint16_t* a2 = a1 + 4096;
    uint16_t* a3 = lds16 + kInferWaves * kWaveLds;  // [16][FC_IN]
    uint16_t* a4 = a3 + 16 * rivnet::FC_IN;         // [16][FC_H]
    float* yl = (float*)(a4 + 16 * rivnet::FC_H);   // [16][16]

    // conv weights as B fragments, loaded once per wave
    auto frag = [&](int f) { return __builtin_bit_cast(v4s, wt.pk[f * 64 + lane]); };
    v4s w1[3], w2[9], w3[9];
#pragma unroll
    for (int ks = 0; ks < 3; ks++) w1[ks] = frag(rivnet::FRAG_W1 + ks);
#pragma unroll
    for (int ks = 0; ks < 9; ks++) { w2[ks] = frag(rivnet::FRAG_W2 + ks); w3[ks] = frag(rivnet::FRAG_W3 + ks); }

    // fc1's and the head's weight fragments, and both blocks' sources, interior patches and
    // predictor inputs, are loaded up front: the computation below waits on no global memory
    v4s w4[rivnet::FC_IN / 16], w5[rivnet::FC_H / 16];
#pragma unroll
    for (int ks = 0; ks < rivnet::FC_IN / 16; ks++) w4[ks] = frag(rivnet::FRAG_W4 + 17 * wv + ks);
#pragma unroll
    for (int ks = 0; ks < rivnet::FC_H / 16; ks++) w5[ks] = wv == 0 ? frag(rivnet::FRAG_W5 + ks) : v4s{0, 0, 0, 0};
    rivnet::Source sq[2];
    int bxq[2], byq[2];
    uint32_t pf[2][8];
    bool inq[2];
    float pin[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int i = blockIdx.x * 16 + 2 * wv + q;
        sq[q] = s0;
        bxq[q] = byq[q] = 0;
        inq[q] = false;
        pin[q] = 0.0f;
        if (i >= n) continue;
        int si, blk;
        sample_of(samples, i, &si, &blk);
        if (srcs) sq[q] = srcs[si];
        bxq[q] = blk % sq[q].bw;
        byq[q] = blk / sq[q].bw;
        inq[q] = patch_interior(sq[q], bxq[q], byq[q]);
        if (inq[q]) patch_fetch(sq[q], bxq[q], byq[q], pf[q], lane);
        if (lane < 16) pin[q] = pred_input(sq[q], bxq[q], byq[q], lane);
    }
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int slot = 2 * wv + q, i = blockIdx.x * 16 + slot;
        uint16_t* row3 = a3 + slot * rivnet::FC_IN;
        if (i >= n) {   // past the end: a zero row, never stored
            for (int k = lane; k < rivnet::FC_IN; k += 64) row3[k] = 0;
            continue;
        }
        if (inq[q]) patch_store(pf[q], x0, lane);
        else patch_edge(sq[q], bxq[q], byq[q], x0, lane);
        if (lane < 16) row3[256 + lane] = f2bf(pin[q]);
        lds_wait();   // this wave's LDS writes done before its reads
        __builtin_amdgcn_wave_barrier();
        // conv1: 256 output pixels (16 tiles), K 48
#pragma unroll 4
        for (int tm = 0; tm < 16; tm++) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 3; ks++) acc = mfma16(conv_a<4, 32>(x0, tm * 16 + (lane & 15), ks, lane), w1[ks], acc);
            store_act<16>(acc, wt.b1, tm * 16, a1, lane);
        }
        lds_wait();
        __builtin_amdgcn_wave_barrier();
        // conv2: 64 output pixels (4 tiles), K 144
#pragma unroll
        for (int tm = 0; tm < 4; tm++) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 9; ks++) acc = mfma16(conv_a<16, 16>(a1, tm * 16 + (lane & 15), ks, lane), w2[ks], acc);
            store_act<8>(acc, wt.b2, tm * 16, a2, lane);
        }
        lds_wait();
        __builtin_amdgcn_wave_barrier();
        // conv3: 16 output pixels, K 144 -> features p * 16 + channel of the fc1 row
        {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int ks = 0; ks < 9; ks++) acc = mfma16(conv_a<16, 8>(a2, lane & 15, ks, lane), w3[ks], acc);
            store_relu(acc, wt.b3, 0, 0, row3, 16, lane);
        }
        lds_wait();
        __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // fc1: 16 rows x 128 columns, K 272; wave w takes columns 16w .. 16w+15
    {
        const int n0 = 16 * wv;
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < rivnet::FC_IN / 16; ks++) {
            const uint2 v = *(const uint2*)(a3 + (lane & 15) * rivnet::FC_IN + 16 * ks + 4 * (lane >> 4));
            v4s a;
            a[0] = (short)(v.x & 0xffff); a[1] = (short)(v.x >> 16); a[2] = (short)(v.y & 0xffff); a[3] = (short)(v.y >> 16);
            acc = mfma16(a, w4[ks], acc);
        }
        store_relu(acc, wt.b4, n0, 0, a4, rivnet::FC_H, lane);
    }
    __syncthreads();
    if (wv != 0) return;
    {   // head: 16 rows x 16 columns, K 128
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < rivnet::FC_H / 16; ks++) {
            const uint2 v = *(const uint2*)(a4 + (lane & 15) * rivnet::FC_H + 16 * ks + 4 * (lane >> 4));
            v4s a;
            a[0] = (short)(v.x & 0xffff); a[1] = (short)(v.x >> 16); a[2] = (short)(v.y & 0xffff); a[3] = (short)(v.y >> 16);
            acc = mfma16(a, w5[ks], acc);
        }
        const int col = lane & 15;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = 4 * (lane >> 4) + j, i = blockIdx.x * 16 + r;
            const float y = acc[j] + wt.b5[col];
            yl[r * 16 + col] = y;
            if (yout && i < n) yout[(size_t)i * rivnet::OUT + col] = y;
        }
    }
    lds_wait();
    __builtin_amdgcn_wave_barrier();
    const int i = blockIdx.x * 16 + lane;
    if (lane < 16 && i < n && prop) {
        const float* y = yl + lane * 16;
        uint32_t c[4];
        for (int k = 0; k < 4; k++) {
            const int cx = (int)fminf(fmaxf(rintf(8.0f * y[2 * k]), (float)-sr), (float)sr);
            const int cy = (int)fminf(fmaxf(rintf(8.0f * y[2 * k + 1]), (float)-sr), (float)sr);
            c[k] = (uint32_t)(uint8_t)cx | ((uint32_t)(uint8_t)cy << 8);
        }
        prop[i] = make_uint4(c[0] | (c[1] << 16), c[2] | (c[3] << 16), y[8] >= conf_logit ? 1u : 0u, __float_as_uint(y[8]));
    }
}

// ============================================================================================
// Launch wrappers
// ============================================================================================
namespace rivnet {

size_t infer_lds_bytes() { return kInferLds; }

void gather(hipStream_t st, const Source* srcs, const int* samples, int n, uint16_t* x0, uint16_t* a3) {
    if (n > 0) k_net_gather<<<n, 256, 0, st>>>(srcs, samples, n, x0, a3);
}
void im2col(hipStream_t st, const uint16_t* x, int B, int Hi, int C, int KP, uint16_t* out) {
    const size_t total = (size_t)B * (Hi / 2) * (Hi / 2) * KP;
    const int grid = (int)std::min<size_t>((total + 255) / 256, 65536);
    if (total) k_im2col<<<grid, 256, 0, st>>>(x, B, Hi, C, KP, out);
}
void gemm(hipStream_t st, const Gemm& g) {
    const int tiles = ((g.M + 15) / 16) * ((g.N + 15) / 16);
    dim3 grid((tiles + 3) / 4, g.splits);
    if (tiles > 0) k_gemm<<<grid, 256, 0, st>>>(g);
}
void sum_splits(hipStream_t st, const float* part, int splits, int count, float* out) {
    if (count > 0) k_sum_splits<<<(count + 255) / 256, 256, 0, st>>>(part, splits, count, out);
}
void col2im_relu(hipStream_t st, const float* dcol, int B, int Hi, int C, int KP, const float* z, uint16_t* dz) {
    const size_t total = (size_t)B * Hi * Hi * C;
    const int grid = (int)std::min<size_t>((total + 255) / 256, 65536);
    if (total) k_col2im_relu<<<grid, 256, 0, st>>>(dcol, B, Hi, C, KP, z, dz);
}
void relu_bwd(hipStream_t st, const float* da, const float* z, int M, int N, int mg, long long ms, int ldc, uint16_t* dz) {
    const size_t total = (size_t)M * N;
    if (total) k_relu_bwd<<<(unsigned)((total + 255) / 256), 256, 0, st>>>(da, z, M, N, mg, ms, ldc, dz);
}
void colsum(hipStream_t st, const uint16_t* dz, int M, int N, float* out) {
    if (N > 0) k_colsum<<<N, 256, 0, st>>>(dz, M, N, out);
}
void loss(hipStream_t st, const float* y, const int8_t* tgt, int n, float* dy, float* l) {
    if (n > 0) k_net_loss<<<(n + 255) / 256, 256, 0, st>>>(y, tgt, n, dy, l);
}
void to_bf16(hipStream_t st, const float* in, size_t n, uint16_t* out) {
    if (n) k_to_bf16<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(in, n, out);
}
void adam(hipStream_t st, float* w, float* m, float* v, const float* g, const uint8_t* mask, size_t n, float lr, float b1, float b2,
          float eps, float c1, float c2, uint16_t* wb) {
    if (n) k_adam<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(w, m, v, g, mask, n, lr, b1, b2, eps, c1, c2, wb);
}
hipError_t pack_fragments(hipStream_t st, const Weights& w, uint2* pk) {
    k_pack_frags<<<NFRAG, 64, 0, st>>>(w, pk);
    return hipGetLastError();
}

hipError_t infer(hipStream_t st, const Source* srcs, const Source& s0, const int* samples, int n, const Weights& w, float conf_logit, int sr,
                 uint4* prop, float* yout) {
    if (n > 0 && !w.pk) return hipErrorInvalidValue;   // no fragment-order weights
    if (n > 0) k_net_infer<<<(n + 15) / 16, 64 * kInferWaves, kInferLds, st>>>(srcs, s0, samples, n, w, conf_logit, sr, prop, yout);
    return hipGetLastError();
}

}  // namespace rivnet
