// Learned MV-proposal network: shapes, parameter layout and the kernel launch wrappers
// (mvnet.hip). The host driver is mvnet_api.cpp.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>

namespace rivnet {

constexpr int C0 = 4;          // input channels (current, reference, difference, zero)
constexpr int F = 16;          // channels of every conv layer
constexpr int K1 = 48;         // conv1 K: 9 taps x 4 channels, padded to a multiple of 16
constexpr int K2 = 144;        // conv2 / conv3 K: 9 taps x 16 channels
constexpr int FC_IN = 272;     // 256 conv features + 16 predictor inputs
constexpr int FC_H = 128;
constexpr int OUT = 16;        // 9 used: four candidates (x, y) and the confidence logit
constexpr int OUT_USED = 9;

// Offsets (in floats) of each tensor in the flat parameter vector; weights are [K][N] row-major.
constexpr size_t OFF_W1 = 0, OFF_B1 = OFF_W1 + K1 * F;
constexpr size_t OFF_W2 = OFF_B1 + F, OFF_B2 = OFF_W2 + K2 * F;
constexpr size_t OFF_W3 = OFF_B2 + F, OFF_B3 = OFF_W3 + K2 * F;
constexpr size_t OFF_W4 = OFF_B3 + F, OFF_B4 = OFF_W4 + (size_t)FC_IN * FC_H;
constexpr size_t OFF_W5 = OFF_B4 + FC_H, OFF_B5 = OFF_W5 + (size_t)FC_H * OUT;
constexpr size_t NPARAM = OFF_B5 + OUT;

// One frame pair the network reads: luma planes in device memory (W x H, W a multiple of 16) and
// the previous P-frame's vectors (x, y as int8 at prev_mv + block * prev_stride), or null.
struct Source {
    const uint8_t* cur;
    const uint8_t* ref;
    const int8_t* prev_mv;
    int prev_stride;
    int W, H, bw, bh;
};

// The bf16 weights in MFMA B-fragment order for the fused inference kernel: fragment f holds,
// for lane l, W[16 ks + 4 (l >> 4) + j][n0 + (l & 15)], j = 0..3, as one 8-byte word at
// pk[64 f + l]. Fragments: conv1 ks 0..2, conv2 ks 0..8, conv3 ks 0..8, fc1 (n-tile t, ks)
// at FRAG_W4 + 17 t + ks, head ks 0..7.
constexpr int FRAG_W1 = 0, FRAG_W2 = 3, FRAG_W3 = 12, FRAG_W4 = 21, FRAG_W5 = 21 + 8 * 17, NFRAG = FRAG_W5 + 8;

// Device pointers of the weights (bf16, [K][N]), biases (f32) and, for inference, the weights in
// fragment order (pk, NFRAG x 64 words, made by pack_fragments from w1 .. w5).
struct Weights {
    const uint16_t *w1, *w2, *w3, *w4, *w5;
    const float *b1, *b2, *b3, *b4, *b5;
    const uint2* pk;
};

struct Gemm {
    const uint16_t* A; long long sam, sak;
    const uint16_t* B; long long sbk, sbn;
    int M, N, K, kc, splits;
    float* C;                 // f32 out (pre-activation), or the split partials
    long long mg, ms, ldc;    // out offset of (m, n): (m / mg) * ms + (m % mg) * ldc + n
    const float* bias;
    uint16_t* Cb;             // bf16 out (after ReLU when relu), or null
    int relu;
};

size_t infer_lds_bytes();
void gather(hipStream_t st, const Source* srcs, const int* samples, int n, uint16_t* x0, uint16_t* a3);
void im2col(hipStream_t st, const uint16_t* x, int B, int Hi, int C, int KP, uint16_t* out);
void gemm(hipStream_t st, const Gemm& g);
void sum_splits(hipStream_t st, const float* part, int splits, int count, float* out);
void col2im_relu(hipStream_t st, const float* dcol, int B, int Hi, int C, int KP, const float* z, uint16_t* dz);
void relu_bwd(hipStream_t st, const float* da, const float* z, int M, int N, int mg, long long ms, int ldc, uint16_t* dz);
void colsum(hipStream_t st, const uint16_t* dz, int M, int N, float* out);
void loss(hipStream_t st, const float* y, const int8_t* tgt, int n, float* dy, float* l);
void to_bf16(hipStream_t st, const float* in, size_t n, uint16_t* out);
void adam(hipStream_t st, float* w, float* m, float* v, const float* g, const uint8_t* mask, size_t n, float lr, float b1, float b2,
          float eps, float c1, float c2, uint16_t* wb);
// Fused forward over n samples (samples: n (source, block) pairs into srcs, or null for blocks
// 0..n-1 of s0, srcs then null too); prop[i]: the four candidates rounded to whole pixels and clamped to +-sr (x, y bytes,
// two per word), the confident flag (logit >= conf_logit) and the logit's bits; yout: the raw
// outputs [n][OUT], or null.
// Writes the fragment-order copy of w.w1 .. w.w5 to pk (NFRAG x 64 words).
hipError_t pack_fragments(hipStream_t st, const Weights& w, uint2* pk);
hipError_t infer(hipStream_t st, const Source* srcs, const Source& s0, const int* samples, int n, const Weights& w, float conf_logit, int sr,
                 uint4* prop, float* yout);

}  // namespace rivnet
