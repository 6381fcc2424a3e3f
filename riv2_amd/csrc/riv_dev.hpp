// Device-side helpers shared by the gfx950 kernels of the RIV backend.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include "riv_common.hpp"

namespace rivdev {

template <typename T> __device__ __forceinline__ T dclamp(T v, T lo, T hi) { return v < lo ? lo : (v > hi ? hi : v); }

// ---- Rust numeric semantics ----
__device__ __forceinline__ int16_t f32_to_i16_sat(float f) {
    if (f != f) return 0;
    if (f <= -32768.0f) return -32768;
    if (f >= 32767.0f) return 32767;
    return (int16_t)f;
}
__device__ __forceinline__ int32_t f32_to_i32_sat(float f) {
    if (f != f) return 0;
    if (f <= -2147483648.0f) return INT32_MIN;
    if (f >= 2147483648.0f) return INT32_MAX;
    return (int32_t)f;
}
__device__ __forceinline__ uint16_t f32_to_u16_sat(float f) {
    if (f != f || f <= 0.0f) return 0;
    if (f >= 65535.0f) return 65535;
    return (uint16_t)f;
}
__device__ __forceinline__ int32_t wadd(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
__device__ __forceinline__ int32_t wsub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }
__device__ __forceinline__ int32_t wmul(int32_t a, int32_t b) { return (int32_t)((uint32_t)a * (uint32_t)b); }
__device__ __forceinline__ int32_t wshl(int32_t a, int s) { return (int32_t)((uint32_t)a << s); }

// ---- glibc-identical logf / exp2f ----
// Restatement of the ARM optimized-routines single-precision logf and exp2f (as shipped in
// glibc >= 2.28, sysdeps/ieee754/flt-32/e_logf.c, e_exp2f.c): double-precision evaluation
// with a 16-entry (logf) / 32-entry (exp2f) table and a degree-3 polynomial. Reproducing the
// host libm bit for bit matters because the reference's adaptive quantiser calls f32::ln and
// f32::exp2, which lower to the host's logf/exp2f. Exhaustively checked against glibc over
// the whole AQ domain (every reachable variance for logf, every float |x| <= 1.25 for exp2f).
// Only the finite-positive / |x| < 128 paths are implemented; callers stay inside them.
__device__ __forceinline__ double u2d(uint64_t u) { return __longlong_as_double((long long)u); }
__device__ __forceinline__ uint64_t d2u(double d) { return (uint64_t)__double_as_longlong(d); }

__device__ __forceinline__ float glibc_logf(float x) {
    const double invc[16] = {0x1.661ec79f8f3bep+0, 0x1.571ed4aaf883dp+0, 0x1.49539f0f010bp+0, 0x1.3c995b0b80385p+0,
                             0x1.30d190c8864a5p+0, 0x1.25e227b0b8eap+0, 0x1.1bb4a4a1a343fp+0, 0x1.12358f08ae5bap+0,
                             0x1.0953f419900a7p+0, 0x1p+0,             0x1.e608cfd9a47acp-1, 0x1.ca4b31f026aap-1,
                             0x1.b2036576afce6p-1, 0x1.9c2d163a1aa2dp-1, 0x1.886e6037841edp-1, 0x1.767dcf5534862p-1};
    const double logc[16] = {-0x1.57bf7808caadep-2, -0x1.2bef0a7c06ddbp-2, -0x1.01eae7f513a67p-2, -0x1.b31d8a68224e9p-3,
                             -0x1.6574f0ac07758p-3, -0x1.1aa2bc79c81p-3,   -0x1.a4e76ce8c0e5ep-4, -0x1.1973c5a611cccp-4,
                             -0x1.252f438e10c1ep-5, 0x0p+0,                0x1.aa5aa5df25984p-5,  0x1.c5e53aa362eb4p-4,
                             0x1.526e57720db08p-3,  0x1.bc2860d22477p-3,   0x1.1058bc8a07ee1p-2,  0x1.4043057b6ee09p-2};
    const double A0 = -0x1.00ea348b88334p-2, A1 = 0x1.5575b0be00b6ap-2, A2 = -0x1.ffffef20a4123p-2;
    const double Ln2 = 0x1.62e42fefa39efp-1;
    uint32_t ix = __float_as_uint(x);
    uint32_t tmp = ix - 0x3f330000u;
    int i = (int)((tmp >> 19) % 16);
    int k = (int32_t)tmp >> 23;
    uint32_t iz = ix - (tmp & 0xff800000u);
    double z = (double)__uint_as_float(iz);
    double r = __dsub_rn(__dmul_rn(z, invc[i]), 1.0);
    double y0 = __dadd_rn(logc[i], __dmul_rn((double)k, Ln2));
    double r2 = __dmul_rn(r, r);
    double y = __dadd_rn(__dmul_rn(A1, r), A2);
    y = __dadd_rn(__dmul_rn(A0, r2), y);
    y = __dadd_rn(__dmul_rn(y, r2), __dadd_rn(y0, r));
    return (float)y;
}
__device__ __forceinline__ float glibc_exp2f(float x) {
    const uint64_t T[32] = {
        0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
        0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
        0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
        0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
        0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
        0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
        0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
        0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull};
    const double C0 = 0x1.c6af84b912394p-5, C1 = 0x1.ebfce50fac4f3p-3, C2 = 0x1.62e42ff0c52d6p-1;
    const double SHIFT = 0x1.8p+52 / 32;
    double xd = (double)x;
    double kd = __dadd_rn(xd, SHIFT);
    uint64_t ki = d2u(kd);
    kd = __dsub_rn(kd, SHIFT);
    double r = __dsub_rn(xd, kd);
    uint64_t t = T[ki % 32];
    t += ki << (52 - 5);
    double s = u2d(t);
    double z = __dadd_rn(__dmul_rn(C0, r), C1);
    double r2 = __dmul_rn(r, r);
    double y = __dadd_rn(__dmul_rn(C2, r), 1.0);
    y = __dadd_rn(__dmul_rn(z, r2), y);
    y = __dmul_rn(y, s);
    return (float)y;
}

// ---- packed-byte helpers ----
// 4 bytes starting at an arbitrary byte offset of a dword-aligned LDS buffer
__device__ __forceinline__ uint32_t lds_read4(const uint32_t* base, int off) {
    uint32_t lo = base[off >> 2], hi = base[(off >> 2) + 1];
    return __builtin_amdgcn_alignbyte(hi, lo, (uint32_t)(off & 3));
}
// per-byte floor((a + b) / 2)
__device__ __forceinline__ uint32_t avg2f(uint32_t a, uint32_t b) { return __builtin_amdgcn_lerp(a, b, 0u); }
// per-byte floor((a + b + c + d) / 4): with h1 = floor((a+b)/2), h2 = floor((c+d)/2), it is
// floor((h1 + h2 + carry) / 2), carry = 1 where both a+b and c+d are odd (v_lerp_u8 adds bit 8i of
// its third operand to byte i before halving)
__device__ __forceinline__ uint32_t avg4f(uint32_t a, uint32_t b, uint32_t c, uint32_t d) {
    return __builtin_amdgcn_lerp(__builtin_amdgcn_lerp(a, b, 0u), __builtin_amdgcn_lerp(c, d, 0u), (a ^ b) & (c ^ d) & 0x01010101u);
}
__device__ __forceinline__ uint32_t sad4(uint32_t a, uint32_t b, uint32_t acc) { return __builtin_amdgcn_sad_u8(a, b, acc); }

// ---- wave64 reduction (DPP within rows of 16, readlane across rows) ----
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, true);    // quad_perm [1,0,3,2]
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, true);    // quad_perm [2,3,0,1]
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x141, 0xF, 0xF, true);   // row_half_mirror
    v += (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x140, 0xF, 0xF, true);   // row_mirror
    uint32_t a = (uint32_t)__builtin_amdgcn_readlane((int)v, 15), b = (uint32_t)__builtin_amdgcn_readlane((int)v, 31);
    uint32_t c = (uint32_t)__builtin_amdgcn_readlane((int)v, 47), d = (uint32_t)__builtin_amdgcn_readlane((int)v, 63);
    return a + b + c + d;
}

// ---- transform tables (closed-form DCT bases, identical to the reference tables) ----
extern __constant__ int32_t c_B16[16][16];   // [sample n][frequency k]
extern __constant__ int32_t c_B8[8][8];

}  // namespace rivdev
