// Kernel-level seams of the C ABI: single-stage GPU entry points on host buffers, used to check
// each stage against the reference functions it replaces.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <mutex>
#include <vector>
#include "host_common.hpp"
#include "riv_kernels.hpp"

using namespace rivhost;

// Scratch of the allocation-free AQ entry point: the per-MB logs of one call and their mean
// (kAqNoAllocMaxBlocks + 1 floats used). Every device the library runs on holds its own copy
// until the process ends; host code reaches it only through the address hipGetSymbolAddress
// returns for the current device (the array's name is a host-side stand-in).
constexpr int kAqNoAllocMaxBlocks = RIV_AQ_NOALLOC_MAX_BLOCKS;
static __device__ __align__(256) float scratch[kAqNoAllocMaxBlocks + 64];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch[kMaxDevices] = {};

// The scratch array's address on `device`, looked up once per device with that device current.
int scratch_on(int device, float** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[device]) {
        void* p = nullptr;
        RIV_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch[device] = (float*)p;
    }
    *out = g_scratch[device];
    return RIV_OK;
}
}  // namespace

extern "C" {

int riv_gpu_build_predicted(int32_t dev, const uint8_t* y, const uint8_t* u, const uint8_t* v, uint32_t w, uint32_t h, const int8_t* mv4,
                            uint8_t* oy, uint8_t* ou, uint8_t* ov) {
    if (!y || !u || !v || !mv4 || !oy || !ou || !ov || w % 16 || h % 16 || !w || !h) return RIV_E_ARG;
    int r = ensure_device(dev);
    if (r) return r;
    const size_t ysz = (size_t)w * h, csz = ysz / 4, nb = ysz / 256;
    DevBuf ref, pred, mvs;
    RIV_HIP(ref.alloc(ysz + 2 * csz)); RIV_HIP(pred.alloc(ysz + 2 * csz)); RIV_HIP(mvs.alloc(4 * nb));
    uint8_t* R = ref.as<uint8_t>(); uint8_t* P = pred.as<uint8_t>();
    RIV_HIP(hipMemcpy(R, y, ysz, hipMemcpyHostToDevice)); RIV_HIP(hipMemcpy(R + ysz, u, csz, hipMemcpyHostToDevice));
    RIV_HIP(hipMemcpy(R + ysz + csz, v, csz, hipMemcpyHostToDevice)); RIV_HIP(hipMemcpy(mvs.p, mv4, 4 * nb, hipMemcpyHostToDevice));
    rivk::mc(0, R, R + ysz, R + ysz + csz, (int)w, (int)h, mvs.as<int8_t>(), 4, P, P + ysz, P + ysz + csz, nullptr, nullptr);
    RIV_HIP(hipGetLastError());
    RIV_HIP(hipMemcpy(oy, P, ysz, hipMemcpyDeviceToHost)); RIV_HIP(hipMemcpy(ou, P + ysz, csz, hipMemcpyDeviceToHost));
    RIV_HIP(hipMemcpy(ov, P + ysz + csz, csz, hipMemcpyDeviceToHost));
    return RIV_OK;
}

int riv_gpu_deblock_yuv420(int32_t dev, uint8_t* y, uint8_t* u, uint8_t* v, uint32_t w, uint32_t h, int32_t level, int32_t intra, const uint8_t* mask) {
    if (!y || !u || !v || w % 16 || h % 16 || !w || !h) return RIV_E_ARG;
    int r = ensure_device(dev);
    if (r) return r;
    const size_t ysz = (size_t)w * h, csz = ysz / 4, nb = ysz / 256;
    DevBuf f, m;
    RIV_HIP(f.alloc(ysz + 2 * csz));
    uint8_t* F = f.as<uint8_t>();
    RIV_HIP(hipMemcpy(F, y, ysz, hipMemcpyHostToDevice)); RIV_HIP(hipMemcpy(F + ysz, u, csz, hipMemcpyHostToDevice));
    RIV_HIP(hipMemcpy(F + ysz + csz, v, csz, hipMemcpyHostToDevice));
    if (mask) { RIV_HIP(m.alloc(nb)); RIV_HIP(hipMemcpy(m.p, mask, nb, hipMemcpyHostToDevice)); }
    rivk::deblock(0, F, F + ysz, F + ysz + csz, (int)w, (int)h, level, intra != 0, mask ? m.as<uint8_t>() : nullptr);
    RIV_HIP(hipGetLastError());
    RIV_HIP(hipMemcpy(y, F, ysz, hipMemcpyDeviceToHost)); RIV_HIP(hipMemcpy(u, F + ysz, csz, hipMemcpyDeviceToHost));
    RIV_HIP(hipMemcpy(v, F + ysz + csz, csz, hipMemcpyDeviceToHost));
    return RIV_OK;
}

int riv_gpu_me_tables(int32_t dev, const uint8_t* cur, const uint8_t* ref, uint32_t w, uint32_t h, int32_t sr, uint16_t* int_sad,
                      uint16_t* hp_sad) {
    if (!cur || !ref || !int_sad || !hp_sad || !w || !h || w % 16 || h % 16 || sr < 0 || sr > 32) return RIV_E_ARG;
    int r = ensure_device(dev);
    if (r) return r;
    const size_t ysz = (size_t)w * h, nb = ysz / 256;
    const int D = 2 * sr + 1, D2 = 4 * sr + 3, RB = rivk::table_region_bytes(sr), TSB = 2 * rivk::sad_table_stride(sr);
    DevBuf dc, dr, dt;
    RIV_HIP(dc.alloc(ysz)); RIV_HIP(dr.alloc(ysz)); RIV_HIP(dt.alloc((size_t)RB * nb));
    RIV_HIP(hipMemcpy(dc.p, cur, ysz, hipMemcpyHostToDevice)); RIV_HIP(hipMemcpy(dr.p, ref, ysz, hipMemcpyHostToDevice));
    uint8_t* T = dt.as<uint8_t>();
    rivk::sad_table(0, dc.as<uint8_t>(), dr.as<uint8_t>(), (int)w, (int)h, sr, (uint16_t*)T, RB / 2);
    rivk::hp_table(0, dc.as<uint8_t>(), dr.as<uint8_t>(), (int)w, (int)h, sr, (uint16_t*)(T + TSB), RB / 2);
    RIV_HIP(hipGetLastError());
    std::vector<uint8_t> t((size_t)RB * nb);
    RIV_HIP(hipMemcpy(t.data(), T, t.size(), hipMemcpyDeviceToHost));
    const int bw = (int)w / 16;
    for (size_t b = 0; b < nb; b++) {
        const uint16_t* it = (const uint16_t*)(t.data() + b * RB);
        const uint16_t* hp = (const uint16_t*)(t.data() + b * RB + TSB);
        std::copy(hp, hp + (size_t)D2 * D2, hp_sad + b * D2 * D2);
        // the search reads an interior block's integer SADs from the half-pel table's even entries
        const bool interior = rivk::mb_is_interior((int)(b % bw) * 16, (int)(b / bw) * 16, (int)w, (int)h, sr);
        for (int dy = -sr; dy <= sr; dy++)
            for (int dx = -sr; dx <= sr; dx++)
                int_sad[b * D * D + (dy + sr) * D + dx + sr] =
                    interior ? hp[(2 * dy + 2 * sr + 1) * D2 + 2 * dx + 2 * sr + 1] : it[(dy + sr) * D + dx + sr];
    }
    return RIV_OK;
}

static int dct_plane(int32_t dev, const int16_t* plane, uint32_t w, uint32_t h, int n8, const float* qs, float dz, int16_t* coeffs, int16_t* recon) {
    const int bs = n8 ? 8 : 16;
    if (!plane || !qs || !coeffs || !recon || !w || !h || w % bs || h % bs) return RIV_E_ARG;
    int r = ensure_device(dev);
    if (r) return r;
    const size_t n = (size_t)w * h, nb = n / (bs * bs);
    DevBuf dp, dq, dc, dr;
    RIV_HIP(dp.alloc(2 * n)); RIV_HIP(dq.alloc(4 * nb)); RIV_HIP(dc.alloc(2 * n)); RIV_HIP(dr.alloc(2 * n));
    RIV_HIP(hipMemcpy(dp.p, plane, 2 * n, hipMemcpyHostToDevice)); RIV_HIP(hipMemcpy(dq.p, qs, 4 * nb, hipMemcpyHostToDevice));
    rivk::dct_plane(0, dp.as<int16_t>(), (int)w, (int)h, n8, dq.as<float>(), dz, dc.as<int16_t>(), dr.as<int16_t>());
    RIV_HIP(hipGetLastError());
    RIV_HIP(hipMemcpy(coeffs, dc.p, 2 * n, hipMemcpyDeviceToHost)); RIV_HIP(hipMemcpy(recon, dr.p, 2 * n, hipMemcpyDeviceToHost));
    return RIV_OK;
}
int riv_gpu_dct8_aq_plane(int32_t dev, const int16_t* plane, uint32_t w, uint32_t h, const float* qs, float dz, int16_t* coeffs, int16_t* recon) {
    return dct_plane(dev, plane, w, h, 1, qs, dz, coeffs, recon);
}
int riv_gpu_dct16_aq_plane(int32_t dev, const int16_t* plane, uint32_t w, uint32_t h, const float* qs, float dz, int16_t* coeffs, int16_t* recon) {
    return dct_plane(dev, plane, w, h, 0, qs, dz, coeffs, recon);
}

int riv_gpu_aq_steps(int32_t dev, const uint8_t* y, uint32_t w, uint32_t h, float base, float* out) {
    if (!y || !out || w % 16 || h % 16 || !w || !h) return RIV_E_ARG;
    int r = ensure_device(dev);
    if (r) return r;
    const size_t ysz = (size_t)w * h, nb = ysz / 256;
    // the variance sums come from the MC kernel; feed it the plane as its own zero-MV prediction
    DevBuf ref, pred, mvs, var, lnv, steps;
    RIV_HIP(ref.alloc(ysz * 3 / 2)); RIV_HIP(pred.alloc(ysz * 3 / 2)); RIV_HIP(mvs.alloc(4 * nb));
    RIV_HIP(var.alloc(8 * nb)); RIV_HIP(lnv.alloc(4 * (nb + 1))); RIV_HIP(steps.alloc(4 * nb));
    RIV_HIP(hipMemset(ref.p, 0, ysz * 3 / 2)); RIV_HIP(hipMemset(mvs.p, 0, 4 * nb));
    RIV_HIP(hipMemcpy(ref.p, y, ysz, hipMemcpyHostToDevice));
    uint8_t* R = ref.as<uint8_t>(); uint8_t* P = pred.as<uint8_t>();
    rivk::mc(0, R, R + ysz, R + ysz + ysz / 4, (int)w, (int)h, mvs.as<int8_t>(), 4, P, P + ysz, P + ysz + ysz / 4, var.as<uint32_t>(), lnv.as<float>());
    rivk::aq(0, (int)nb, base, lnv.as<float>(), steps.as<float>());
    RIV_HIP(hipGetLastError());
    RIV_HIP(hipMemcpy(out, steps.p, 4 * nb, hipMemcpyDeviceToHost));
    return RIV_OK;
}

int riv_gpu_aq_steps_noalloc(int32_t dev, const uint8_t* y, uint32_t w, uint32_t h, float base, float* steps, void* stream) {
    if (!y || !steps || !w || !h || w % 16 || h % 16) { set_error("luma plane and steps are required, dimensions multiples of 16"); return RIV_E_ARG; }
    if (((uintptr_t)y & 3u) || ((uintptr_t)steps & 3u)) { set_error("luma plane and steps must be 4-byte aligned"); return RIV_E_ARG; }
    const size_t nb = (size_t)(w / 16) * (h / 16);
    if (nb > (size_t)kAqNoAllocMaxBlocks) {
        set_error("riv_gpu_aq_steps_noalloc: " + std::to_string(nb) + " macroblocks exceed its scratch of " +
                  std::to_string(kAqNoAllocMaxBlocks) + " (RIV_AQ_NOALLOC_MAX_BLOCKS)");
        return RIV_E_ARG;
    }
    // device selection only: no table upload, no allocation, nothing synchronous
    int n = 0;
    RIV_HIP(hipGetDeviceCount(&n));
    if (dev < 0 || dev >= n || dev >= kMaxDevices) { set_error("device_id out of range"); return RIV_E_ARG; }
    RIV_HIP(hipSetDevice(dev));
    float* lnv = nullptr;
    int r = scratch_on(dev, &lnv);
    if (r) return r;
    hipStream_t st = (hipStream_t)stream;
    rivk::aq_logs(st, y, (int)w, (int)h, lnv);   // writes every log the mean then reads: nothing to clear
    rivk::aq(st, (int)nb, base, lnv, steps);
    RIV_HIP(hipGetLastError());
    return RIV_OK;
}

int riv_gpu_aq_mean(int32_t dev, const float* logs, size_t n, float base, float* mean, float* steps) {
    if (!logs || !mean || !n || n > (size_t)INT32_MAX - 1) return RIV_E_ARG;
    int r = ensure_device(dev);
    if (r) return r;
    DevBuf lnv, st;
    RIV_HIP(lnv.alloc(4 * (n + 1))); RIV_HIP(st.alloc(4 * n));
    RIV_HIP(hipMemcpy(lnv.p, logs, 4 * n, hipMemcpyHostToDevice));
    rivk::aq(0, (int)n, base, lnv.as<float>(), st.as<float>());
    RIV_HIP(hipGetLastError());
    RIV_HIP(hipMemcpy(mean, lnv.as<float>() + n, 4, hipMemcpyDeviceToHost));
    if (steps) RIV_HIP(hipMemcpy(steps, st.p, 4 * n, hipMemcpyDeviceToHost));
    return RIV_OK;
}

int riv_gpu_logf_exp2f(int32_t dev, const float* in, size_t n, int32_t which, float* out) {
    if (!in || !out) return RIV_E_ARG;
    if (!n) return RIV_OK;
    int r = ensure_device(dev);
    if (r) return r;
    DevBuf di, dout;
    RIV_HIP(di.alloc(4 * n)); RIV_HIP(dout.alloc(4 * n));
    RIV_HIP(hipMemcpy(di.p, in, 4 * n, hipMemcpyHostToDevice));
    rivk::logf_exp2f(0, di.as<float>(), n, which, dout.as<float>());
    RIV_HIP(hipGetLastError());
    RIV_HIP(hipMemcpy(out, dout.p, 4 * n, hipMemcpyDeviceToHost));
    return RIV_OK;
}

}  // extern "C"
