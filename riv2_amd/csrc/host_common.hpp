// Host-side helpers shared by the encoder and decoder drivers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>
#include "../../include/riv_gpu.h"
#include "riv_common.hpp"

namespace rivhost {

void set_error(const std::string& msg);
const char* last_error();

// Selects `device` and uploads the per-device constant tables (DCT bases) once.
int ensure_device(int device);

#define RIV_HIP(call)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess) {                                                                     \
            rivhost::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                  \
            return RIV_E_GPU;                                                                       \
        }                                                                                           \
    } while (0)

// quant_step_from_quality (reitero_residual/src/residual.rs:13-23)
float quant_step(uint8_t quality);
// deblock_level_from_quant_step (reitero_video_common/src/deblock.rs:34-36)
int deblock_level(float qs);
// build_intra_y_quant_table / build_intra_uv_quant_table (residual.rs:58-96)
void intra_tables(float qs, uint16_t* yt, uint16_t* ct);

// LambdaEntry (reitero_encode/src/rdo.rs:23-86) and its skip decision (:140-178).
struct RdoModel {
    double lambda, retention, slope, intercept, skip_bits, keep_bits;
    uint8_t skip_threshold;
    RdoModel(uint8_t quality, double mult, uint8_t skip_threshold);
    bool skip(int64_t sad, bool integer_aligned) const;
};

// (dx, dy, sx, sy) of a motion vector
struct Mv4 {
    int8_t v[4];
    bool operator==(const Mv4& o) const { return std::memcmp(v, o.v, 4) == 0; }
    int hx() const { return v[0] * 2 + v[2]; }
    int hy() const { return v[1] * 2 + v[3]; }
};

// Predictors of block (bx, by) from the reconstructed MVs of the current frame (the first
// `known` entries, raster order) and the previous frame (mv_predictor.rs:59-272).
struct Neighbours {
    bool has_l = false, has_t = false, has_tr = false, has_tl = false, has_tp = false;
    Mv4 l{}, t{}, tr{}, tl{}, tp{};
    Mv4 nearest{}, nearp{}, temporal{};
};
Neighbours neighbours(const Mv4* cur, int known, const Mv4* prev, int bw, int bh, int bx, int by);

// Container (reitero_video_common/src/format.rs)
void put_header(std::vector<uint8_t>& out, const riv_video_header& h);
inline void put_u32(std::vector<uint8_t>& b, uint32_t v) { for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i))); }
inline void put_u64(std::vector<uint8_t>& b, uint64_t v) { for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i))); }
inline uint32_t get_u32(const uint8_t* p) { return (uint32_t)p[0] | ((uint32_t)p[1] << 8) | ((uint32_t)p[2] << 16) | ((uint32_t)p[3] << 24); }
inline uint64_t get_u64(const uint8_t* p) { return (uint64_t)get_u32(p) | ((uint64_t)get_u32(p + 4) << 32); }

// device allocation owned by a driver object
struct DevBuf {
    void* p = nullptr;
    size_t n = 0;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t bytes) {
        if (p) { (void)hipFree(p); p = nullptr; }
        n = bytes;
        return hipMalloc(&p, bytes ? bytes : 1);
    }
    template <typename T> T* as() const { return (T*)p; }
};
// pinned host allocation that kernels may also write (mapped; d is its device-side address)
struct HostBuf {
    void* p = nullptr;
    void* d = nullptr;
    ~HostBuf() { if (p) (void)hipHostFree(p); }
    hipError_t alloc(size_t bytes) {
        if (p) { (void)hipHostFree(p); p = nullptr; d = nullptr; }
        hipError_t r = hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocMapped);
        if (r != hipSuccess) return r;
        // the host's first touch of a page costs about 1.5 us; pay it here, not in the frame loop
        std::memset(p, 0, bytes ? bytes : 1);
        return hipHostGetDevicePointer(&d, p, 0);
    }
    template <typename T> T* as() const { return (T*)p; }
    template <typename T> T* dev() const { return (T*)d; }
};

}  // namespace rivhost
