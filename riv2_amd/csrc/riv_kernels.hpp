// Host-side launch wrappers of the gfx950 kernels (kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <vector>
#include "riv_common.hpp"

namespace rivk {
hipError_t upload_bases(const int32_t* b16, const int32_t* b8);
void rgb_to_yuv420(hipStream_t st, const uint8_t* rgb, int dw, int dh, int sw, int sh, uint8_t* Y, uint8_t* U, uint8_t* V);
void yuv420_to_rgb(hipStream_t st, const uint8_t* Y, const uint8_t* U, const uint8_t* V, int sw, int dw, int dh, uint8_t* rgb);
int sad_table_stride(int sr);
int hp_table_stride(int sr);
int table_region_bytes(int sr);
bool mb_is_interior(int x0, int y0, int W, int H, int sr);   // its integer SADs are the half-pel table's even entries
void sad_table(hipStream_t st, const uint8_t* cur, const uint8_t* ref, int W, int H, int sr, uint16_t* out, int stride);
void hp_table(hipStream_t st, const uint8_t* cur, const uint8_t* ref, int W, int H, int sr, uint16_t* out, int stride);
size_t search_lds_bytes(int sr, int waves, int nstreams, int nslot);
// Launch shape of the motion search for nb blocks over nstreams streams: workgroups (at most one
// per CU), waves per workgroup, window slots per wave (nres of them resident) and LDS bytes.
// False when the blocks do not fit in resident waves or the LDS budget. Workgroups have at most
// max_waves waves (16, the most a workgroup can have).
struct SearchShape { int waves, groups, nslot, nres; size_t lds; };
bool search_shape(int sr, int nb, int nstreams, int cus, size_t budget, SearchShape* out, int max_waves = 16);
// The launch shapes an autotuner may try for the same search (search_shape's first); every one
// reaches the same fixed point, so the choice changes only the time.
std::vector<SearchShape> search_shapes(int sr, int nb, int nstreams, int cus, size_t budget);
// With ev0 / ev1, the launch records them when the kernel starts / ends (profiling).
hipError_t me_search(hipStream_t st, const riv::SearchParams& p, int groups, int waves, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void mc(hipStream_t st, const uint8_t* rY, const uint8_t* rU, const uint8_t* rV, int W, int H, const int8_t* mvs, int mv_stride,
        uint8_t* pY, uint8_t* pU, uint8_t* pV, uint32_t* var_sums, float* lnv);
// lnv: nb logs written by mc, plus one float for their mean; one launch computes the mean and, when
// steps is not null, the steps
void aq(hipStream_t st, int nb, float base, float* lnv, float* steps);
// lnv: (W/16)*(H/16) logs of the luma plane Y's block variances (Y 4-byte aligned), aq's input
void aq_logs(hipStream_t st, const uint8_t* Y, int W, int H, float* lnv);
void logf_exp2f(hipStream_t st, const float* in, size_t n, int which, float* out);
// Each MB's quantiser step from lnv (its log variance; the frame's mean at lnv[nb], as aq leaves
// it) and base, also written to steps.
void inter_residual(hipStream_t st, const uint8_t* cY, const uint8_t* cU, const uint8_t* cV, const uint8_t* rY, const uint8_t* rU,
                    const uint8_t* rV, uint8_t* pY, uint8_t* pU, uint8_t* pV, int W, int H, const riv::BlockRec* rec, const float* lnv,
                    float base, float* steps, float dz, uint8_t* oY, uint8_t* oU, uint8_t* oV, uint8_t* main, uint8_t* over, uint8_t* mask,
                    uint8_t* host_opt_skip, uint8_t* dpack = nullptr, uint8_t* dcoded = nullptr);
// r*: the reference planes; p*: out, the motion-compensated prediction (with the search MV)
// main/over: coded MBs' packed coefficients, MB b's masks (and its values if there are at most
// PACK_MAIN_VALUES) at main + b * PACK_MAIN, else its values at over + b * PACK_STRIDE + 48
// mask: per MB the deblock filter flag (not optimised-skip, or a non-zero reconstructed MV; encoder.rs:831-845)
// dpack / dcoded (optional, device memory): each coded MB's full record (PACK_STRIDE bytes: the six
// masks, then the values) and per MB 1 when its residual is coded, for the GPU residual modelling;
// main and over may then be null (no host copy of the records)
// pack: riv::PACK_STRIDE bytes per MB (mapped host memory); the layout is in riv_common.hpp
void intra(hipStream_t st, const uint8_t* cY, const uint8_t* cU, const uint8_t* cV, int W, int H, const uint16_t* yt, const uint16_t* ct,
           uint8_t* oY, uint8_t* oU, uint8_t* oV, uint8_t* pack);
void inter_recon(hipStream_t st, const uint8_t* pY, const uint8_t* pU, const uint8_t* pV, int W, int H, const uint8_t* skip,
                 const float* steps, const int16_t* coeffs, uint8_t* oY, uint8_t* oU, uint8_t* oV);
void intra_recon(hipStream_t st, int W, int H, const uint16_t* yt, const uint16_t* ct, const int16_t* coeffs, uint8_t* oY, uint8_t* oU,
                 uint8_t* oV);
void dct_plane(hipStream_t st, const int16_t* plane, int w, int h, int n8, const float* qs, float dz, int16_t* coeffs, int16_t* recon);
void deblock(hipStream_t st, uint8_t* Y, uint8_t* U, uint8_t* V, int W, int H, int level, bool intra, const uint8_t* mask);

// Context modelling of a P-frame's residual symbols on the GPU (resmodel.hip): the frame's symbol
// sequence as the host coder would buffer it for rANS32, from inter_residual's dpack / dcoded.
struct ResModelBufs {
    int nb = 0;
    uint32_t cap = 0;               // symbols a frame may have; sym and the host buffer hold cap + 8
    const uint8_t* pack = nullptr;  // nb * PACK_STRIDE
    const uint8_t* coded = nullptr; // nb
    uint16_t* cnt = nullptr;        // 3 * nb: symbols per plane
    uint8_t* last = nullptr;        // 3 * nb: last non-zero position per plane
    uint8_t* ke = nullptr;          // 3 * nb: its class
    uint16_t* bcnt = nullptr;       // 7 * nb: non-zeros per band
    uint16_t* wst = nullptr;        // 6 * nb: per MB and mask word, the MB-relative position of its first significance symbol
    uint32_t* sigw = nullptr;       // 384 * nb: the significance chains' words (rows: luma position, then U, V)
    uint32_t* eobw = nullptr;       // 20 * nb: the EOB chains' words (rows: luma contexts 0..7, U and V 0..5)
    uint8_t* bad = nullptr;         // nb: a magnitude beyond the class table
    uint32_t* off = nullptr;        // nb: the MB's first symbol, relative to its tile of 1024 MBs
    uint32_t* boff = nullptr;       // 7 * nb: the MB's first event per band, relative to its tile
    uint32_t* tt = nullptr;         // 9 per tile: its totals (symbols, seven bands) and a bad flag
    uint32_t* tp = nullptr;         // 8 per tile: the totals of the tiles before it
    uint32_t* ev = nullptr;         // 384 * nb: band event lists
    uint16_t* sym = nullptr;        // cap + 8: the frame's symbols (device memory the caller copies out)
    uint32_t* hdr = nullptr;        // 16 words
};
// host_hdr[0]: the frame's symbols, host_hdr[1]: status (0: b.sym holds them; 1: a magnitude
// beyond the class table; 2: more than cap symbols: host_pack then holds the MBs' records, nb *
// PACK_STRIDE bytes); both in mapped host memory.
hipError_t res_model(hipStream_t st, const ResModelBufs& b, uint32_t* host_hdr, uint8_t* host_pack);
}  // namespace rivk
