// Structures shared by the host driver and the gfx950 kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace riv {

// MV prediction modes (reitero_residual/src/mv_predictor.rs:5-33, as_u8 order)
enum Mode : uint8_t { M_ZERO = 0, M_NEAREST = 1, M_NEAR = 2, M_TOPRIGHT = 3, M_TOPLEFT = 4, M_NEW = 5, M_TEMPORAL = 6 };

// One record per 16x16 macroblock, written by the mode decision after the motion search.
// Bytes 4..7 (the reconstructed MV) seed the next frame's temporal candidates.
struct BlockRec {
    int8_t s[4];        // search MV (dx, dy, sx, sy), canonical (motion/scalar.rs:619-634)
    int8_t r[4];        // reconstructed MV as the decoder rebuilds it (encoder.rs:667-718)
    uint8_t mode;       // Mode
    uint8_t new_base;   // 0 nearest, 1 near, 2 top-right, 3 top-left, 4 temporal
    int8_t raw_dx, raw_dy;   // NEW delta before the frame bias (encoder.rs:651-660)
    uint16_t sad;       // best half-pel SAD (motion/scalar.rs:759-839)
    uint8_t rdo_skip;   // RdoContext::decide (rdo.rs:140-178)
    uint8_t flags;
};
static_assert(sizeof(BlockRec) == 16, "BlockRec layout");

// Quantised coefficients of one MB as the host entropy stage receives them: six 64-bit words
// marking the non-zero coefficients (bit i of word k: coefficient 64k+i of the MB's 384, Y then
// U then V, row-major), then the non-zero values as int16 in index order. Slots are PACK_STRIDE
// bytes apart; only the used part of a slot is written.
constexpr int PACK_STRIDE = 48 + 768;
// Inter frames split the record: the masks, and the values when there are at most
// PACK_MAIN_VALUES, go to a compact slot of PACK_MAIN bytes; more values go to the MB's
// PACK_STRIDE slot, after its (unwritten) masks.
constexpr int PACK_MAIN_VALUES = 40;
constexpr int PACK_MAIN = 48 + 2 * PACK_MAIN_VALUES;

// Pass barrier of the motion search: per pass, kBarrierLeaves leaf counters (workgroup w arrives
// at leaf w % kBarrierLeaves) and a root counter; then one release word. A bank holds
// barrier_words(maxp) counters; the search alternates between two banks.
constexpr int kBarrierLeaves = 16;
__host__ __device__ inline int barrier_words(int maxp) { return maxp * (kBarrierLeaves + 1) + 1; }

// One frame of one stream in a motion-search launch. A launch covers the frames of one or more
// independent streams; their blocks are numbered consecutively, stream by stream (blk0).
struct SearchStream {
    const uint8_t* cur;          // current luma plane, W x H
    const uint8_t* ref;          // reference luma plane (previous reconstruction), W x H
    const int8_t* prev_mv;       // previous P-frame's reconstructed MVs (dx,dy,sx,sy) or nullptr
    BlockRec* rec;               // per block out: the mode decision's record
    BlockRec* host_rec;          // the same records, written to mapped host memory
    const uint32_t* rdo_bits;    // bit s set -> RDO chooses skip at SAD s (integer-aligned MV)
    int* passes;                 // out: [0] passes the launch ran, [1] left as the host cleared it (0) unless the
                                 // stream's search failed: 1 barrier timeout, 3 pass limit, 2 mode-decision
                                 // identity violated (the host reads it after the frame)
    int W, H, bw, bh;
    int blk0;                    // number of the stream's first block in the launch
    int prev_stride;             // bytes between consecutive MVs of prev_mv
    int rdo_max;                 // largest SAD that can skip, -1 when skipping is disabled
    int pad_;
    float* lnv;                  // out: per block, the log variance of its luma prediction (AQ), or nullptr
};
static_assert(sizeof(SearchStream) == 96, "SearchStream layout");

// LDS layout of the motion search: the MV-cost table (TD x TD ints, TD = 2*sr + 4), 4 flag words,
// the launch's stream descriptors, then per wave its window slots, each an NR x ND dword window
// (NR = 2*sr + 18 rows of ND dwords) and the block's 16 rows (64 dwords).
// The row pitch ND is ((2*sr + 20) >> 2) + 1 dwords rounded up to odd: the 16 lanes of a row
// group read 16 consecutive window rows, which an odd pitch spreads over distinct LDS banks.
__host__ __device__ inline int search_pitch(int sr) { return (((2 * sr + 20) >> 2) + 1) | 1; }
__host__ __device__ inline int search_window_dwords(int sr) {
    return (((2 * sr + 18) * search_pitch(sr)) + 3) & ~3;
}
__host__ __device__ inline int search_slot_dwords(int sr) { return search_window_dwords(sr) + 64; }
__host__ __device__ inline size_t search_slots_offset(int sr, int nstreams) {
    const int TD = 2 * sr + 4;
    return ((((size_t)((TD * TD + 3) & ~3) + 4) * 4 + (size_t)nstreams * sizeof(SearchStream)) + 15) & ~(size_t)15;
}

// Parameters of the motion search by relaxation passes (one persistent launch per frame step).
// Every stream of a launch shares the search range, lambda and thresholds.
// Waves a motion-search launch can have at most (workgroups <= 256 CUs, 16 waves each).
constexpr int kMaxSearchWaves = 256 * 16;

struct SearchParams {
    SearchStream s0;             // the stream of a single-stream launch (streams == nullptr)
    const SearchStream* streams; // device array of nstreams streams, ascending blk0, or nullptr
    int nstreams;
    int nb;                      // blocks of all streams
    int sr;
    uint32_t* field;             // 2 * nb MV words: the field after even / odd passes; bytes
                                 // (dx, dy, sx, sy)
    unsigned* arrive;            // 2 banks of barrier_words(maxp) pass-barrier counters
    int bank;                    // bank of this launch (zero at launch); the launch clears the other
    int maxp;                    // pass bound: 2*(bh-1) + bw + 2 passes of the largest stream always
                                 // reach the fixed point (it also lays out the barrier counters)
    int plimit;                  // passes this launch may run (<= maxp); not settling within them fails it,
                                 // unless `truncate`
    int truncate;                // learned form only: after plimit passes keep the field reached (a valid
                                 // field: the mode decision follows it) instead of failing
    int nslot, nres;             // LDS window slots per wave: lane k < nres keeps its block's window in
                                 // slot k; when nslot = nres + 1 the other lanes share slot nres
    const double* bits;          // 256 entries: 1.0 for 0, else 2*log2(a)+2 (motion/scalar.rs:495-525)
    double lam2;                 // lambda * 2.0
    int c_zero, c_pred;          // (lam2*0.5) as i64, (lam2*1.5) as i64; c_pred is also the half-pel penalty
    int lam_zero;                // lambda == 0.0 -> every MV cost is 0
    long long zthr, pthr;        // me_zero_mv_threshold, me_predictor_threshold
    // learned form (single-stream launches): per block the network's proposals (candidates 0|1,
    // 2|3 as (x, y) bytes, confident flag, logit bits), or null for the plain search; wave w of
    // the launch accumulates (blocks, confident, proposal taken) at lstat[3w .. 3w+2]
    const uint4* prop;
    unsigned long long* lstat;
};

}  // namespace riv
