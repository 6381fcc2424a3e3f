// Host-side entropy coding of the RIV bitstream: adaptive binary rANS core, residual
// coefficient coder and motion-vector coder. These stay on the CPU (inherently serial per
// frame); the GPU produces their inputs.
#pragma once
#include <cstdint>
#include <cstddef>
#include <functional>
#include <vector>
#include <array>
#include <stdexcept>

namespace rivh {

struct CorruptStream : std::runtime_error { using std::runtime_error::runtime_error; };

// Reciprocals for P(0) without a divide: x / d == (x * kP0Rcp[d]) >> 32 for x < 2^16, d <= 510
// (kP0Rcp[d] = floor(2^32 / d) + 1: the error x * (kP0Rcp[d] - 2^32/d) / 2^32 < 2^-16 stays
// below the distance 1/d from x/d to the next integer above it). kP0Rcp[0] = 0.
extern const uint32_t kP0Rcp[512];

// Adaptive binary context: two saturating 8-bit counters packed in 16 bits
// (reitero_video_common/src/rans.rs:25-85).
struct Ctx {
    uint16_t c = 0x0101;
    // P(bit == 0) * 256: zeros*256/(zeros+ones) truncated to u8 (rans.rs:30-44), 0 read as 1
    uint32_t p0() const {
        const uint32_t z = c >> 8, d = z + (c & 0xffu);
        const uint32_t p = (uint32_t)(((uint64_t)(z << 8) * kP0Rcp[d]) >> 32) & 0xffu;
        return p ? p : 1u;
    }
    // BinProb::observe (rans.rs:73-84): the counter of the observed symbol sits in the high byte;
    // on overflow both counters are halved (rounding up), the observed one re-based.
    // The observed counter steps by one; only a saturated one (255, rare) takes the general rule.
    void update(bool bit) {
        if (__builtin_expect(((c >> (bit ? 0 : 8)) & 0xffu) == 0xffu, 0)) update_saturated(bit);
        else c = (uint16_t)(c + (bit ? 1u : 0x100u));
    }
    void update_saturated(bool bit) {
        const uint32_t v = bit ? (uint32_t)(((c << 8) | (c >> 8)) & 0xffffu) : c;
        const uint32_t halved = (((v + 0x101u) & 0xffffu) >> 1) | (v == 0xff01u ? 0xff00u : 0x8100u);
        const uint32_t n = v >= 0xff00u ? halved : v + 0x100u;
        c = (uint16_t)(bit ? ((n << 8) | (n >> 8)) : n);
    }
};

// Symbol-batched rANS encoder with two interleaved 32-bit states and u16 renormalisation
// (rans.rs:88-298): symbols are coded in reverse in blocks of 16386, every block from fresh
// states, and the blocks' outputs follow each other in order. So the whole stream's symbols are
// buffered (as start << 8 | freq) and the blocks are coded at flush().
class RansWriter {
public:
    explicit RansWriter(std::vector<uint8_t>& out);
    ~RansWriter();
    RansWriter(const RansWriter&) = delete;
    RansWriter& operator=(const RansWriter&) = delete;
    void bit(bool b, Ctx& c) {
        const uint32_t p = c.p0();
        c.update(b);
        push(b ? ((p << 8) | (256u - p)) : p);
    }
    void bypass(bool b) { push(b ? 0x8080u : 0x80u); }
    // Direct writes: room for n more symbols at the returned pointer; commit(end) adopts them.
    // Once sixteen full blocks are buffered they are coded and dropped from the buffer (the
    // blocks are coded independently and emitted in order), which keeps the buffer small.
    uint16_t* reserve(size_t n) {
        if (cap_ - n_ < n || n_ >= 16 * (size_t)kBatch) make_room(n);
        return sym_ + n_;
    }
    void commit(const uint16_t* end) { n_ = (size_t)(end - sym_); }
    void flush();
    // flush() with the blocks of a stream of two to seven blocks coded by parallel tasks, one block
    // each, through pfor(n, fn) (fn(0) .. fn(n - 1)); the bytes are flush()'s, as every block is
    // coded from fresh states and they are emitted in order. Other streams, or no pfor: flush().
    void flush_parallel(const std::function<void(int, const std::function<void(int)>&)>& pfor);
    // Codes n symbols held by the caller (start << 8 | freq each) as a whole stream: the blocks
    // flush() would code, appended to the output; nothing may be buffered.
    void code_symbols(const uint16_t* sym, size_t n);
    // Fills the calling thread's buffer pool (when empty) with buffers for `symbols` symbols, their
    // pages touched, so that the thread's first frames do not pay for growing them.
    static void prewarm(size_t symbols);
private:
    void push(uint32_t s) {
        if (n_ == cap_) grow();
        sym_[n_++] = (uint16_t)s;
    }
    void grow();
    void make_room(size_t n);
    void drain_group(size_t b, int nblocks);   // 16 or 8 full blocks from symbol b, vector coder
    void drain(const uint16_t* const* sym, const int* n, int nblk);
    void emit_block(uint32_t r0, uint32_t r1, const uint16_t* wp, const uint16_t* end);
    static constexpr int kBatch = 16386;
    static constexpr int kLanes = 4;   // blocks coded together
    std::vector<uint8_t>& out_;
    // buffers come from a per-thread pool, so a worker coding frame after frame reuses warm memory
    std::vector<uint16_t> buf_;
    std::vector<uint16_t> tmp_[kLanes];   // emitted words of the blocks being coded, filled from the end
    std::vector<uint16_t> tmp8_[16];      // the same for the vector coder's blocks (up to sixteen)
    uint16_t* sym_ = nullptr;
    size_t n_ = 0, cap_ = 0;
    bool emitted_ = false;   // blocks were already coded by make_room
};

class RansReader {
public:
    RansReader(const uint8_t* data, size_t len);
    bool bit(Ctx& c);
    bool bypass();
private:
    uint32_t state();
    uint16_t word();
    void step();
    const uint8_t* d_; size_t len_, pos_ = 0;
    uint32_t a_ = 0, b_ = 0;
    int count_ = 0;
};

// Residual coefficient coder (reitero_residual/src/rans.rs:1-460). One instance per frame
// (contexts are fresh every frame). `dc_pred` is the intra DC-prediction mode.
struct ResidualCtx {
    Ctx sig_y[256], sig_c[64];
    Ctx mag_y[4][13], mag_c[3][13];
    Ctx eob_y[9], eob_c[7];
    int16_t prev_dc[3] = {0, 0, 0};
};
// Codes the coefficients of the MBs with coded[b] != 0, in raster order. coeffs: nb*384.
std::vector<uint8_t> encode_residuals(const int16_t* coeffs, const uint8_t* coded, int nb, bool dc_pred);
// The same coding from packed per-MB slots (six 64-bit non-zero masks, then the non-zero values
// in index order), `stride` bytes apart.
std::vector<uint8_t> encode_residuals_packed(const uint8_t* pack, size_t stride, const uint8_t* coded, int nb, bool dc_pred);
// The inter form from split records: MB b's masks at main + b * PACK_MAIN, its values after them
// when there are at most PACK_MAIN_VALUES, else at over + b * PACK_STRIDE + 48 (coded MBs only).
std::vector<uint8_t> encode_residuals_split(const uint8_t* main, const uint8_t* over, const uint8_t* coded, int nb);
// Runs fn(0) .. fn(n - 1), possibly on several threads at once, and returns when all have run.
using ParallelFor = std::function<void(int n, const std::function<void(int)>& fn)>;
// encode_residuals_split, modelled and coded by parallel tasks, byte for byte the same stream.
// The luma and chroma contexts are disjoint, and within a plane so are the significance and
// magnitude contexts of the position ranges [0,16), [16,64), [64,256) (luma) and [0,8), [8,64)
// (chroma), whose magnitude bands do not straddle them; so each range is modelled over every
// coded MB by its own task (the EOB contexts go with the first range). The tasks' symbols are
// interleaved back into the coder's order (per MB: Y, U, V, each plane's ranges in position
// order), and the stream's 16386-symbol rANS blocks, each coded from fresh states, are coded in
// parallel groups.
std::vector<uint8_t> encode_residuals_split_par(const uint8_t* main, const uint8_t* over, const uint8_t* coded, int nb,
                                                const ParallelFor& pfor);
// rANS32 coding of a stream's whole symbol sequence (n symbols, each start << 8 | freq, as the
// residual coder buffers them): its 16386-symbol blocks, each coded from fresh states, in groups
// of group_blocks (16 or 8), by parallel tasks when pfor is set; the bytes do not depend on the
// grouping. An empty sequence gives an empty stream (no coder is instantiated when every block
// is skipped, residual.rs:670-674).
std::vector<uint8_t> code_symbol_stream(const uint16_t* sym, size_t n, const ParallelFor& pfor, int group_blocks = 16);
// One MB of the split form as 384 coefficients.
void unpack_split(const uint8_t* main, const uint8_t* over, int b, int16_t* q);
void unpack_coeffs(const uint8_t* slot, int16_t* q384);
// Inverse; fills coeffs (nb*384, caller-zeroed) of the MBs with coded[b] != 0.
void decode_residuals(const uint8_t* data, size_t len, const uint8_t* coded, int nb, bool dc_pred, int16_t* coeffs);

// Motion-vector coder state (reitero_residual/src/mv_rans.rs, mv_rans/encoder.rs, decoder.rs):
// contexts and the NEW-delta candidate history persist across P-frames, reset at I-frames.
struct MvBlock {
    uint8_t mode, new_base;
    int8_t dx, dy;           // NEW delta after the frame bias
    int8_t sx, sy;           // sub-pel of a NEW MV: 0, +1 (-1 is rejected by the coder)
    bool skip;
};
class MvCoder {
public:
    void reset();
    std::vector<uint8_t> encode(const std::vector<MvBlock>& blocks, int bw, int bh);
    // The frame's symbols into w (not flushed) and the history update: the part of encode that
    // depends on the previous frames; w.flush() then yields the bytes encode returns. It is
    // events() then contexts(), which touch disjoint state: frame k's contexts may run beside
    // frame k + 1's events (each in frame order).
    void model(const std::vector<MvBlock>& blocks, int bw, int bh, RansWriter& w);
    // The frame's decisions in coding order as (context, bit) events into evs (grown as needed;
    // returns how many), the candidate list of the vectors so far deciding the hit events, then
    // the history update for the next frame.
    size_t events(const std::vector<MvBlock>& blocks, int bw, std::vector<uint16_t>& evs);
    // The adaptive contexts over a frame's n events, in order, each giving its symbol into w.
    void contexts(const uint16_t* ev, size_t n, RansWriter& w);
    void reset_history();    // reset() is both of these
    void reset_contexts();
    std::vector<MvBlock> decode(const uint8_t* data, size_t len, int bw, int bh);
private:
    void refresh_history(const std::vector<MvBlock>& blocks, bool from_decoder);
    int lookup(int8_t dx, int8_t dy, int8_t sx, int8_t sy) const;
    // the adaptive contexts, in one array so that a coded decision can name its context by index
    enum : int {
        kSkip = 0,
        kMode = 1,      // [6][4]: is_new, nonzero, is_near, is_temporal, is_top_right, is_top_left x mode context
        kBase = 25,     // [4][4]: nearest, near, top_right, top_left x mode context
        kCls = 41,      // [2][4]: per axis, by depth of the class tree
        kSign = 49,     // [2]
        kMag = 51,      // [2][10]
        kFrac = 71,     // [2]
        kHit = 73,      // [3]
        kIdx = 76,      // [4]
        kNumCtx = 80
    };
    Ctx ctx_[kNumCtx];
    // candidate history: key = (dx, dy, sub-pel order x, sub-pel order y)
    std::vector<std::array<int, 4>> cand_;
    uint32_t cand_key_[16];   // the candidates' packed keys (below), the unused ones ~0u (matches no key)
    // counts of the packed keys (dx+128, dy+128, sub-pel order x, y as bytes, high to low, so that
    // unsigned order is tuple order): a list of the live ones and an open-addressing index into it
    std::vector<std::pair<uint32_t, uint32_t>> live_;   // (key, count)
    std::vector<uint32_t> index_;                       // slot: live_ position + 1, 0 = empty
    uint32_t total_ = 0;
    uint32_t* find_count(uint32_t key, bool insert);
};

int mv_class(int magnitude);   // mv_class_from_magnitude (mv_rans.rs:81-87)

}  // namespace rivh
