/*
 * riv_gpu.h -- C ABI of the MI355X (gfx950) backend for the RIV (.riv) video codec.
 *
 * The encoder/decoder objects mirror the reference Rust API one-for-one; each
 * declaration names the reference item it replaces (paths relative to the
 * reference tree, under reitero_video/). Plain pointers and sizes only; no C++
 * exceptions cross this boundary; every call returns a riv_status.
 *
 * Threading: one handle per thread. Each handle owns one HIP stream on its device
 * and keeps its reference frame, motion vectors and scratch in that device's HBM.
 */
#ifndef RIV_GPU_H
#define RIV_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 1..5 map EncodeError (reitero_encode/src/error.rs:4-19),
 * 11..14 map DecodeError (reitero_decode/src/error.rs:5-20). */
typedef enum {
    RIV_OK = 0,
    RIV_E_IO = 1,                 /* EncodeError::Io / DecodeError::Io */
    RIV_E_INVALID_CONFIG = 2,     /* EncodeError::InvalidConfig */
    RIV_E_ENCODING_FAILED = 3,    /* EncodeError::EncodingFailed */
    RIV_E_UNSUPPORTED_CODEC = 4,  /* EncodeError::UnsupportedCodec */
    RIV_E_FRAME = 5,              /* EncodeError::FrameError */
    RIV_E_INVALID_HEADER = 11,    /* DecodeError::InvalidHeader */
    RIV_E_DECODING_FAILED = 12,   /* DecodeError::DecodingFailed (also corrupt entropy data, where the reference panics) */
    RIV_E_INVALID_FRAME = 13,     /* DecodeError::InvalidFrame */
    RIV_E_END_OF_STREAM = 14,     /* DecodeError::EndOfStream */
    RIV_E_GPU = 20,               /* HIP runtime / kernel failure (no reference counterpart) */
    RIV_E_UNSUPPORTED = 21,       /* valid reference config outside this backend's limits (search_range > 63) */
    RIV_E_ARG = 22                /* NULL handle / bad argument */
} riv_status;

/* KeyframeInterval (reitero_encode/src/config.rs:5-19) */
typedef enum { RIV_KEYFRAME_FIXED = 0, RIV_KEYFRAME_ALL_INTRA = 1, RIV_KEYFRAME_AUTOMATIC = 2 } riv_keyframe_mode;

/* EncoderConfig (reitero_encode/src/config.rs:24-53) + device selection */
typedef struct {
    uint32_t display_width, display_height;
    uint32_t storage_width, storage_height;   /* round_up(display, 16) (config.rs:57-58) */
    uint32_t fps;
    int32_t keyframe_mode;                    /* riv_keyframe_mode */
    uint64_t keyframe_interval;               /* N for RIV_KEYFRAME_FIXED */
    uint8_t intra_quality, inter_quality;
    uint8_t search_range;
    uint8_t mv_deflate_level;                 /* vestigial in the reference, validated only */
    uint8_t skip_threshold;
    int64_t me_zero_mv_threshold, me_predictor_threshold;
    double rdo_lambda_mult;
    float inter_dead_zone;
    int32_t device_id;                        /* HIP device ordinal */
    /* 0: encode_frame returns after the frame is written (reference semantics). N > 0: up to N
     * frames in flight; the GPU analyses frame k+1 while host threads entropy-code earlier frames,
     * frames are written in order, stats arrive through riv_encoder_pop_stats. Same bitstream. */
    int32_t pipeline_depth;
} riv_encoder_config;

/* EncoderConfig::new(w, h, fps) defaults (config.rs:56-77) */
void riv_encoder_config_default(riv_encoder_config* cfg, uint32_t display_width, uint32_t display_height, uint32_t fps);
/* EncoderConfig::validate (config.rs:129-164) */
int riv_encoder_config_validate(const riv_encoder_config* cfg);

/* VideoWriter trait (reitero_encode/src/writer.rs:5-18). Callbacks return 0 on success. */
typedef struct {
    int (*write)(void* user, const uint8_t* data, size_t len);
    uint64_t (*position)(void* user);
    int (*seek)(void* user, uint64_t pos);
    int (*flush)(void* user);
} riv_writer_vtbl;

/* VideoReader trait (reitero_decode/src/reader.rs:4-13): read returns bytes read, 0 at end. */
typedef struct {
    size_t (*read)(void* user, uint8_t* buf, size_t cap);
    uint64_t (*position)(void* user);
    int (*seek)(void* user, uint64_t pos);
} riv_reader_vtbl;

/* VideoHeader (reitero_video_common/src/format.rs:34-88) */
typedef struct {
    uint32_t display_width, display_height, storage_width, storage_height, fps;
    uint64_t frame_count;
} riv_video_header;

/* EncodeFrameStats (reitero_encode/src/encoder.rs:58-124), the fields this backend fills. */
typedef struct {
    uint64_t frame_index, timestamp_ms;
    int32_t frame_type;                       /* 1 = intra, 2 = inter (FrameType, format.rs:12-17) */
    uint32_t total_bytes, mv_bytes, residual_bytes;
    uint32_t storage_width, storage_height;
    uint32_t blocks_total, blocks_skipped;
    uint32_t mv_mode_counts[7];               /* Zero, Nearest, Near, TopRight, TopLeft, New, Temporal */
    uint32_t mv_new_blocks;
    int32_t mv_bias_dx, mv_bias_dy;
    double rdo_lambda;
    uint32_t rdo_blocks_evaluated, rdo_blocks_selected_skip, rdo_forced_fractional, rdo_forced_threshold, rdo_forced_disabled;
    double gpu_ms;                            /* device time of this frame's kernels when stage profiling (level 2) is
                                                 on, else 0 (no reference counterpart) */
    double host_entropy_ms;                   /* host residual rANS time of this frame (no reference counterpart) */
    double host_tail_ms;                      /* host ordered tail: MV rANS, packing, write (no reference counterpart) */
    uint32_t me_passes;                       /* relaxation passes of the motion search, 0 for intra (no reference counterpart) */
    /* Host timeline of the frame, ms since the encoder was created (no reference counterpart):
     * submitted (its encode call), its GPU work seen finished, its residuals coded, written. */
    double t_submit_ms, t_gpu_done_ms, t_coded_ms, t_written_ms;
    /* 1 when the frame's residual symbols were modelled on the GPU and only rANS-coded on the host,
     * 0 when the host modelled them too (no reference counterpart) */
    uint32_t residual_symbols_on_gpu;
} riv_frame_stats;

typedef struct riv_encoder riv_encoder;
typedef struct riv_decoder riv_decoder;

/* Encoder::new (reitero_encode/src/encoder.rs:162-188): validates, writes the 36-byte header. */
int riv_encoder_create(const riv_encoder_config* cfg, const riv_writer_vtbl* writer, void* user, riv_encoder** out);
/* Encoder::new with VecWriter (writer.rs:21-77): output kept in memory, see riv_encoder_output. */
int riv_encoder_create_mem(const riv_encoder_config* cfg, riv_encoder** out);
/* Encoder::encode_frame / encode_frame_with_stats (encoder.rs:229-256). rgb: host RGB24,
 * width*height*3 bytes, pageable or pinned; the frame has been copied to the GPU when the call
 * returns, so the caller may reuse its memory at once. stats may be NULL. */
int riv_encode_frame(riv_encoder* enc, const uint8_t* rgb, size_t len, uint32_t width, uint32_t height,
                     uint64_t timestamp_ms, riv_frame_stats* stats);
/* Same, with the RGB24 frame already resident in device memory of the encoder's device. */
int riv_encode_frame_device(riv_encoder* enc, const void* d_rgb, size_t len, uint32_t width, uint32_t height,
                            uint64_t timestamp_ms, riv_frame_stats* stats);
/* Batched form (no reference counterpart; it stands for running one reference Encoder per stream):
 * one frame for each of n encoders, whose streams may differ in size. The GPU work of all n frames
 * is queued on encs[0]'s stream, and the motion searches of the inter frames run as one persistent
 * launch over all their blocks. Every stream's output is byte-identical to encoding it alone.
 * The encoders must share device, search range, inter quality / rdo_lambda_mult and the ME
 * thresholds, and appear once each; rgb[i] is a host (on_device = 0) or device (on_device = 1)
 * RGB24 frame of len[i] bytes at encs[i]'s display size; host frames have been copied when the call
 * returns. timestamp_ms and stats (n entries) may be NULL. With pipeline_depth > 0 the call returns
 * once the work is queued. */
int riv_encode_batch_step(riv_encoder* const* encs, size_t n, const void* const* rgb, const size_t* len, const uint64_t* timestamp_ms,
                          int on_device, riv_frame_stats* stats);
/* Blocks until every submitted frame has been written (pipelined mode; otherwise frames are written
 * by their encode call) and the encoder's own HIP streams are idle; other streams are not waited for. */
int riv_encoder_flush(riv_encoder* enc);
/* Completed frames' stats in encode order (pipelined mode); *n receives the number copied. */
int riv_encoder_pop_stats(riv_encoder* enc, riv_frame_stats* out, size_t cap, size_t* n);
/* Encoder::finish (encoder.rs:950-961): patches frame_count into the header, flushes. */
int riv_encoder_finish(riv_encoder* enc);
/* Encoder::frame_count / header (encoder.rs:969-981) */
uint64_t riv_encoder_frame_count(const riv_encoder* enc);
int riv_encoder_header(const riv_encoder* enc, riv_video_header* out);
/* Encoder::config (encoder.rs:974-976): the configuration the encoder was created with. */
int riv_encoder_get_config(const riv_encoder* enc, riv_encoder_config* out);
/* VecWriter::into_inner (writer.rs:33): bytes written so far (mem writer only). */
int riv_encoder_output(const riv_encoder* enc, const uint8_t** data, size_t* len);
void riv_encoder_destroy(riv_encoder* enc);

/* Per-block record of the last encoded P-frame: search MV (dx,dy,sx,sy), reconstructed MV
 * (dx,dy,sx,sy), mode, new_base, raw NEW delta (dx,dy) -- the values of
 * motion/scalar.rs:546-636 and encoder.rs:447-733. rec12: nblocks*12 int8, sad: nblocks. */
int riv_encoder_debug_blocks(riv_encoder* enc, int8_t* rec12, int64_t* sad, uint8_t* rdo_skip, uint8_t* opt_skip);
/* Last frame's planes: which = 0 current (after RGB->YUV), 1 motion-compensated prediction,
 * 2 reconstruction after deblock (the next reference). */
int riv_encoder_debug_yuv(riv_encoder* enc, int which, uint8_t* y, uint8_t* u, uint8_t* v);
/* Last frame's quantised coefficients, nblocks*384 int16 (Y 256 | U 64 | V 64, row-major), zero for skipped blocks. */
int riv_encoder_debug_coeffs(riv_encoder* enc, int16_t* coeffs);
/* Last P-frame's adaptive quant steps (residual.rs:116-167), nblocks floats. */
int riv_encoder_debug_aq(riv_encoder* enc, float* steps);
/* Test seam of the motion search's failure path (no reference counterpart): caps the relaxation
 * passes a search may run (0 restores the proven bound 2*(bh-1) + bw + 2). A search that has not
 * settled within the cap fails its frame: the frame returns (synchronous mode) or the next call
 * and riv_encoder_flush return (pipelined mode) RIV_E_GPU, and no byte of that frame is written. */
int riv_encoder_debug_pass_limit(riv_encoder* enc, int32_t max_passes);
/* Test seam (no reference counterpart): how P-frame residuals are coded. 0 (the default): their
 * symbols are modelled on the GPU and the host codes them with rANS32 (batch steps, and frames
 * with more symbols than the GPU buffers hold, take mode 3); 1: modelled and coded on the host,
 * serially; 2: on the host, split across its workers; 3: on the host, split when at least two
 * workers are idle. Every mode writes the same bytes. */
int riv_encoder_debug_split_coding(riv_encoder* enc, int32_t mode);
/* Test seam (no reference counterpart): the symbols a P-frame's GPU residual modelling may hold,
 * at most the encoder's own limit (256 per MB; 0 restores it). A frame with more is coded on the
 * host, byte for byte the same. */
int riv_encoder_debug_model_cap(riv_encoder* enc, uint32_t cap);

/* Decoder::new (reitero_decode/src/decoder.rs:112-171): reads and validates the header. */
int riv_decoder_create(const riv_reader_vtbl* reader, void* user, int32_t device_id, riv_decoder** out);
int riv_decoder_create_mem(const uint8_t* data, size_t len, int32_t device_id, riv_decoder** out);

/* DecodedFrame (decoder.rs:36-49) metadata; the RGB24 pixels go to the caller's buffer. */
typedef struct {
    uint32_t width, height;
    uint64_t timestamp_ms, frame_index;
    int32_t frame_type;
} riv_decoded_info;

/* Decoder::decode_frame (decoder.rs:174-192): RGB24 cropped to display size. */
int riv_decode_frame(riv_decoder* dec, uint8_t* rgb_out, size_t cap, riv_decoded_info* info);
/* Decoder::decode_frame_null (decoder.rs:463-470) */
int riv_decoder_decode_frame_null(riv_decoder* dec);
/* Decoder::has_more_frames / header / set_skip_residuals / skip_residuals / current_frame
 * (decoder.rs:470-499) */
int riv_decoder_has_more_frames(const riv_decoder* dec);
int riv_decoder_header(const riv_decoder* dec, riv_video_header* out);
int riv_decoder_set_skip_residuals(riv_decoder* dec, int skip);
int riv_decoder_skip_residuals(const riv_decoder* dec);
uint64_t riv_decoder_current_frame(const riv_decoder* dec);
/* DecodePhaseTimings (decoder.rs:73-87), nanoseconds summed since the last drain. Host phases, on
 * whichever thread ran them: read_bits (reading frame records), parse_frame (record headers and
 * vector reconstruction), mv_decode (MV rANS), residual_rans (residual rANS). Device phases, from
 * HIP events on the decoder's stream, only while riv_decoder_set_profiling is on (else 0):
 * build_pred (motion compensation), residual_apply (dequantisation, inverse transforms and the
 * add: one fused kernel, so residual_dct_y and residual_dct_uv stay 0; with the deblocking
 * filter), yuv_to_rgb (conversion and the copy out). residual_deinterleave is 0: the residual
 * decoder writes coefficients in place. residual = the residual counters' sum, as the reference
 * defines it. */
typedef struct {
    uint64_t read_bits_ns, parse_frame_ns, mv_decode_ns, build_pred_ns, residual_ns, residual_rans_ns, residual_deinterleave_ns,
        residual_dct_y_ns, residual_dct_uv_ns, residual_apply_ns, yuv_to_rgb_ns;
} riv_decode_timings;
/* Decoder::drain_timings (decoder.rs:485-489): the timings, then zero. */
int riv_decoder_drain_timings(riv_decoder* dec, riv_decode_timings* out);
/* Device-phase timing of riv_decoder_drain_timings (no reference counterpart): 1 on, 0 off (the
 * default; each timed phase costs the stream two HIP events). */
int riv_decoder_set_profiling(riv_decoder* dec, int on);
/* Last decoded frame's storage-size YUV planes (before RGB conversion). */
int riv_decoder_debug_yuv(riv_decoder* dec, uint8_t* y, uint8_t* u, uint8_t* v);
void riv_decoder_destroy(riv_decoder* dec);

/* ---- kernel-level seams (reference internal plug points), host buffers in/out ---- */
/* build_predicted (reitero_video_common/src/fast_motion.rs:103-108); mv4 = nblocks*(dx,dy,sx,sy) */
int riv_gpu_build_predicted(int32_t device_id, const uint8_t* y, const uint8_t* u, const uint8_t* v, uint32_t w, uint32_t h,
                            const int8_t* mv4, uint8_t* out_y, uint8_t* out_u, uint8_t* out_v);
/* deblock_yuv420 (reitero_video_common/src/deblock.rs:189-202); mask may be NULL (all edges) */
int riv_gpu_deblock_yuv420(int32_t device_id, uint8_t* y, uint8_t* u, uint8_t* v, uint32_t w, uint32_t h, int32_t level,
                           int32_t intra, const uint8_t* mask);
/* encode_plane_8x8_aq + decode_plane_8x8_aq (reitero_dct/src/dct8.rs:80-135, fast_dct8.rs:69-110) on one plane */
int riv_gpu_dct8_aq_plane(int32_t device_id, const int16_t* plane, uint32_t w, uint32_t h, const float* qs, float dead_zone,
                          int16_t* coeffs, int16_t* recon);
/* encode_plane_16x16_aq + decode_plane_16x16_aq (reitero_dct/src/dct16.rs:274-412) on one plane */
int riv_gpu_dct16_aq_plane(int32_t device_id, const int16_t* plane, uint32_t w, uint32_t h, const float* qs, float dead_zone,
                           int16_t* coeffs, int16_t* recon);
/* compute_aq_quant_steps (reitero_residual/src/residual.rs:116-167) on a luma plane */
int riv_gpu_aq_steps(int32_t device_id, const uint8_t* y, uint32_t w, uint32_t h, float base_step, float* out);
/* No reference counterpart: compute_aq_quant_steps (reitero_residual/src/residual.rs:116-167) on
 * a luma plane already in device memory, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Same steps as riv_gpu_aq_steps.
 * It allocates nothing: its scratch (the per-MB logs and their mean) is a static device array of
 * RIV_AQ_NOALLOC_MAX_BLOCKS + 64 floats that each device holds until the process ends, and larger
 * planes are rejected with RIV_E_ARG. y: w x h luma plane, steps: (w/16)*(h/16) floats out, both
 * device memory of device_id, 4-byte aligned; w, h multiples of 16. stream: a hipStream_t of that
 * device (NULL: the null stream); the call returns once the work is queued. Every call on a
 * device shares that device's scratch, so calls that overlap on one device (other streams,
 * other threads) race; calls queued in order on one stream do not. */
#define RIV_AQ_NOALLOC_MAX_BLOCKS 32768
int riv_gpu_aq_steps_noalloc(int32_t device_id, const uint8_t* y, uint32_t w, uint32_t h, float base_step, float* steps, void* stream);
/* The adaptive quantiser's mean of n per-MB logs (the sequential f32 sum of
 * reitero_residual/src/residual.rs:150-155, divided by n) and the steps base_step derives from it;
 * every log must be positive, as the logs of floored variances are. steps may be NULL. */
int riv_gpu_aq_mean(int32_t device_id, const float* logs, size_t n, float base_step, float* mean, float* steps);
/* The motion search's SAD tables (sad_block_int_limit, satd_scalar.rs:4-99, and
 * sad_block_halfpel_limit_luma, satd_scalar.rs:171-223, at every candidate the search can visit):
 * per 16x16 block, int_sad[(dy+sr)*(2sr+1) + dx+sr] for dx,dy in [-sr,sr] and
 * hp_sad[(hy+2sr+1)*(4sr+3) + hx+2sr+1] for hx,hy in [-2sr-1,2sr+1] (half-pel units), blocks in
 * raster order. `cur` and `ref` are w x h luma planes (multiples of 16); sr <= 32. */
int riv_gpu_me_tables(int32_t device_id, const uint8_t* cur, const uint8_t* ref, uint32_t w, uint32_t h, int32_t sr,
                      uint16_t* int_sad, uint16_t* hp_sad);
/* glibc-identical logf / exp2f on the device (used by AQ), n values */
int riv_gpu_logf_exp2f(int32_t device_id, const float* in, size_t n, int32_t which, float* out);

/* Per-stage device timing of the encoder (HIP events on the encoder's stream around each stage;
 * no reference counterpart). Stages: 0 RGB->YUV, 1 motion search (relaxation passes, mode
 * decision and the prediction variances), 2 motion compensation (on the encoder's P-frame path
 * formed inside stage 4, so not recorded), 3 adaptive quantiser, 4 residual prediction/transform/
 * quantise/reconstruct, 5 deblock, 6 intra transform, 7 the whole motion-search launch of batch
 * steps this encoder leads (all streams; the stage-1 figures of encoders in a batch are not the
 * search alone), 8 the GPU context modelling of a P-frame's residual symbols (their copy to host
 * memory is the host stage's). `on`: 0 off, 1 the motion-search
 * stage only (two events per frame), 2 every stage and each frame's gpu_ms. Every event costs the
 * stream a few microseconds. */
#define RIV_NUM_STAGES 9
int riv_encoder_set_profiling(riv_encoder* enc, int on);
/* Profile only the frames whose index is a multiple of `every` (1, the default: every frame), a
 * sample of the stage times that costs the stream fewer events (no reference counterpart). */
int riv_encoder_set_profiling_sample(riv_encoder* enc, uint32_t every);
/* Accumulated milliseconds and launch counts per stage since profiling was switched on. */
int riv_encoder_stage_times(riv_encoder* enc, double* ms, uint32_t* counts);

/* Batch-step launch-shape autotuner (no reference counterpart; the reference has no batched
 * form). Per (blocks, streams, search range) of a batch, the first batch steps led by `enc` (the
 * first encoder of the batch) each launch one candidate motion-search shape (waves per
 * workgroup, workgroups, window slots per wave), timed with events read without waiting; after
 * three samples of every candidate the fastest is kept. The shape changes no output. On by
 * default; off keeps the occupancy-derived shape. */
int riv_encoder_set_batch_autotune(riv_encoder* enc, int32_t on);
/* The latest batch step's shape: out[0] waves per workgroup, [1] workgroups, [2] window slots per
 * wave, [3] candidates, [4] 1 once tuning has settled, [5] the chosen candidate (or -1). */
int riv_encoder_batch_tuning(riv_encoder* enc, int32_t* out);

/* ---- learned form: a BF16 MFMA MV-proposal network seeding the motion search ----
 * The reference has no learned form; the search it seeds is process_block_hex
 * (reitero_encode/src/motion/scalar.rs:546-636). Per 16x16 block the network reads the current
 * block with an 8-pixel border, the co-located reference patch and the previous P-frame's vectors
 * around the block, and proposes four integer vectors and a confidence. In the encoder, a block
 * whose confidence clears the threshold scores the four proposals (SAD + lambda * bits, as the
 * search scores candidates); when the best is no worse than the best predictor seed it starts a
 * small-diamond refine (at most 8 steps) instead of the hexagon search, then the usual half-pel
 * refine. Every other block runs the plain search. The bitstream format is unchanged: any
 * decoder of the plain form decodes it. */
typedef struct riv_mvnet riv_mvnet;
/* A frame pair the network reads: luma planes in device memory (width x height, multiples of
 * 16) and the previous P-frame's vectors (x, y as int8 at prev_mv + block * prev_stride) or NULL. */
typedef struct {
    const uint8_t* cur_y;
    const uint8_t* ref_y;
    const int8_t* prev_mv;
    int32_t prev_stride;
    uint32_t width, height;
} riv_mv_source;
/* Number of f32 parameters (weights [K][N] row-major and biases, in layer order). */
size_t riv_mvnet_param_count(void);
/* A network with parameters initialised from `seed`, on device_id. */
int riv_mvnet_create(int32_t device_id, uint64_t seed, riv_mvnet** out);
void riv_mvnet_destroy(riv_mvnet* net);
int riv_mvnet_get_params(riv_mvnet* net, float* out, size_t n);
/* Replaces the parameters (and resets the optimiser state). */
int riv_mvnet_set_params(riv_mvnet* net, const float* in, size_t n);
/* The frame pairs that samples index; the planes must stay valid while they are used. */
int riv_mvnet_set_sources(riv_mvnet* net, const riv_mv_source* src, size_t n);
/* One Adam step on n samples: samples = n (source, block) int32 pairs, targets = n (x, y) int8
 * vectors (the plain search's integer vectors). loss (may be NULL) receives the batch's mean
 * loss; with NULL the call does not wait for the GPU. */
int riv_mvnet_train_step(riv_mvnet* net, const int32_t* samples, const int8_t* targets, size_t n, float lr, float* loss);
/* The loss's gradient with respect to every parameter (riv_mvnet_param_count floats), no update. */
int riv_mvnet_gradients(riv_mvnet* net, const int32_t* samples, const int8_t* targets, size_t n, float* grad, float* loss);
/* Raw outputs, n x 16 floats (9 used), by the layer-wise kernels (fused 0) or the fused
 * inference kernel (fused 1). */
int riv_mvnet_forward(riv_mvnet* net, const int32_t* samples, size_t n, int32_t fused, float* out);
/* Proposals as the encoder receives them, n x 4 words: candidates 0 and 1, candidates 2 and 3
 * (x, y bytes, whole pixels clamped to +-sr), the confident flag, the confidence logit's bits. */
int riv_mvnet_propose(riv_mvnet* net, const int32_t* samples, size_t n, float confidence, int32_t sr, uint32_t* prop);
/* Switches the encoder to the learned form with `net` (NULL: back to the plain form) from its
 * next P-frame. confidence in (0, 1] is the probability threshold (1: never confident, so every
 * block takes the plain search). The network must outlive its use and keep its parameters while
 * frames are in flight. In a batch step each stream follows its own encoder's setting. */
int riv_encoder_set_mvnet(riv_encoder* enc, riv_mvnet* net, float confidence);
/* Learned form only: the motion search runs `passes` relaxation passes and keeps the field it
 * reached instead of relaxing to the reference's raster-order fixed point (0, the default: relax to
 * it). The field's vectors are searched as the learned form searches them, and the mode decision
 * and the bitstream follow the field, so any decoder of the plain form decodes the stream; a plain
 * encoder always relaxes to the fixed point. In a batch step a fixed pass count applies only when
 * every inter stream of the step is learned with the same count. */
int riv_encoder_set_learned_passes(riv_encoder* enc, int32_t passes);
/* Learned-form counts since the encoder was created: counts[0] P-frame blocks searched, [1] those
 * with a confident proposal, [2] those that took the proposal (the rest fell back). For the first
 * encoder of batch steps with learned streams, the counts also cover every stream of those steps. */
int riv_encoder_learned_stats(riv_encoder* enc, uint64_t* counts);

/* Last error message of the calling thread (static storage). */
const char* riv_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* RIV_GPU_H */
