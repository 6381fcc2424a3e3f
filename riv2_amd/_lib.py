"""ctypes binding of the C ABI declared in include/riv_gpu.h (libriv_gpu.so, built in-tree).

Loading fails loudly when the library has not been built: there is no CPU fallback.
"""
import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_double, c_float, c_int, c_int8, c_int16, c_int32, c_int64, c_size_t, \
    c_uint8, c_uint16, c_uint32, c_uint64, c_void_p

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libriv_gpu.so")
SYNTH_PATH = os.path.join(_HERE, "libriv_synth.so")

RIV_OK = 0
STATUS_NAMES = {
    1: "Io", 2: "InvalidConfig", 3: "EncodingFailed", 4: "UnsupportedCodec", 5: "FrameError",
    11: "InvalidHeader", 12: "DecodingFailed", 13: "InvalidFrame", 14: "EndOfStream",
    20: "Gpu", 21: "Unsupported", 22: "Argument",
}


class EncoderConfigC(Structure):
    _fields_ = [
        ("display_width", c_uint32), ("display_height", c_uint32),
        ("storage_width", c_uint32), ("storage_height", c_uint32),
        ("fps", c_uint32),
        ("keyframe_mode", c_int32),
        ("keyframe_interval", c_uint64),
        ("intra_quality", c_uint8), ("inter_quality", c_uint8),
        ("search_range", c_uint8), ("mv_deflate_level", c_uint8), ("skip_threshold", c_uint8),
        ("me_zero_mv_threshold", c_int64), ("me_predictor_threshold", c_int64),
        ("rdo_lambda_mult", c_double),
        ("inter_dead_zone", c_float),
        ("device_id", c_int32),
        ("pipeline_depth", c_int32),
    ]


class VideoHeaderC(Structure):
    _fields_ = [("display_width", c_uint32), ("display_height", c_uint32), ("storage_width", c_uint32),
                ("storage_height", c_uint32), ("fps", c_uint32), ("frame_count", c_uint64)]


class FrameStatsC(Structure):
    _fields_ = [
        ("frame_index", c_uint64), ("timestamp_ms", c_uint64),
        ("frame_type", c_int32),
        ("total_bytes", c_uint32), ("mv_bytes", c_uint32), ("residual_bytes", c_uint32),
        ("storage_width", c_uint32), ("storage_height", c_uint32),
        ("blocks_total", c_uint32), ("blocks_skipped", c_uint32),
        ("mv_mode_counts", c_uint32 * 7),
        ("mv_new_blocks", c_uint32),
        ("mv_bias_dx", c_int32), ("mv_bias_dy", c_int32),
        ("rdo_lambda", c_double),
        ("rdo_blocks_evaluated", c_uint32), ("rdo_blocks_selected_skip", c_uint32), ("rdo_forced_fractional", c_uint32),
        ("rdo_forced_threshold", c_uint32), ("rdo_forced_disabled", c_uint32),
        ("gpu_ms", c_double), ("host_entropy_ms", c_double), ("host_tail_ms", c_double),
        ("me_passes", c_uint32),
        ("t_submit_ms", c_double), ("t_gpu_done_ms", c_double), ("t_coded_ms", c_double), ("t_written_ms", c_double),
        ("residual_symbols_on_gpu", c_uint32),
    ]


class DecodedInfoC(Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("timestamp_ms", c_uint64), ("frame_index", c_uint64),
                ("frame_type", c_int32)]


WRITE_FN = ctypes.CFUNCTYPE(c_int, c_void_p, POINTER(c_uint8), c_size_t)
POS_FN = ctypes.CFUNCTYPE(c_uint64, c_void_p)
SEEK_FN = ctypes.CFUNCTYPE(c_int, c_void_p, c_uint64)
FLUSH_FN = ctypes.CFUNCTYPE(c_int, c_void_p)
READ_FN = ctypes.CFUNCTYPE(c_size_t, c_void_p, POINTER(c_uint8), c_size_t)


class MvSourceC(Structure):
    _fields_ = [("cur_y", c_void_p), ("ref_y", c_void_p), ("prev_mv", c_void_p), ("prev_stride", c_int32),
                ("width", c_uint32), ("height", c_uint32)]


class DecodeTimingsC(Structure):
    _fields_ = [(n, c_uint64) for n in ("read_bits_ns", "parse_frame_ns", "mv_decode_ns", "build_pred_ns", "residual_ns", "residual_rans_ns",
                                        "residual_deinterleave_ns", "residual_dct_y_ns", "residual_dct_uv_ns", "residual_apply_ns",
                                        "yuv_to_rgb_ns")]


class WriterVtblC(Structure):
    _fields_ = [("write", WRITE_FN), ("position", POS_FN), ("seek", SEEK_FN), ("flush", FLUSH_FN)]


class ReaderVtblC(Structure):
    _fields_ = [("read", READ_FN), ("position", POS_FN), ("seek", SEEK_FN)]


# (name, restype, argtypes) for every function of include/riv_gpu.h and include/riv_host.h
PROTOTYPES = [
    ("riv_encoder_config_default", None, [POINTER(EncoderConfigC), c_uint32, c_uint32, c_uint32]),
    ("riv_encoder_config_validate", c_int, [POINTER(EncoderConfigC)]),
    ("riv_encoder_create", c_int, [POINTER(EncoderConfigC), POINTER(WriterVtblC), c_void_p, POINTER(c_void_p)]),
    ("riv_encoder_create_mem", c_int, [POINTER(EncoderConfigC), POINTER(c_void_p)]),
    ("riv_encode_frame", c_int, [c_void_p, c_void_p, c_size_t, c_uint32, c_uint32, c_uint64, POINTER(FrameStatsC)]),
    ("riv_encode_frame_device", c_int, [c_void_p, c_void_p, c_size_t, c_uint32, c_uint32, c_uint64, POINTER(FrameStatsC)]),
    ("riv_encode_batch_step", c_int, [POINTER(c_void_p), c_size_t, POINTER(c_void_p), POINTER(c_size_t), POINTER(c_uint64), c_int,
                                      POINTER(FrameStatsC)]),
    ("riv_encoder_flush", c_int, [c_void_p]),
    ("riv_encoder_pop_stats", c_int, [c_void_p, POINTER(FrameStatsC), c_size_t, POINTER(c_size_t)]),
    ("riv_encoder_finish", c_int, [c_void_p]),
    ("riv_encoder_frame_count", c_uint64, [c_void_p]),
    ("riv_encoder_header", c_int, [c_void_p, POINTER(VideoHeaderC)]),
    ("riv_encoder_get_config", c_int, [c_void_p, POINTER(EncoderConfigC)]),
    ("riv_encoder_output", c_int, [c_void_p, POINTER(POINTER(c_uint8)), POINTER(c_size_t)]),
    ("riv_encoder_destroy", None, [c_void_p]),
    ("riv_encoder_debug_blocks", c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("riv_encoder_debug_yuv", c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    ("riv_encoder_debug_coeffs", c_int, [c_void_p, c_void_p]),
    ("riv_encoder_debug_aq", c_int, [c_void_p, c_void_p]),
    ("riv_encoder_debug_pass_limit", c_int, [c_void_p, c_int32]),
    ("riv_encoder_debug_split_coding", c_int, [c_void_p, c_int32]),
    ("riv_encoder_debug_model_cap", c_int, [c_void_p, c_uint32]),
    ("riv_decoder_create", c_int, [POINTER(ReaderVtblC), c_void_p, c_int32, POINTER(c_void_p)]),
    ("riv_decoder_create_mem", c_int, [c_void_p, c_size_t, c_int32, POINTER(c_void_p)]),
    ("riv_decode_frame", c_int, [c_void_p, c_void_p, c_size_t, POINTER(DecodedInfoC)]),
    ("riv_decoder_decode_frame_null", c_int, [c_void_p]),
    ("riv_decoder_has_more_frames", c_int, [c_void_p]),
    ("riv_decoder_header", c_int, [c_void_p, POINTER(VideoHeaderC)]),
    ("riv_decoder_set_skip_residuals", c_int, [c_void_p, c_int]),
    ("riv_decoder_skip_residuals", c_int, [c_void_p]),
    ("riv_decoder_current_frame", c_uint64, [c_void_p]),
    ("riv_decoder_drain_timings", c_int, [c_void_p, POINTER(DecodeTimingsC)]),
    ("riv_decoder_set_profiling", c_int, [c_void_p, c_int]),
    ("riv_decoder_debug_yuv", c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    ("riv_decoder_destroy", None, [c_void_p]),
    ("riv_gpu_build_predicted", c_int, [c_int32, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_void_p, c_void_p, c_void_p, c_void_p]),
    ("riv_gpu_deblock_yuv420", c_int, [c_int32, c_void_p, c_void_p, c_void_p, c_uint32, c_uint32, c_int32, c_int32, c_void_p]),
    ("riv_gpu_dct8_aq_plane", c_int, [c_int32, c_void_p, c_uint32, c_uint32, c_void_p, c_float, c_void_p, c_void_p]),
    ("riv_gpu_dct16_aq_plane", c_int, [c_int32, c_void_p, c_uint32, c_uint32, c_void_p, c_float, c_void_p, c_void_p]),
    ("riv_gpu_aq_steps", c_int, [c_int32, c_void_p, c_uint32, c_uint32, c_float, c_void_p]),
    ("riv_gpu_aq_steps_noalloc", c_int, [c_int32, c_void_p, c_uint32, c_uint32, c_float, c_void_p, c_void_p]),
    ("riv_gpu_aq_mean", c_int, [c_int32, c_void_p, c_size_t, c_float, c_void_p, c_void_p]),
    ("riv_gpu_me_tables", c_int, [c_int32, c_void_p, c_void_p, c_uint32, c_uint32, c_int32, c_void_p, c_void_p]),
    ("riv_gpu_logf_exp2f", c_int, [c_int32, c_void_p, c_size_t, c_int32, c_void_p]),
    ("riv_encoder_set_profiling", c_int, [c_void_p, c_int]),
    ("riv_encoder_set_profiling_sample", c_int, [c_void_p, c_uint32]),
    ("riv_encoder_stage_times", c_int, [c_void_p, c_void_p, c_void_p]),
    ("riv_encoder_set_batch_autotune", c_int, [c_void_p, c_int32]),
    ("riv_encoder_batch_tuning", c_int, [c_void_p, c_void_p]),
    ("riv_mvnet_param_count", c_size_t, []),
    ("riv_mvnet_create", c_int, [c_int32, c_uint64, POINTER(c_void_p)]),
    ("riv_mvnet_destroy", None, [c_void_p]),
    ("riv_mvnet_get_params", c_int, [c_void_p, c_void_p, c_size_t]),
    ("riv_mvnet_set_params", c_int, [c_void_p, c_void_p, c_size_t]),
    ("riv_mvnet_set_sources", c_int, [c_void_p, c_void_p, c_size_t]),
    ("riv_mvnet_train_step", c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_float, c_void_p]),
    ("riv_mvnet_gradients", c_int, [c_void_p, c_void_p, c_void_p, c_size_t, c_void_p, c_void_p]),
    ("riv_mvnet_forward", c_int, [c_void_p, c_void_p, c_size_t, c_int32, c_void_p]),
    ("riv_mvnet_propose", c_int, [c_void_p, c_void_p, c_size_t, c_float, c_int32, c_void_p]),
    ("riv_encoder_set_mvnet", c_int, [c_void_p, c_void_p, c_float]),
    ("riv_encoder_set_learned_passes", c_int, [c_void_p, c_int32]),
    ("riv_encoder_learned_stats", c_int, [c_void_p, c_void_p]),
    ("riv_last_error", c_char_p, []),
]

_lib = None


def lib():
    """The loaded backend library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"RIV GPU backend not built: {LIB_PATH} is missing (run __graft_entry__.build())")
        # PyTorch ships its own HIP runtime under the same soname (libamdhip64.so.7). The first one
        # loaded serves the whole process, and PyTorch cannot use the GPU through another's, so
        # where PyTorch is installed it is loaded first and this library shares its runtime.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in PROTOTYPES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def last_error():
    msg = lib().riv_last_error()
    return msg.decode(errors="replace") if msg else ""
