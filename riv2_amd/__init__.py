"""MI355X (gfx950) backend of the RIV video codec.

The encoder/decoder objects mirror reitero_encode / reitero_decode; all per-frame analysis and
reconstruction runs in hand-written HIP kernels (riv2_amd/csrc/kernels.hip) behind the C ABI of
include/riv_gpu.h, with the rANS entropy coder on the host.
"""
from .api import (Decoder, DecodedFrame, DecodeError, EncodeError, EncodeFrameStats, Encoder, EncoderConfig, Frame,
                  KeyframeInterval, RivError, VecWriter, FRAME_TYPE_INTER, FRAME_TYPE_INTRA, MODE_NAMES, encode_batch_step)
from ._lib import lib, last_error, LIB_PATH, SYNTH_PATH

__all__ = ["Decoder", "DecodedFrame", "DecodeError", "EncodeError", "EncodeFrameStats", "Encoder", "EncoderConfig",
           "Frame", "KeyframeInterval", "RivError", "VecWriter", "FRAME_TYPE_INTER", "FRAME_TYPE_INTRA", "MODE_NAMES",
           "lib", "last_error", "LIB_PATH", "SYNTH_PATH", "encode_batch_step"]
