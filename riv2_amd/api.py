"""Python mirror of the reference encoder/decoder API on top of the C ABI.

Names, argument meaning and error behaviour follow reitero_encode (EncoderConfig, KeyframeInterval,
Frame, Encoder, VecWriter; reitero_encode/src/{config,encoder,writer,error}.rs) and reitero_decode
(Decoder, DecodedFrame; reitero_decode/src/decoder.rs). Every call runs the gfx950 backend.
"""
import ctypes
from ctypes import byref, c_size_t, c_uint8, c_uint64, c_void_p, POINTER
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib as L

FRAME_TYPE_INTRA = 1
FRAME_TYPE_INTER = 2
MODE_NAMES = ["Zero", "Nearest", "Near", "TopRight", "TopLeft", "New", "Temporal"]


class RivError(Exception):
    """A non-zero status from the C ABI; `kind` names the reference error variant."""

    def __init__(self, status: int, message: str = ""):
        self.status = status
        self.kind = L.STATUS_NAMES.get(status, f"status{status}")
        super().__init__(f"{self.kind}: {message}" if message else self.kind)


class EncodeError(RivError):
    pass


class DecodeError(RivError):
    pass


def _check(status: int, exc=RivError):
    if status != L.RIV_OK:
        raise exc(status, L.last_error())


class KeyframeInterval:
    """reitero_encode/src/config.rs:5-19"""

    def __init__(self, mode: int, interval: int = 0):
        self.mode, self.interval = mode, interval

    @staticmethod
    def fixed(n: int) -> "KeyframeInterval":
        return KeyframeInterval(0, n)

    @staticmethod
    def all_intra() -> "KeyframeInterval":
        return KeyframeInterval(1)

    @staticmethod
    def automatic() -> "KeyframeInterval":
        return KeyframeInterval(2)


class EncoderConfig:
    """EncoderConfig::new + with_* builders (reitero_encode/src/config.rs:24-127)."""

    def __init__(self, display_width: int, display_height: int, fps: int):
        c = L.EncoderConfigC()
        L.lib().riv_encoder_config_default(byref(c), display_width, display_height, fps)
        self._c = c

    def __getattr__(self, name):
        if name == "keyframe_interval":
            c = self.__dict__["_c"]
            return KeyframeInterval(c.keyframe_mode, c.keyframe_interval)
        if name != "_c" and name in dict(L.EncoderConfigC._fields_):
            return getattr(self.__dict__["_c"], name)
        raise AttributeError(name)

    def __setattr__(self, name, value):
        if name == "_c":
            object.__setattr__(self, name, value)
        elif name == "keyframe_interval":
            self._c.keyframe_mode, self._c.keyframe_interval = value.mode, value.interval
        elif name in dict(L.EncoderConfigC._fields_):
            setattr(self._c, name, value)
        else:
            raise AttributeError(name)

    def _with(self, **kw):
        for k, v in kw.items():
            setattr(self, k, v)
        return self

    def with_rdo_lambda_mult(self, m): return self._with(rdo_lambda_mult=m)
    def with_inter_dead_zone(self, d): return self._with(inter_dead_zone=d)
    def with_intra_quality(self, q): return self._with(intra_quality=q)
    def with_inter_quality(self, q): return self._with(inter_quality=q)
    def with_keyframe_interval(self, k): return self._with(keyframe_interval=k)
    def with_search_range(self, r): return self._with(search_range=r)
    def with_mv_deflate_level(self, v): return self._with(mv_deflate_level=v)
    def with_skip_threshold(self, t): return self._with(skip_threshold=t)
    def with_me_zero_mv_threshold(self, t): return self._with(me_zero_mv_threshold=t)
    def with_me_predictor_threshold(self, t): return self._with(me_predictor_threshold=t)
    def with_device(self, d): return self._with(device_id=d)
    def with_pipeline_depth(self, n): return self._with(pipeline_depth=n)

    def validate(self):
        _check(L.lib().riv_encoder_config_validate(byref(self._c)), EncodeError)


@dataclass
class Frame:
    """Raw RGB24 input frame (reitero_encode/src/encoder.rs:18-34)."""
    data: object
    width: int
    height: int
    timestamp: int = 0


@dataclass
class EncodeFrameStats:
    frame_index: int
    timestamp_ms: int
    frame_type: int
    total_bytes: int
    mv_bytes: int
    residual_bytes: int
    storage_width: int
    storage_height: int
    blocks_total: int
    blocks_skipped: int
    mv_mode_counts: List[int]
    mv_new_blocks: int
    mv_bias_dx: int
    mv_bias_dy: int
    rdo_lambda: float
    rdo_blocks_evaluated: int
    rdo_blocks_selected_skip: int
    gpu_ms: float
    host_entropy_ms: float
    host_tail_ms: float = 0.0
    me_passes: int = 0
    # host timeline, ms since the encoder was created: submitted, GPU work seen done, coded, written
    timeline: tuple = ()
    residual_symbols_on_gpu: bool = False

    @staticmethod
    def from_c(s):
        return EncodeFrameStats(s.frame_index, s.timestamp_ms, s.frame_type, s.total_bytes, s.mv_bytes, s.residual_bytes,
                                s.storage_width, s.storage_height, s.blocks_total, s.blocks_skipped, list(s.mv_mode_counts),
                                s.mv_new_blocks, s.mv_bias_dx, s.mv_bias_dy, s.rdo_lambda, s.rdo_blocks_evaluated,
                                s.rdo_blocks_selected_skip, s.gpu_ms, s.host_entropy_ms, s.host_tail_ms, s.me_passes,
                                (s.t_submit_ms, s.t_gpu_done_ms, s.t_coded_ms, s.t_written_ms), bool(s.residual_symbols_on_gpu))


def encode_batch_step(encoders, frames, timestamps=None, on_device=False):
    """One frame for each encoder (streams of any sizes) in one batched GPU step: the inter frames'
    motion searches run as one launch. frames[i]: HxWx3 uint8 arrays, or device pointers (ints) with
    on_device=True. Returns the per-frame stats (complete only for synchronous encoders)."""
    n = len(encoders)
    if n == 0 or len(frames) != n:
        raise ValueError("one frame per encoder")
    hs = (c_void_p * n)(*[e._h for e in encoders])
    keep, ptrs, lens = [], (c_void_p * n)(), (c_size_t * n)()
    for i, (e, f) in enumerate(zip(encoders, frames)):
        if on_device:
            ptrs[i] = c_void_p(int(f))
            lens[i] = e.config.display_width * e.config.display_height * 3
        else:
            buf = _as_u8(f)
            keep.append(buf)
            ptrs[i] = c_void_p(buf.ctypes.data)
            lens[i] = buf.size
    ts = (c_uint64 * n)(*(timestamps if timestamps is not None else [0] * n))
    st = (L.FrameStatsC * n)()
    _check(L.lib().riv_encode_batch_step(hs, n, ptrs, lens, ts, 1 if on_device else 0, st), EncodeError)
    return [EncodeFrameStats.from_c(st[i]) for i in range(n)]


class VecWriter:
    """In-memory VideoWriter (reitero_encode/src/writer.rs:21-77)."""

    def __init__(self):
        self.buffer = bytearray()
        self.pos = 0

    def write(self, data: bytes):
        end = self.pos + len(data)
        if end > len(self.buffer):
            self.buffer.extend(b"\0" * (end - len(self.buffer)))
        self.buffer[self.pos:end] = data
        self.pos = end

    def position(self) -> int:
        return self.pos

    def seek(self, pos: int):
        if pos > len(self.buffer):
            self.buffer.extend(b"\0" * (pos - len(self.buffer)))
        self.pos = pos

    def flush(self):
        pass

    def into_inner(self) -> bytes:
        return bytes(self.buffer)


def _as_u8(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8).reshape(-1)
    return np.frombuffer(bytes(data), dtype=np.uint8)


class Encoder:
    """Encoder<W: VideoWriter> (reitero_encode/src/encoder.rs:40-981).

    With `writer=None` the encoder writes to an internal VecWriter-like buffer (see `output()`);
    a Python object with write/position/seek/flush is driven through the C writer vtable.
    """

    def __init__(self, config: EncoderConfig, writer=None):
        self._h = c_void_p()
        self._writer = writer
        self.config = config
        if writer is None:
            _check(L.lib().riv_encoder_create_mem(byref(config._c), byref(self._h)), EncodeError)
        else:
            def w(_, p, n):
                try:
                    writer.write(ctypes.string_at(p, n)); return 0
                except Exception:
                    return 1

            def pos(_):
                return writer.position()

            def sk(_, p):
                try:
                    writer.seek(p); return 0
                except Exception:
                    return 1

            def fl(_):
                try:
                    writer.flush(); return 0
                except Exception:
                    return 1
            self._vt = L.WriterVtblC(L.WRITE_FN(w), L.POS_FN(pos), L.SEEK_FN(sk), L.FLUSH_FN(fl))
            _check(L.lib().riv_encoder_create(byref(config._c), byref(self._vt), None, byref(self._h)), EncodeError)
        self._bw = config.storage_width // 16
        self._bh = config.storage_height // 16

    def close(self):
        if self._h:
            L.lib().riv_encoder_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def encode_frame_with_stats(self, frame: Frame) -> EncodeFrameStats:
        buf = _as_u8(frame.data)
        s = L.FrameStatsC()
        _check(L.lib().riv_encode_frame(self._h, buf.ctypes.data, buf.size, frame.width, frame.height, frame.timestamp, byref(s)), EncodeError)
        return EncodeFrameStats.from_c(s)

    def encode_frame(self, frame: Frame):
        self.encode_frame_with_stats(frame)

    def encode_frame_device(self, ptr: int, nbytes: int, width: int, height: int, timestamp: int = 0) -> EncodeFrameStats:
        """RGB24 frame already in device memory of the encoder's GPU (e.g. a torch tensor's data_ptr())."""
        s = L.FrameStatsC()
        _check(L.lib().riv_encode_frame_device(self._h, c_void_p(ptr), nbytes, width, height, timestamp, byref(s)), EncodeError)
        return EncodeFrameStats.from_c(s)

    def flush(self):
        """Waits until every submitted frame is written (pipelined mode)."""
        _check(L.lib().riv_encoder_flush(self._h), EncodeError)

    def pop_stats(self, cap: int = 256):
        """Stats of completed frames in encode order (pipelined mode)."""
        arr = (L.FrameStatsC * cap)()
        n = c_size_t()
        _check(L.lib().riv_encoder_pop_stats(self._h, arr, cap, byref(n)), EncodeError)
        return [EncodeFrameStats.from_c(arr[i]) for i in range(n.value)]

    def finish(self):
        _check(L.lib().riv_encoder_finish(self._h), EncodeError)

    def frame_count(self) -> int:
        return int(L.lib().riv_encoder_frame_count(self._h))

    def header(self):
        h = L.VideoHeaderC()
        _check(L.lib().riv_encoder_header(self._h, byref(h)), EncodeError)
        return h

    def get_config(self) -> EncoderConfig:
        """Encoder::config (reitero_encode/src/encoder.rs:974-976): the configuration in use."""
        c = EncoderConfig(1, 1, 1)
        _check(L.lib().riv_encoder_get_config(self._h, byref(c._c)), EncodeError)
        return c

    def output(self) -> bytes:
        p = POINTER(c_uint8)()
        n = c_size_t()
        _check(L.lib().riv_encoder_output(self._h, byref(p), byref(n)), EncodeError)
        return ctypes.string_at(p, n.value)

    STAGES = ["rgb_to_yuv", "motion_search", "mc", "aq", "residual", "deblock", "intra", "batch_motion_search", "residual_model"]

    def set_profiling(self, on=True):
        """0/False off, 1 the motion-search stage only, 2/True every stage."""
        level = 2 if on is True else 0 if on is False else int(on)
        _check(L.lib().riv_encoder_set_profiling(self._h, level))

    def set_profiling_sample(self, every: int = 1):
        """Profile only the frames whose index is a multiple of `every`."""
        _check(L.lib().riv_encoder_set_profiling_sample(self._h, int(every)))

    def stage_times(self):
        ms = np.zeros(len(self.STAGES), np.float64)   # RIV_NUM_STAGES entries
        n = np.zeros(len(self.STAGES), np.uint32)
        _check(L.lib().riv_encoder_stage_times(self._h, ms.ctypes.data, n.ctypes.data))
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(self.STAGES)}

    # ---- batch-step launch-shape autotuner (no reference counterpart) ----
    def set_batch_autotune(self, on: bool = True):
        _check(L.lib().riv_encoder_set_batch_autotune(self._h, 1 if on else 0), EncodeError)

    def batch_tuning(self):
        """The latest batch step led by this encoder: dict of waves, groups, slots, candidates,
        settled, chosen."""
        o = np.zeros(6, np.int32)
        _check(L.lib().riv_encoder_batch_tuning(self._h, o.ctypes.data), EncodeError)
        return {"waves": int(o[0]), "groups": int(o[1]), "slots": int(o[2]), "candidates": int(o[3]), "settled": bool(o[4]),
                "chosen": int(o[5])}

    # ---- learned form (no reference counterpart; see riv2_amd.learned) ----
    def set_mvnet(self, net, confidence: float = 0.5):
        """Seeds the motion search with `net`'s proposals from the next P-frame (None: plain form)."""
        self._mvnet = net   # the network must outlive its use
        _check(L.lib().riv_encoder_set_mvnet(self._h, net._h if net is not None else None, float(confidence)), EncodeError)

    def set_learned_passes(self, passes: int):
        """Learned form: relaxation passes per motion search (0: to the raster fixed point)."""
        _check(L.lib().riv_encoder_set_learned_passes(self._h, int(passes)), EncodeError)

    def learned_stats(self):
        """(P-frame blocks searched, blocks with a confident proposal, blocks that took it)."""
        c = np.zeros(3, np.uint64)
        _check(L.lib().riv_encoder_learned_stats(self._h, c.ctypes.data), EncodeError)
        return tuple(int(v) for v in c)

    # ---- last-frame introspection for parity tests ----
    def debug_blocks(self):
        nb = self._bw * self._bh
        rec = np.zeros((nb, 12), np.int8)
        sad = np.zeros(nb, np.int64)
        rdo = np.zeros(nb, np.uint8)
        opt = np.zeros(nb, np.uint8)
        _check(L.lib().riv_encoder_debug_blocks(self._h, rec.ctypes.data, sad.ctypes.data, rdo.ctypes.data, opt.ctypes.data))
        return rec, sad, rdo, opt

    def debug_yuv(self, which: int):
        W, H = self.config.storage_width, self.config.storage_height
        y = np.zeros((H, W), np.uint8)
        u = np.zeros((H // 2, W // 2), np.uint8)
        v = np.zeros((H // 2, W // 2), np.uint8)
        _check(L.lib().riv_encoder_debug_yuv(self._h, which, y.ctypes.data, u.ctypes.data, v.ctypes.data))
        return y, u, v

    def debug_coeffs(self):
        c = np.zeros((self._bw * self._bh, 384), np.int16)
        _check(L.lib().riv_encoder_debug_coeffs(self._h, c.ctypes.data))
        return c

    def debug_pass_limit(self, max_passes: int):
        """Caps the motion search's relaxation passes (0: the proven bound); a search that does not
        settle within the cap fails its frame (no bytes are written for it)."""
        _check(L.lib().riv_encoder_debug_pass_limit(self._h, int(max_passes)), EncodeError)

    def debug_model_cap(self, cap: int):
        """Symbols a P-frame's GPU residual modelling may hold (0: the encoder's limit); a frame
        with more is coded on the host, with the same bytes."""
        _check(L.lib().riv_encoder_debug_model_cap(self._h, int(cap)), EncodeError)

    def debug_split_coding(self, mode: int):
        """How P-frame residuals are coded: 0 modelled on the GPU, rANS32-coded on the host (the
        default); on the host: 1 serially, 2 split across its workers, 3 split when two workers are
        idle. Same bytes."""
        _check(L.lib().riv_encoder_debug_split_coding(self._h, int(mode)), EncodeError)

    def debug_aq(self):
        a = np.zeros(self._bw * self._bh, np.float32)
        _check(L.lib().riv_encoder_debug_aq(self._h, a.ctypes.data))
        return a


@dataclass
class DecodedFrame:
    """reitero_decode/src/decoder.rs:36-49"""
    data: np.ndarray
    width: int
    height: int
    timestamp: int
    frame_index: int
    frame_type: int


class Decoder:
    """Decoder<R: VideoReader> (reitero_decode/src/decoder.rs:89-499) over an in-memory stream, or,
    with `reader`, over any object with read(n) -> bytes (empty at the end), position() and
    seek(pos), driven through the C reader vtable (reitero_decode/src/reader.rs:4-13)."""

    def __init__(self, data: bytes = None, device: int = 0, reader=None):
        self._h = c_void_p()
        if reader is None:
            self._data = bytes(data)
            _check(L.lib().riv_decoder_create_mem(self._data, len(self._data), device, byref(self._h)), DecodeError)
        else:
            self._reader = reader

            def rd(_, buf, cap):
                try:
                    chunk = reader.read(cap)
                except Exception:
                    return 0
                ctypes.memmove(buf, chunk, len(chunk))
                return len(chunk)

            def pos(_):
                return reader.position()

            def sk(_, p):
                try:
                    reader.seek(p); return 0
                except Exception:
                    return 1
            self._vt = L.ReaderVtblC(L.READ_FN(rd), L.POS_FN(pos), L.SEEK_FN(sk))
            _check(L.lib().riv_decoder_create(byref(self._vt), None, device, byref(self._h)), DecodeError)
        h = self.header()
        self._w, self._hh = h.display_width, h.display_height
        self._sw, self._sh = h.storage_width, h.storage_height

    def close(self):
        if self._h:
            L.lib().riv_decoder_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def header(self):
        h = L.VideoHeaderC()
        _check(L.lib().riv_decoder_header(self._h, byref(h)), DecodeError)
        return h

    def decode_frame(self) -> DecodedFrame:
        out = np.zeros((self._hh, self._w, 3), np.uint8)
        info = L.DecodedInfoC()
        _check(L.lib().riv_decode_frame(self._h, out.ctypes.data, out.size, byref(info)), DecodeError)
        return DecodedFrame(out, info.width, info.height, info.timestamp_ms, info.frame_index, info.frame_type)

    def decode_frame_null(self):
        _check(L.lib().riv_decoder_decode_frame_null(self._h), DecodeError)

    def has_more_frames(self) -> bool:
        return bool(L.lib().riv_decoder_has_more_frames(self._h))

    def set_skip_residuals(self, skip: bool):
        _check(L.lib().riv_decoder_set_skip_residuals(self._h, int(bool(skip))), DecodeError)

    def skip_residuals(self) -> bool:
        return bool(L.lib().riv_decoder_skip_residuals(self._h))

    def current_frame(self) -> int:
        """Index of the next frame decode_frame returns (decoder.rs:474-477)."""
        return int(L.lib().riv_decoder_current_frame(self._h))

    def set_profiling(self, on: bool = True):
        """Time the device phases for drain_timings (HIP events on the decoder's stream)."""
        _check(L.lib().riv_decoder_set_profiling(self._h, int(bool(on))), DecodeError)

    def drain_timings(self) -> dict:
        """DecodePhaseTimings (decoder.rs:73-87, :485-489): nanoseconds per phase since the last drain."""
        t = L.DecodeTimingsC()
        _check(L.lib().riv_decoder_drain_timings(self._h, byref(t)), DecodeError)
        return {name: int(getattr(t, name)) for name, _ in L.DecodeTimingsC._fields_}

    def debug_yuv(self):
        y = np.zeros((self._sh, self._sw), np.uint8)
        u = np.zeros((self._sh // 2, self._sw // 2), np.uint8)
        v = np.zeros((self._sh // 2, self._sw // 2), np.uint8)
        _check(L.lib().riv_decoder_debug_yuv(self._h, y.ctypes.data, u.ctypes.data, v.ctypes.data), DecodeError)
        return y, u, v
