"""ctypes wrapper of the parity oracle (oracle/riv_oracle.cpp -> oracle/liboracle_riv.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg as the checker / CPU baseline; never by the product package.
"""
import ctypes
import os
from ctypes import POINTER, Structure, byref, c_double, c_float, c_int, c_int8, c_int32, c_int64, c_size_t, c_uint8, \
    c_uint16, c_uint32, c_uint64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle_riv.so")


class ROConfig(Structure):
    _fields_ = [("display_width", c_uint32), ("display_height", c_uint32), ("fps", c_uint32),
                ("keyframe_mode", c_int32), ("keyframe_interval", c_uint64),
                ("intra_quality", c_uint8), ("inter_quality", c_uint8), ("search_range", c_uint8), ("skip_threshold", c_uint8),
                ("me_zero_mv_threshold", c_int64), ("me_predictor_threshold", c_int64),
                ("rdo_lambda_mult", c_double), ("inter_dead_zone", c_float), ("me_threads", c_int32)]


class ROStats(Structure):
    _fields_ = [("frame_index", c_uint64), ("frame_type", c_int32), ("total_bytes", c_uint32), ("mv_bytes", c_uint32),
                ("residual_bytes", c_uint32), ("blocks_total", c_uint32), ("blocks_skipped", c_uint32), ("rdo_skips", c_uint32),
                ("mode_counts", c_uint32 * 7), ("bias_dx", c_int32), ("bias_dy", c_int32), ("new_blocks", c_uint32)]


class ROInfo(Structure):
    _fields_ = [("timestamp", c_uint64), ("frame_index", c_uint64), ("frame_type", c_int32), ("width", c_uint32), ("height", c_uint32)]


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"oracle not built: {LIB}")
        L = ctypes.CDLL(LIB)
        P = c_void_p
        sig = {
            "ro_encoder_create": (c_int, [POINTER(ROConfig), POINTER(c_void_p)]),
            "ro_encode_frame": (c_int, [P, P, c_size_t, c_uint32, c_uint32, c_uint64, POINTER(ROStats)]),
            "ro_encoder_finish": (c_int, [P]),
            "ro_encoder_output": (c_size_t, [P, POINTER(POINTER(c_uint8))]),
            "ro_encoder_destroy": (None, [P]),
            "ro_encoder_blocks": (c_uint32, [P]),
            "ro_encoder_last_frame_bytes": (None, [P, POINTER(POINTER(c_uint8)), POINTER(c_size_t)]),
            "ro_encoder_last_yuv": (None, [P, c_int, P, P, P]),
            "ro_encoder_last_blocks": (None, [P, P, P]),
            "ro_encoder_last_skip": (None, [P, P, P]),
            "ro_encoder_last_aq": (None, [P, P]),
            "ro_encoder_last_coeffs": (None, [P, P]),
            "ro_decoder_create": (c_int, [P, c_size_t, POINTER(c_void_p)]),
            "ro_decode_frame": (c_int, [P, P, c_size_t, POINTER(ROInfo)]),
            "ro_decoder_header": (None, [P, P]),
            "ro_decoder_last_yuv": (None, [P, P, P, P]),
            "ro_decoder_destroy": (None, [P]),
            "ro_decoder_set_skip_residuals": (None, [P, c_int]),
            "ro_dct8_aq_plane": (None, [P, c_int, c_int, P, c_float, P, P]),
            "ro_dct16_aq_plane": (None, [P, c_int, c_int, P, c_float, P, P]),
            "ro_basis": (None, [P, P]),
            "ro_intra_tables": (None, [c_float, P, P]),
            "ro_quant_step": (c_float, [c_uint8]),
            "ro_deblock_level": (c_int, [c_float]),
            "ro_deblock": (None, [P, P, P, c_int, c_int, c_int, c_int, P]),
            "ro_rgb_to_yuv420": (None, [P, c_int, c_int, P, P, P]),
            "ro_yuv420_to_rgb": (None, [P, P, P, c_int, c_int, P]),
            "ro_build_predicted": (None, [P, P, P, c_int, c_int, P, P, P, P]),
            "ro_hex_search": (None, [P, P, c_int, c_int, c_int, P, c_int64, c_int64, c_double, c_int, P, P]),
            "ro_sad_int": (c_int64, [P, P, c_int, c_int, c_int, c_int, c_int, c_int]),
            "ro_sad_hp": (c_int64, [P, P, c_int, c_int, c_int, c_int, c_int, c_int]),
            "ro_mv_cost": (c_int64, [c_int] * 8 + [c_double]),
            "ro_lambda": (c_double, [c_uint8, c_double]),
            "ro_rdo_skip": (c_int, [c_uint8, c_double, c_uint8, c_int64, c_int]),
            "ro_aq_steps": (None, [P, c_int, c_int, c_float, P]),
            "ro_binprob_observe": (c_uint16, [c_uint16, c_int]),
            "ro_binprob_p0": (c_uint8, [c_uint16]),
            "ro_rans_encode": (c_size_t, [P, P, c_size_t, P, c_size_t]),
            "ro_rans_decode": (c_int, [P, c_size_t, P, c_size_t, P]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype, f.argtypes = r, a
        _L = L
    return _L


def config(width, height, fps=30, intra_quality=85, inter_quality=80, search_range=12, skip_threshold=3,
           keyframe_interval=1000, me_threads=0, zero_thr=0, pred_thr=0, lambda_mult=0.49, dead_zone=0.75, keyframe_mode=0):
    """keyframe_mode: 0 Fixed(keyframe_interval), 1 AllIntra, 2 Automatic (config.rs:5-19)."""
    return ROConfig(width, height, fps, keyframe_mode, keyframe_interval, intra_quality, inter_quality, search_range, skip_threshold,
                    zero_thr, pred_thr, lambda_mult, dead_zone, me_threads)


class OracleEncoder:
    def __init__(self, cfg: ROConfig):
        self.cfg = cfg
        self._h = c_void_p()
        rc = lib().ro_encoder_create(byref(cfg), byref(self._h))
        if rc:
            raise ValueError(f"oracle encoder create failed ({rc})")
        self.sw = (cfg.display_width + 15) // 16 * 16
        self.sh = (cfg.display_height + 15) // 16 * 16
        self.nb = (self.sw // 16) * (self.sh // 16)

    def encode(self, rgb: np.ndarray, ts: int = 0) -> ROStats:
        buf = np.ascontiguousarray(rgb, np.uint8)
        st = ROStats()
        rc = lib().ro_encode_frame(self._h, buf.ctypes.data, buf.size, self.cfg.display_width, self.cfg.display_height, ts, byref(st))
        if rc:
            raise RuntimeError(f"oracle encode failed ({rc})")
        return st

    def finish(self) -> bytes:
        lib().ro_encoder_finish(self._h)
        p = POINTER(c_uint8)()
        n = lib().ro_encoder_output(self._h, byref(p))
        return ctypes.string_at(p, n)

    def last_frame_bytes(self) -> bytes:
        p = POINTER(c_uint8)(); n = c_size_t()
        lib().ro_encoder_last_frame_bytes(self._h, byref(p), byref(n))
        return ctypes.string_at(p, n.value)

    def last_yuv(self, which):
        y = np.zeros((self.sh, self.sw), np.uint8); u = np.zeros((self.sh // 2, self.sw // 2), np.uint8); v = np.zeros_like(u)
        lib().ro_encoder_last_yuv(self._h, which, y.ctypes.data, u.ctypes.data, v.ctypes.data)
        return y, u, v

    def last_blocks(self):
        rec = np.zeros((self.nb, 12), np.int8); sad = np.zeros(self.nb, np.int64)
        lib().ro_encoder_last_blocks(self._h, rec.ctypes.data, sad.ctypes.data)
        rdo = np.zeros(self.nb, np.uint8); opt = np.zeros(self.nb, np.uint8)
        lib().ro_encoder_last_skip(self._h, rdo.ctypes.data, opt.ctypes.data)
        return rec, sad, rdo, opt

    def last_aq(self):
        a = np.zeros(self.nb, np.float32)
        lib().ro_encoder_last_aq(self._h, a.ctypes.data)
        return a

    def last_coeffs(self):
        c = np.zeros((self.nb, 384), np.int16)
        lib().ro_encoder_last_coeffs(self._h, c.ctypes.data)
        return c

    def __del__(self):
        try:
            if self._h:
                lib().ro_encoder_destroy(self._h)
        except Exception:
            pass


class OracleDecoder:
    def __init__(self, data: bytes):
        self._data = bytes(data)
        self._h = c_void_p()
        rc = lib().ro_decoder_create(self._data, len(self._data), byref(self._h))
        if rc:
            raise ValueError(f"oracle decoder create failed ({rc})")
        hdr = (c_uint32 * 6)()
        lib().ro_decoder_header(self._h, hdr)
        self.dw, self.dh, self.sw, self.sh, self.fps, self.frame_count = list(hdr)

    def decode(self):
        out = np.zeros((self.dh, self.dw, 3), np.uint8)
        info = ROInfo()
        rc = lib().ro_decode_frame(self._h, out.ctypes.data, out.size, byref(info))
        if rc:
            raise RuntimeError(f"oracle decode failed ({rc})")
        return out, info

    def set_skip_residuals(self, skip: bool):
        lib().ro_decoder_set_skip_residuals(self._h, 1 if skip else 0)

    def last_yuv(self):
        y = np.zeros((self.sh, self.sw), np.uint8); u = np.zeros((self.sh // 2, self.sw // 2), np.uint8); v = np.zeros_like(u)
        lib().ro_decoder_last_yuv(self._h, y.ctypes.data, u.ctypes.data, v.ctypes.data)
        return y, u, v

    def __del__(self):
        try:
            if self._h:
                lib().ro_decoder_destroy(self._h)
        except Exception:
            pass
