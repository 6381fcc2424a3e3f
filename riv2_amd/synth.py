"""Seeded synthetic RGB24 sequences (libriv_synth.so, plain C++; no GPU needed).

Presets follow the workload descriptions: the CIF reference case (352x288, 30 frames, global
motion up to 3 px at half-pel steps, 4 rectangles of 32-96 px at up to 6 px/frame, noise sigma 2)
and the 1080p case (1920x1080, 50 frames, 8 px, 12 rectangles of 64-320 px at up to 16 px/frame).
"""
import ctypes
import os
from ctypes import POINTER, Structure, c_float, c_int, c_uint32, c_uint64, c_void_p

import numpy as np

from ._lib import SYNTH_PATH


class SynthParams(Structure):
    _fields_ = [("width", c_uint32), ("height", c_uint32), ("frames", c_uint32), ("seed", c_uint64),
                ("global_motion", c_float), ("n_rects", c_uint32), ("rect_min", c_uint32), ("rect_max", c_uint32),
                ("rect_velocity", c_float), ("noise_sigma", c_float), ("scene_cut_interval", c_uint32),
                ("rect_zoom", c_float), ("rect_rotate_deg", c_float)]


_synth = None


def _lib():
    global _synth
    if _synth is None:
        if not os.path.exists(SYNTH_PATH):
            raise RuntimeError(f"synthetic generator not built: {SYNTH_PATH} is missing")
        L = ctypes.CDLL(SYNTH_PATH)
        L.riv_synth_create.restype = c_int
        L.riv_synth_create.argtypes = [POINTER(SynthParams), POINTER(c_void_p)]
        L.riv_synth_frames.restype = c_int
        L.riv_synth_frames.argtypes = [c_void_p, c_uint32, c_uint32, c_void_p, c_int]
        L.riv_synth_destroy.restype = None
        L.riv_synth_destroy.argtypes = [c_void_p]
        _synth = L
    return _synth


def cif_params(seed: int = 1, frames: int = 30) -> SynthParams:
    return SynthParams(352, 288, frames, seed, 3.0, 4, 32, 96, 6.0, 2.0, 0)


def hd_params(seed: int = 1, frames: int = 50) -> SynthParams:
    return SynthParams(1920, 1080, frames, seed, 8.0, 12, 64, 320, 16.0, 2.0, 0)


def custom_params(width, height, frames, seed=1, global_motion=3.0, n_rects=4, rect_min=32, rect_max=96,
                  rect_velocity=6.0, noise_sigma=2.0, scene_cut_interval=0, rect_zoom=0.0, rect_rotate_deg=0.0) -> SynthParams:
    return SynthParams(width, height, frames, seed, global_motion, n_rects, rect_min, rect_max, rect_velocity,
                       noise_sigma, scene_cut_interval, rect_zoom, rect_rotate_deg)


def stress_params(width, height, frames, seed=1) -> SynthParams:
    """High-motion stress content: global motion up to 24 px per frame, 4 rectangles of 32-96 px
    moving up to 32 px per frame, each zooming up to 0.5% and rotating up to 1 degree per frame, a
    hard scene cut every 25 frames, noise sigma 4."""
    return custom_params(width, height, frames, seed, global_motion=24.0, n_rects=4, rect_min=32, rect_max=96,
                         rect_velocity=32.0, noise_sigma=4.0, scene_cut_interval=25, rect_zoom=0.005, rect_rotate_deg=1.0)


class Sequence:
    """A generated sequence; frames are rendered on demand, byte-identically for a given seed."""

    def __init__(self, params: SynthParams):
        self.params = params
        self._h = c_void_p()
        rc = _lib().riv_synth_create(ctypes.byref(params), ctypes.byref(self._h))
        if rc != 0:
            raise ValueError("invalid synthetic sequence parameters")

    def frames(self, start: int = 0, count: int = None, threads: int = None) -> np.ndarray:
        p = self.params
        if count is None:
            count = p.frames - start
        out = np.empty((count, p.height, p.width, 3), np.uint8)
        rc = _lib().riv_synth_frames(self._h, start, count, out.ctypes.data, threads or min(16, os.cpu_count() or 1))
        if rc != 0:
            raise ValueError("frame range outside the sequence")
        return out

    def __del__(self):
        try:
            if self._h:
                _lib().riv_synth_destroy(self._h)
        except Exception:
            pass
