"""Single-stage GPU entry points (the kernel-level seams of the C ABI) on host numpy arrays.

Each mirrors a reference internal plug point (build_predicted, deblock_yuv420, the reitero_dct
plane functions, compute_aq_quant_steps, the motion search's SAD kernels) so that a stage can be
checked on its own against the function it replaces.
"""
import numpy as np

from . import _lib as L
from .api import _check


def _p(a):
    return a.ctypes.data


def build_predicted(y, u, v, mv4, device=0):
    """Motion-compensated YUV420 prediction; mv4 is (nblocks, 4) int8 (dx, dy, sx, sy)."""
    h, w = y.shape
    y, u, v = (np.ascontiguousarray(a, np.uint8) for a in (y, u, v))
    mv = np.ascontiguousarray(mv4, np.int8)
    oy, ou, ov = np.zeros_like(y), np.zeros_like(u), np.zeros_like(v)
    _check(L.lib().riv_gpu_build_predicted(device, _p(y), _p(u), _p(v), w, h, _p(mv), _p(oy), _p(ou), _p(ov)))
    return oy, ou, ov


def deblock(y, u, v, level, intra, mask=None, device=0):
    """In-loop deblocking on copies of the planes; returns the filtered planes."""
    h, w = y.shape
    y, u, v = (np.array(a, np.uint8, copy=True, order="C") for a in (y, u, v))
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    _check(L.lib().riv_gpu_deblock_yuv420(device, _p(y), _p(u), _p(v), w, h, level, int(intra), _p(m) if m is not None else None))
    return y, u, v


def dct_plane(plane, qs, dead_zone, n=16, device=0):
    """encode + decode of one residual plane in n x n blocks: (coefficients, reconstruction)."""
    h, w = plane.shape
    p = np.ascontiguousarray(plane, np.int16)
    q = np.ascontiguousarray(qs, np.float32)
    c = np.zeros(h * w, np.int16)
    r = np.zeros_like(p)
    f = L.lib().riv_gpu_dct8_aq_plane if n == 8 else L.lib().riv_gpu_dct16_aq_plane
    _check(f(device, _p(p), w, h, _p(q), float(dead_zone), _p(c), _p(r)))
    return c, r


def aq_steps(y, base_step, device=0):
    h, w = y.shape
    y = np.ascontiguousarray(y, np.uint8)
    out = np.zeros((w // 16) * (h // 16), np.float32)
    _check(L.lib().riv_gpu_aq_steps(device, _p(y), w, h, float(base_step), _p(out)))
    return out


def aq_mean(logs, base_step, device=0):
    """The adaptive quantiser's mean of positive per-MB logs and the steps derived from it."""
    x = np.ascontiguousarray(logs, np.float32)
    mean = np.zeros(1, np.float32)
    steps = np.zeros_like(x)
    _check(L.lib().riv_gpu_aq_mean(device, _p(x), x.size, float(base_step), _p(mean), _p(steps)))
    return mean[0], steps


def logf_exp2f(x, which, device=0):
    """which 0: logf, 1: exp2f, elementwise on the device."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    _check(L.lib().riv_gpu_logf_exp2f(device, _p(x), x.size, int(which), _p(out)))
    return out


def me_tables(cur, ref, sr, device=0):
    """(int_sad, hp_sad) per block: (nb, (2sr+1)^2) and (nb, (4sr+3)^2) uint16, row-major by dy / hy."""
    h, w = cur.shape
    cur = np.ascontiguousarray(cur, np.uint8)
    ref = np.ascontiguousarray(ref, np.uint8)
    nb = (w // 16) * (h // 16)
    it = np.zeros((nb, (2 * sr + 1) ** 2), np.uint16)
    hp = np.zeros((nb, (4 * sr + 3) ** 2), np.uint16)
    _check(L.lib().riv_gpu_me_tables(device, _p(cur), _p(ref), w, h, sr, _p(it), _p(hp)))
    return it, hp
