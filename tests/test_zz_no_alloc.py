"""The allocation-free AQ entry point (riv_gpu_aq_steps_noalloc: device buffers in and out, its
scratch in a static device array) gives the same steps, bit for bit, as riv_gpu_aq_steps (host
buffers, allocating) on two calls in a row, the second on a smaller plane, so that the shared
scratch's leftovers from the first call must not leak into the second; and it rejects planes
beyond its scratch without touching the GPU."""
import ctypes

import numpy as np
import pytest

from riv2_amd import _lib

pytestmark = pytest.mark.gpu


def _plane(w, h, seed):
    rng = np.random.default_rng(seed)
    # textured with flat patches, so variances span the floor (4) and large values
    y = rng.integers(16, 236, size=(h, w), dtype=np.uint8)
    y[: h // 3, : w // 4] = 128
    return np.ascontiguousarray(y)


def _reference_steps(y, base):
    h, w = y.shape
    out = np.zeros((h // 16) * (w // 16), np.float32)
    rc = _lib.lib().riv_gpu_aq_steps(0, y.ctypes.data, w, h, ctypes.c_float(base), out.ctypes.data)
    assert rc == 0, _lib.last_error()
    return out


def _noalloc_steps(y, base):
    import torch
    h, w = y.shape
    dy = torch.from_numpy(y).to("cuda:0")
    out = torch.full(((h // 16) * (w // 16),), -1.0, dtype=torch.float32, device="cuda:0")
    torch.cuda.synchronize()
    rc = _lib.lib().riv_gpu_aq_steps_noalloc(0, ctypes.c_void_p(dy.data_ptr()), w, h, ctypes.c_float(base),
                                             ctypes.c_void_p(out.data_ptr()), None)
    assert rc == 0, _lib.last_error()
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_noalloc_matches_allocating_entry_point_twice():
    base = 23.424242
    for k, (w, h) in enumerate([(1920, 1088), (352, 288)]):
        print(f"[no-alloc] call {k}: {w}x{h}", flush=True)
        y = _plane(w, h, 40 + k)
        want = _reference_steps(y, base)
        got = _noalloc_steps(y, base)
        assert got.view(np.uint32).tolist() == want.view(np.uint32).tolist(), f"call {k} ({w}x{h}): steps differ"


def test_noalloc_rejects_planes_beyond_its_scratch():
    w, h = 4096, 2304   # 36,864 macroblocks > RIV_AQ_NOALLOC_MAX_BLOCKS
    rc = _lib.lib().riv_gpu_aq_steps_noalloc(0, ctypes.c_void_p(4096), w, h, ctypes.c_float(1.0), ctypes.c_void_p(4096), None)
    assert rc == 22 and "exceed" in _lib.last_error()
