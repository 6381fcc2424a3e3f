"""Each GPU stage on its own against the oracle (and, where the reference fixes a value, against
that value): SAD tables of the motion search, transforms, deblocking, motion compensation,
adaptive quantisation and the device logf/exp2f restatement."""
import ctypes

import numpy as np
import pytest

import oracle
from riv2_amd import stages

pytestmark = pytest.mark.gpu
O = oracle.lib()


def vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def frame_pair(w, h, seed, shift=(3, -2)):
    rng = np.random.default_rng(seed)
    cur = rng.integers(0, 256, (h, w), dtype=np.uint8)
    ref = np.roll(np.roll(cur, shift[1], 0), shift[0], 1)
    ref = np.clip(ref.astype(int) + rng.integers(-20, 21, ref.shape), 0, 255).astype(np.uint8)
    return cur, ref


@pytest.mark.parametrize("w,h,sr", [(96, 64, 4), (96, 64, 12), (80, 48, 20), (64, 64, 32)])
def test_me_tables_every_entry(w, h, sr):
    """Every integer and half-pel SAD the search can read, against the oracle's SAD kernels."""
    print(f"[tables] {w}x{h} sr {sr}", flush=True)
    cur, ref = frame_pair(w, h, sr)
    it, hp = stages.me_tables(cur, ref, sr)
    bw = w // 16
    D, D2 = 2 * sr + 1, 4 * sr + 3
    for b in range(it.shape[0]):
        x0, y0 = 16 * (b % bw), 16 * (b // bw)
        want_i = np.array([O.ro_sad_int(vp(ref), vp(cur), w, h, x0, y0, dx, dy)
                           for dy in range(-sr, sr + 1) for dx in range(-sr, sr + 1)], np.int64)
        assert np.array_equal(it[b].astype(np.int64), want_i), f"integer table, block {b}"
        want_h = np.array([O.ro_sad_hp(vp(ref), vp(cur), w, h, x0, y0, hx, hy)
                           for hy in range(-2 * sr - 1, 2 * sr + 2) for hx in range(-2 * sr - 1, 2 * sr + 2)], np.int64)
        assert np.array_equal(hp[b].astype(np.int64), want_h), f"half-pel table, block {b}"


def test_dct8_known_answer_and_oracle():
    kat = np.array([16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69,
                    56, 14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81,
                    104, 113, 92, 49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99], np.int16).reshape(8, 8)
    c, _ = stages.dct_plane(kat, [1.0], 0.5, n=8)
    assert (c[0], c[1], c[8]) == (461, -169, -195)   # reitero_dct/tests/roundtrip.rs:366-400
    rng = np.random.default_rng(1)
    for n in (8, 16):
        plane = rng.integers(-255, 256, (64, 96)).astype(np.int16)
        qs = rng.uniform(0.5, 40.0, (64 // n) * (96 // n)).astype(np.float32)
        gc, gr = stages.dct_plane(plane, qs, 0.75, n=n)
        oc = np.zeros_like(gc)
        orr = np.zeros_like(plane)
        f = O.ro_dct8_aq_plane if n == 8 else O.ro_dct16_aq_plane
        f(vp(plane), 96, 64, vp(qs), 0.75, vp(oc), vp(orr))
        assert np.array_equal(gc, oc) and np.array_equal(gr, orr), n


def test_deblock_against_oracle():
    rng = np.random.default_rng(2)
    w, h = 96, 64
    for level, intra, masked in ((14, False, True), (11, True, False), (40, False, False), (3, False, True)):
        y = rng.integers(0, 256, (h, w), dtype=np.uint8)
        y = (y // 32 * 32 + rng.integers(0, 8, (h, w))).astype(np.uint8)   # blocky content
        u = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
        v = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
        mask = rng.integers(0, 2, (w // 16) * (h // 16)).astype(np.uint8) if masked else None
        gy, gu, gv = stages.deblock(y, u, v, level, intra, mask)
        oy, ou, ov = y.copy(), u.copy(), v.copy()
        O.ro_deblock(vp(oy), vp(ou), vp(ov), w, h, level, int(intra), vp(mask) if mask is not None else None)
        assert np.array_equal(gy, oy) and np.array_equal(gu, ou) and np.array_equal(gv, ov), (level, intra)


def test_motion_compensation_against_oracle():
    rng = np.random.default_rng(3)
    w, h = 96, 64
    y = rng.integers(0, 256, (h, w), dtype=np.uint8)
    u = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
    v = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
    nb = (w // 16) * (h // 16)
    for trial in range(4):
        mv = np.stack([rng.integers(-40, 41, nb), rng.integers(-40, 41, nb), rng.integers(0, 2, nb), rng.integers(0, 2, nb)], 1).astype(np.int8)
        gy, gu, gv = stages.build_predicted(y, u, v, mv)
        oy, ou, ov = np.zeros_like(y), np.zeros_like(u), np.zeros_like(v)
        O.ro_build_predicted(vp(y), vp(u), vp(v), w, h, vp(mv), vp(oy), vp(ou), vp(ov))
        assert np.array_equal(gy, oy) and np.array_equal(gu, ou) and np.array_equal(gv, ov), trial


def test_aq_steps_against_oracle():
    rng = np.random.default_rng(4)
    w, h = 352, 288
    y = rng.integers(0, 256, (h, w), dtype=np.uint8)
    y[:32, :48] = 90
    for base in (23.424242, 4.0, 60.0):
        g = stages.aq_steps(y, base)
        o = np.zeros_like(g)
        O.ro_aq_steps(vp(y), w, h, np.float32(base), vp(o))
        assert np.array_equal(g.view(np.uint32), o.view(np.uint32)), base


def test_aq_mean_is_the_sequential_f32_sum():
    """The device computes the mean of the logs in parallel; it must equal the sequential f32 sum
    (a left fold, as numpy's add.accumulate is) divided by n, bit for bit, also on inputs made of
    few distinct values, where halfway cases (ties) are common, and on frames of 4K and above."""
    rng = np.random.default_rng(9)
    cases = []
    for n in (1, 31, 396, 8160, 32400, 130000):
        cases.append(rng.uniform(np.log(4.0), np.log(65025.0), n).astype(np.float32))
        cases.append((np.float32(np.log(4.0)) + rng.integers(0, 64, n).astype(np.float32) / np.float32(8)).astype(np.float32))
        cases.append(np.ldexp(rng.integers(1, 4096, n), -rng.integers(0, 12, n)).astype(np.float32))
    cases.append(np.full(8160, np.float32(np.log(4.0)), np.float32))
    for i, x in enumerate(cases):
        mean, steps = stages.aq_mean(x, 23.424242)
        want = np.add.accumulate(x, dtype=np.float32)[-1] / np.float32(x.size)
        assert np.float32(mean).view(np.uint32) == np.float32(want).view(np.uint32), (i, x.size, mean, want)


def test_device_logf_exp2f_match_libm():
    """The device restatement of glibc's logf/exp2f agrees with the host libm bit for bit over the
    adaptive quantiser's domains (log of variances in [4, 65025], exp2 of offsets in [-2, 2])."""
    libm = ctypes.CDLL("libm.so.6")
    libm.logf.restype = ctypes.c_float
    libm.logf.argtypes = [ctypes.c_float]
    libm.exp2f.restype = ctypes.c_float
    libm.exp2f.argtypes = [ctypes.c_float]
    rng = np.random.default_rng(5)
    x = np.concatenate([np.linspace(4.0, 65025.0, 20000, dtype=np.float32),
                        rng.uniform(4.0, 65025.0, 20000).astype(np.float32),
                        (np.arange(4, 65026, 7, dtype=np.float32) / np.float32(1.0))])
    g = stages.logf_exp2f(x, 0)
    want = np.array([libm.logf(float(v)) for v in x], np.float32)
    assert np.array_equal(g.view(np.uint32), want.view(np.uint32))
    e = np.concatenate([np.linspace(-2.0, 2.0, 20001, dtype=np.float32), rng.uniform(-2.0, 2.0, 20000).astype(np.float32)])
    g = stages.logf_exp2f(e, 1)
    want = np.array([libm.exp2f(float(v)) for v in e], np.float32)
    assert np.array_equal(g.view(np.uint32), want.view(np.uint32))
