"""Pins of the oracle (oracle/riv_oracle.cpp, the C++ restatement of the reference codec).

Each test checks an oracle function against something the reference fixes independently of the
oracle's own code: a known answer from the reference's tests, a worked example from its SPEC, a
closed form, or a second, direct restatement of the reference routine on inputs small enough for
brute force. CPU only; the GPU path is compared with the oracle in test_gpu_parity.py.
"""
import ctypes
import math

import numpy as np
import pytest

import oracle

L = oracle.lib()


def vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def textured(w, h, seed, smooth=True):
    """A 1/f^2 field scaled to [16, 235] (smooth) or white noise."""
    rng = np.random.default_rng(seed)
    if not smooth:
        return rng.integers(0, 256, (h, w), dtype=np.uint8)
    f = np.fft.fft2(rng.standard_normal((h, w)))
    ky = np.fft.fftfreq(h)[:, None]
    kx = np.fft.fftfreq(w)[None, :]
    k2 = kx * kx + ky * ky
    k2[0, 0] = 1.0
    x = np.real(np.fft.ifft2(f / k2))
    x = (x - x.min()) / max(1e-9, np.ptp(x))
    return np.clip(16 + 219 * x, 0, 255).astype(np.uint8)


# ---------------------------------------------------------------------------------------------
# DCT (reitero_dct/tests/roundtrip.rs)
# ---------------------------------------------------------------------------------------------
KAT_8X8 = [
    16, 11, 10, 16, 24, 40, 51, 61, 12, 12, 14, 19, 26, 58, 60, 55, 14, 13, 16, 24, 40, 57, 69,
    56, 14, 17, 22, 29, 51, 87, 80, 62, 18, 22, 37, 56, 68, 109, 103, 77, 24, 35, 55, 64, 81,
    104, 113, 92, 49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99,
]


def dct_plane(fn, plane, qs, dz):
    h, w = plane.shape
    n = 8 if fn == "8" else 16
    coeffs = np.zeros((h // n) * (w // n) * n * n, np.int16)
    recon = np.zeros_like(plane)
    q = np.full((h // n) * (w // n), qs, np.float32)
    f = L.ro_dct8_aq_plane if fn == "8" else L.ro_dct16_aq_plane
    f(vp(plane), w, h, vp(q), dz, vp(coeffs), vp(recon))
    return coeffs, recon


def test_dct8_known_answer():
    """roundtrip.rs:366-400: DC 461, [1] -169, [8] -195; roundtrip error <= 2."""
    plane = np.array(KAT_8X8, np.int16).reshape(8, 8)
    c, r = dct_plane("8", plane, 1.0, 0.5)
    assert (c[0], c[1], c[8]) == (461, -169, -195)
    assert np.abs(r.astype(int) - plane).max() <= 2


def test_dct8_roundtrip_bound():
    """roundtrip.rs:48-92: random residuals in [-255, 255], step 1.0 -> every sample within 2."""
    rng = np.random.default_rng(3)
    plane = rng.integers(-255, 256, (64, 64)).astype(np.int16)
    _, r = dct_plane("8", plane, 1.0, 0.5)
    assert np.abs(r.astype(int) - plane).max() <= 2


def test_dct16_roundtrip_bound():
    """roundtrip.rs:95-147: random residuals, step 2.0 -> every sample within 4."""
    rng = np.random.default_rng(4)
    plane = rng.integers(-255, 256, (64, 64)).astype(np.int16)
    _, r = dct_plane("16", plane, 2.0, 0.5)
    assert np.abs(r.astype(int) - plane).max() <= 4


def test_dct_bases_closed_form():
    """The tables of dct16.rs:7-74 / dct8.rs:6-25 equal round(2^18 c_k cos((2n+1) k pi / 2N))."""
    b16 = np.zeros((16, 16), np.int32)
    b8 = np.zeros((8, 8), np.int32)
    L.ro_basis(vp(b16), vp(b8))
    for N, b in ((16, b16), (8, b8)):
        ref = np.zeros((N, N), np.int64)
        for n in range(N):
            for k in range(N):
                ck = math.sqrt(1.0 / N) if k == 0 else math.sqrt(2.0 / N)
                ref[n, k] = round(2 ** 18 * ck * math.cos((2 * n + 1) * k * math.pi / (2 * N)))
        assert np.array_equal(b.astype(np.int64), ref), N


def test_dead_zone_and_rounding():
    """residual.rs quantiser: |c/q| < dz -> 0, else round half away from zero. A DC-only block
    with value v has DC coefficient 16 v (orthonormal 16x16 basis, 2^18 scale, >> 18 twice)."""
    for v, qs, dz, want in ((1, 24.0, 0.75, 0), (2, 24.0, 0.75, 1), (3, 24.0, 0.75, 2), (-3, 24.0, 0.75, -2),
                            (1, 16.0, 0.5, 1), (-1, 32.0, 0.5, -1), (1, 33.0, 0.5, 0), (5, 32.0, 0.5, 3)):
        plane = np.full((16, 16), v, np.int16)
        c, _ = dct_plane("16", plane, qs, dz)
        assert c[0] == want, (v, qs, dz, c[0])


# ---------------------------------------------------------------------------------------------
# Deblocking (reitero_video_common/src/deblock.rs:204-289)
# ---------------------------------------------------------------------------------------------
def flat_frame(w, h, left, right):
    y = np.ascontiguousarray(np.broadcast_to(np.where(np.arange(w) < w // 2, left, right).astype(np.uint8), (h, w)))
    u = np.full((h // 2, w // 2), 128, np.uint8)
    v = np.full((h // 2, w // 2), 128, np.uint8)
    return y, u, v


def deblock(frame, level, intra, mask=None):
    y, u, v = frame
    h, w = y.shape
    m = None if mask is None else np.array(mask, np.uint8)
    L.ro_deblock(vp(y), vp(u), vp(v), w, h, level, int(intra), vp(m) if m is not None else None)
    return y.reshape(-1)


def test_deblock_small_step_is_smoothed_wide():
    y = deblock(flat_frame(32, 32, 100, 106), 20, False)
    assert y[15] > 100 and y[16] < 106
    assert 100 <= y[14] <= y[15] and y[16] <= y[17] <= 106
    assert y[8] == 100 and y[24] == 106


def test_deblock_large_step_is_preserved():
    y = deblock(flat_frame(32, 32, 40, 200), 20, False)
    assert (y[15], y[16]) == (40, 200)


def test_deblock_rough_interior_is_not_filtered():
    f = flat_frame(32, 32, 100, 106)
    f[0][:, 13] = 160
    y = deblock(f, 8, False)
    assert (y[15], y[16]) == (100, 106)


def test_deblock_masks():
    y = deblock(flat_frame(32, 32, 100, 106), 20, False, [0, 0, 0, 0])
    assert (y[15], y[16]) == (100, 106)
    y = deblock(flat_frame(32, 32, 100, 106), 20, False, [0, 1, 0, 1])
    assert y[15] > 100 and y[16] < 106


def test_deblock_zero_level_is_noop():
    f = flat_frame(32, 32, 100, 106)
    before = f[0].copy()
    deblock(f, 0, False)
    assert np.array_equal(f[0], before)


# ---------------------------------------------------------------------------------------------
# BinProb and rANS (SPEC.md 13.1-13.2, reitero_video_common/src/rans.rs)
# ---------------------------------------------------------------------------------------------
def test_binprob_p0_closed_form():
    """p_zero = clamp(floor(z * 256 / (z + o)), 1, 255) for every reachable state."""
    for z in range(1, 256):
        for o in range(1, 256):
            want = min(255, max(1, (z * 256) // (z + o)))
            assert L.ro_binprob_p0((z << 8) | o) == want, (z, o)


@pytest.mark.parametrize("state,bit,after", [
    ((1, 1), 0, (2, 1)), ((1, 1), 1, (1, 2)), ((254, 7), 0, (255, 7)),
    ((255, 2), 0, (129, 1)), ((255, 100), 0, (129, 50)), ((255, 254), 0, (129, 127)),
    ((2, 255), 1, (1, 129)), ((255, 1), 0, (255, 1)), ((1, 255), 1, (1, 255)),
])
def test_binprob_update_examples(state, bit, after):
    z, o = state
    p = L.ro_binprob_observe((z << 8) | o, bit)
    assert ((p >> 8) & 255, p & 255) == after


def test_rans_roundtrip_across_batches():
    """Adaptive contexts and bypass bits, more than two 16386-symbol batches."""
    rng = np.random.default_rng(11)
    n = 40000
    ctx = rng.integers(-1, 8, n).astype(np.int8)
    p1 = np.where(ctx < 0, 0.5, (ctx + 1) / 10.0)
    bits = (rng.random(n) < p1).astype(np.uint8)
    cap = n
    out = np.zeros(cap, np.uint8)
    size = L.ro_rans_encode(vp(bits), vp(ctx), n, vp(out), cap)
    assert size <= cap
    back = np.zeros(n, np.uint8)
    assert L.ro_rans_decode(vp(out), size, vp(ctx), n, vp(back)) == 0
    assert np.array_equal(back, bits)
    # skewed contexts compress: well under one byte per 8 symbols
    assert size < n // 8


# ---------------------------------------------------------------------------------------------
# Derived constants (rdo.rs:75-86, SPEC 5.3.1 quant_step, SPEC 14.1 deblock level)
# ---------------------------------------------------------------------------------------------
def test_quality_constants():
    assert L.ro_quant_step(80) == pytest.approx(23.424242, abs=1e-5)
    assert L.ro_quant_step(85) == pytest.approx(17.818182, abs=1e-5)
    assert L.ro_deblock_level(L.ro_quant_step(80)) == 14
    assert L.ro_deblock_level(L.ro_quant_step(85)) == 11
    # quant_step_from_quality (residual.rs:13-23): linear between the q100 and q1 steps
    s100, s1 = L.ro_quant_step(100), L.ro_quant_step(1)
    for q in (2, 37, 80, 99):
        assert L.ro_quant_step(q) == np.float32(s100 + np.float32((np.float32(100.0) - np.float32(q)) / np.float32(99.0)) * np.float32(s1 - s100))
    assert L.ro_lambda(80, 0.49) == pytest.approx(0.49, abs=1e-12)   # qp = round(20 * 0.6) = 12


def test_mv_cost_special_cases():
    """estimate_mv_cost (motion/scalar.rs:453-544) at lambda 0.49: zero 0, predictor 1,
    NEW (0.98 * (6 + b(dx) + b(dy))) as i64 with b(0) = 1, b(d) = 2 log2 d + 2."""
    lam = 0.49
    assert L.ro_mv_cost(0, 0, 5, 5, 7, 7, 9, 9, lam) == 0
    assert L.ro_mv_cost(5, 5, 5, 5, 7, 7, 9, 9, lam) == 1
    assert L.ro_mv_cost(9, 9, 5, 5, 7, 7, 9, 9, lam) == 1
    assert L.ro_mv_cost(1, 2, 0, 0, 0, 0, 0, 0, 0.0) == 0
    # (3, 0) vs nearest (0,0): bits 2*log2(3)+2 = 5.17 and 1 -> 0.98 * 12.17 = 11.9 -> 11
    assert L.ro_mv_cost(3, 0, 0, 0, 100, 100, 100, 100, lam) == 11
    # (4, 4) vs nearest (0,0): 6 + 6 + 6 = 18 -> 17.64 -> 17
    assert L.ro_mv_cost(4, 4, 0, 0, 100, 100, 100, 100, lam) == 17


# ---------------------------------------------------------------------------------------------
# SAD kernels (encode/motion/satd_scalar.rs) against direct restatements
# ---------------------------------------------------------------------------------------------
def sad_int_ref(prev, cur, x0, y0, dx, dy):
    """satd_scalar.rs:4-99: sixteen 4x4 sub-blocks; each reference origin clamped as a unit."""
    h, w = prev.shape
    s = 0
    for by in range(4):
        for bx in range(4):
            sx, sy = x0 + 4 * bx, y0 + 4 * by
            rx = min(max(sx + dx, 0), w - 4)
            ry = min(max(sy + dy, 0), h - 4)
            s += int(np.abs(cur[sy:sy + 4, sx:sx + 4].astype(int) - prev[ry:ry + 4, rx:rx + 4].astype(int)).sum())
    return s


def sample_hp_fixed(prev, x_hp, y_hp):
    """sample_luma_halfpel (satd_scalar.rs:113-166) on the x256 fixed-point samples, >> 8."""
    h, w = prev.shape
    x_hp = min(max(x_hp, 0), 2 * (w - 1))
    y_hp = min(max(y_hp, 0), 2 * (h - 1))
    x0, y0 = x_hp // 2, y_hp // 2
    x1, y1 = min(x0 + 1, w - 1), min(y0 + 1, h - 1)
    a, b, d, e = (int(prev[y0, x0]) << 8, int(prev[y0, x1]) << 8, int(prev[y1, x0]) << 8, int(prev[y1, x1]) << 8)
    xo, yo = x_hp & 1, y_hp & 1
    if xo and not yo:
        f = (a + b + 1) // 2
    elif yo and not xo:
        f = (a + d + 1) // 2
    elif xo and yo:
        f = (a + b + d + e + 2) // 4
    else:
        f = a
    return f >> 8


def sad_hp_ref(prev, cur, x0, y0, dxh, dyh):
    s = 0
    for yy in range(16):
        for xx in range(16):
            cx, cy = x0 + xx, y0 + yy
            s += abs(int(cur[cy, cx]) - sample_hp_fixed(prev, 2 * cx + dxh, 2 * cy + dyh))
    return s


def test_sad_int_matches_restatement():
    rng = np.random.default_rng(5)
    prev = rng.integers(0, 256, (48, 64), dtype=np.uint8)
    cur = rng.integers(0, 256, (48, 64), dtype=np.uint8)
    for (x0, y0) in ((0, 0), (48, 32), (16, 16), (48, 0)):
        for dx, dy in ((0, 0), (-6, -6), (6, 5), (-3, 6), (5, -6), (1, -1)):
            assert L.ro_sad_int(vp(prev), vp(cur), 64, 48, x0, y0, dx, dy) == sad_int_ref(prev, cur, x0, y0, dx, dy)


def test_sad_hp_matches_restatement():
    rng = np.random.default_rng(6)
    prev = rng.integers(0, 256, (32, 48), dtype=np.uint8)
    cur = rng.integers(0, 256, (32, 48), dtype=np.uint8)
    for (x0, y0) in ((0, 0), (32, 16), (16, 0)):
        for dxh, dyh in ((0, 0), (-1, -1), (1, 0), (0, 1), (3, -5), (-7, 7), (13, 2)):
            assert L.ro_sad_hp(vp(prev), vp(cur), 48, 32, x0, y0, dxh, dyh) == sad_hp_ref(prev, cur, x0, y0, dxh, dyh)


# ---------------------------------------------------------------------------------------------
# Motion search: a direct restatement of the serial reference walk on brute-force SADs
# ---------------------------------------------------------------------------------------------
HEX = [(2, 0), (1, 2), (-1, 2), (-2, 0), (-1, -2), (1, -2)]
CARD = [(1, 0), (-1, 0), (0, 1), (0, -1)]


def mv_cost_ref(dx, dy, nearest, near, median, lam):
    """estimate_mv_cost, motion/scalar.rs:453-544 (Python floats are IEEE doubles)."""
    if lam == 0.0:
        return 0
    best = None

    def take(c):
        nonlocal best
        best = c if best is None or c < best else best

    for p in (nearest, near, median):
        if (dx, dy) == p:
            take(int(lam * 2.0 * 1.5))
    if (dx, dy) == (0, 0):
        take(int(lam * 2.0 * 0.5))
    for p in (nearest, near):
        ax, ay = abs(dx - p[0]), abs(dy - p[1])
        bx = 1.0 if ax == 0 else 2.0 * math.log2(ax) + 2.0
        by = 1.0 if ay == 0 else 2.0 * math.log2(ay) + 2.0
        take(int(lam * 2.0 * (6.0 + bx + by)))
    return best


def median3(a, b, c):
    return sorted((a, b, c))[1]


def canon(hx, hy):
    """Truncating /2 and %2, then as_canonicalized (motion_vector.rs:103-116)."""
    out = []
    for h in (max(-255, min(255, hx)), max(-255, min(255, hy))):
        d = int(h / 2)
        f = h - 2 * d
        d = max(-128, min(127, d))
        if d == -128 and f == -1:
            f = 0
        if f == -1:
            f, d = 1, max(-128, d - 1)
        out += [d, f]
    return out[0], out[2], out[1], out[3]


def hex_search_ref(prev, cur, sr, prev_mvs, lam):
    h, w = prev.shape
    bw, bh = w // 16, h // 16
    mvs = [(0, 0, 0, 0)] * (bw * bh)
    sads = [0] * (bw * bh)
    for by in range(bh):
        for bx in range(bw):
            idx = by * bw + bx
            preds = []
            if bx > 0:
                preds.append(mvs[idx - 1][:2])
            if by > 0:
                preds.append(mvs[idx - bw][:2])
                if bx + 1 < bw:
                    preds.append(mvs[idx - bw + 1][:2])
                if bx > 0:
                    preds.append(mvs[idx - bw - 1][:2])
            if prev_mvs is not None:
                preds.append(prev_mvs[idx][:2])
            uniq = []
            for p in preds:
                if p not in uniq:
                    uniq.append(p)
            while len(uniq) < 2:
                uniq.append((0, 0))
            nearest, near = uniq[0], uniq[1]
            left = mvs[idx - 1][:2] if bx > 0 else (0, 0)
            top = mvs[idx - bw][:2] if by > 0 else (0, 0)
            tr = mvs[idx - bw + 1][:2] if by > 0 and bx + 1 < bw else (mvs[idx - bw - 1][:2] if by > 0 and bx > 0 else (0, 0))
            median = (median3(left[0], top[0], tr[0]), median3(left[1], top[1], tr[1]))
            cands = []
            for p in [(0, 0)] + preds:
                if abs(p[0]) <= sr and abs(p[1]) <= sr and p not in cands:
                    cands.append(p)
            x0, y0 = bx * 16, by * 16
            sad = lambda d: sad_int_ref(prev, cur, x0, y0, d[0], d[1])
            cost = lambda d: mv_cost_ref(d[0], d[1], nearest, near, median, lam)
            best, bc = (0, 0), sad((0, 0)) + cost((0, 0))
            for c in cands:
                if c == (0, 0):
                    continue
                t = sad(c) + cost(c)
                if t < bc:
                    best, bc = c, t
            while True:
                improved = False
                cx, cy = best
                for ox, oy in HEX:
                    d = (cx + ox, cy + oy)
                    if abs(d[0]) > sr or abs(d[1]) > sr:
                        continue
                    t = sad(d) + cost(d)
                    if t < bc:
                        best, bc, improved = d, t, True
                if not improved:
                    break
            cx, cy = best
            for ox, oy in CARD:
                d = (cx + ox, cy + oy)
                if abs(d[0]) > sr or abs(d[1]) > sr:
                    continue
                t = sad(d) + cost(d)
                if t < bc:
                    best, bc = d, t
            bhx, bhy = 2 * best[0], 2 * best[1]
            tot = sad_hp_ref(prev, cur, x0, y0, bhx, bhy)
            bsad, hx, hy = tot, bhx, bhy
            pen = int(lam * 2.0 * 1.5)
            for sy in (-1, 0, 1):
                for sx in (-1, 0, 1):
                    if sx == 0 and sy == 0:
                        continue
                    s = sad_hp_ref(prev, cur, x0, y0, bhx + sx, bhy + sy)
                    if s + pen < tot:
                        tot, bsad, hx, hy = s + pen, s, bhx + sx, bhy + sy
            mvs[idx] = canon(hx, hy)
            sads[idx] = bsad
    return mvs, sads


def run_hex(prev, cur, sr, prev_mvs, lam):
    h, w = prev.shape
    nb = (w // 16) * (h // 16)
    mv = np.zeros(nb * 4, np.int8)
    sad = np.zeros(nb, np.int64)
    pm = None if prev_mvs is None else np.array(prev_mvs, np.int8).reshape(-1)
    L.ro_hex_search(vp(prev), vp(cur), w, h, sr, vp(pm) if pm is not None else None, 0, 0, lam, 0, vp(mv), vp(sad))
    return [tuple(int(v) for v in mv[4 * i:4 * i + 4]) for i in range(nb)], [int(s) for s in sad]


@pytest.mark.parametrize("seed,sr,lam,temporal", [(1, 4, 0.49, False), (2, 6, 2.0, True), (3, 5, 0.0, True)])
def test_hex_search_matches_restatement(seed, sr, lam, temporal):
    cur = textured(64, 48, seed)
    shift = (seed % 3 + 1, -(seed % 2) - 1)
    prev = np.roll(np.roll(cur, shift[1], 0), shift[0], 1)
    prev = np.clip(prev.astype(int) + np.random.default_rng(seed).integers(-3, 4, prev.shape), 0, 255).astype(np.uint8)
    pm = None
    if temporal:
        rng = np.random.default_rng(seed + 10)
        pm = [(int(a), int(b), int(c), int(d)) for a, b, c, d in zip(rng.integers(-3, 4, 12), rng.integers(-3, 4, 12),
                                                                     rng.integers(0, 2, 12), rng.integers(0, 2, 12))]
    got = run_hex(prev, cur, sr, pm, lam)
    want = hex_search_ref(prev, cur, sr, pm, lam)
    assert got == want


def test_hex_search_identical_frames():
    cur = textured(64, 48, 9, smooth=False)
    mvs, sads = run_hex(cur, cur.copy(), 12, None, 0.49)
    assert all(m == (0, 0, 0, 0) for m in mvs) and not any(sads)


# ---------------------------------------------------------------------------------------------
# Motion compensation (fast_motion.rs:103-352; motion.rs:28-117)
# ---------------------------------------------------------------------------------------------
def sample_round(p, x_hp, y_hp):
    h, w = p.shape
    x_hp = min(max(x_hp, 0), 2 * (w - 1))
    y_hp = min(max(y_hp, 0), 2 * (h - 1))
    x0, y0 = x_hp // 2, y_hp // 2
    x1, y1 = min(x0 + 1, w - 1), min(y0 + 1, h - 1)
    a, b, d, e = int(p[y0, x0]), int(p[y0, x1]), int(p[y1, x0]), int(p[y1, x1])
    xo, yo = x_hp & 1, y_hp & 1
    if xo and not yo:
        return (a + b + 1) // 2
    if yo and not xo:
        return (a + d + 1) // 2
    if xo and yo:
        return (a + b + d + e + 2) // 4
    return a


def mc_ref(Y, U, V, mvs):
    h, w = Y.shape
    bw = w // 16
    pY, pU, pV = np.zeros_like(Y), np.zeros_like(U), np.zeros_like(V)
    for i, (dx, dy, sx, sy) in enumerate(mvs):
        bx, by = i % bw, i // bw
        x0, y0 = 16 * bx, 16 * by
        mx, my = 2 * dx + sx, 2 * dy + sy
        if mx % 2 == 0 and my % 2 == 0 and 0 <= x0 + mx // 2 and 0 <= y0 + my // 2 \
                and x0 + mx // 2 + 15 < w and y0 + my // 2 + 15 < h:
            ax, ay = x0 + mx // 2, y0 + my // 2
            pY[y0:y0 + 16, x0:x0 + 16] = Y[ay:ay + 16, ax:ax + 16]
            pU[y0 // 2:y0 // 2 + 8, x0 // 2:x0 // 2 + 8] = U[ay // 2:ay // 2 + 8, ax // 2:ax // 2 + 8]
            pV[y0 // 2:y0 // 2 + 8, x0 // 2:x0 // 2 + 8] = V[ay // 2:ay // 2 + 8, ax // 2:ax // 2 + 8]
            continue
        for yy in range(16):
            for xx in range(16):
                pY[y0 + yy, x0 + xx] = sample_round(Y, 2 * (x0 + xx) + mx, 2 * (y0 + yy) + my)
        for cy in range(8):
            for cx in range(8):
                X, Yc = x0 // 2 + cx, y0 // 2 + cy
                pU[Yc, X] = sample_round(U, (4 * X + mx) >> 1, (4 * Yc + my) >> 1)
                pV[Yc, X] = sample_round(V, (4 * X + mx) >> 1, (4 * Yc + my) >> 1)
    return pY, pU, pV


def test_mc_matches_restatement():
    rng = np.random.default_rng(8)
    w, h = 64, 48
    Y = rng.integers(0, 256, (h, w), dtype=np.uint8)
    U = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
    V = rng.integers(0, 256, (h // 2, w // 2), dtype=np.uint8)
    mvs = [(3, -2, 0, 0), (-1, 0, 0, 0), (5, 5, 1, 0), (-4, 1, 0, 1), (0, 0, 1, 1), (1, 1, 0, 0),
           (-20, -20, 1, 1), (7, -3, 0, 0), (2, 2, 1, 1), (-1, -1, 1, 0), (30, 4, 0, 0), (0, -6, 0, 1)]
    mv = np.array(mvs, np.int8).reshape(-1)
    oy, ou, ov = np.zeros_like(Y), np.zeros_like(U), np.zeros_like(V)
    L.ro_build_predicted(vp(Y), vp(U), vp(V), w, h, vp(mv), vp(oy), vp(ou), vp(ov))
    ry, ru, rv = mc_ref(Y, U, V, mvs)
    assert np.array_equal(oy, ry) and np.array_equal(ou, ru) and np.array_equal(ov, rv)


# ---------------------------------------------------------------------------------------------
# Adaptive quantisation (residual.rs:116-167) with the host libm
# ---------------------------------------------------------------------------------------------
def test_aq_steps_with_libm():
    libm = ctypes.CDLL("libm.so.6")
    libm.logf.restype = ctypes.c_float
    libm.logf.argtypes = [ctypes.c_float]
    libm.exp2f.restype = ctypes.c_float
    libm.exp2f.argtypes = [ctypes.c_float]
    w, h = 64, 48
    y = textured(w, h, 12)
    y[:16, :16] = 77   # a flat block: variance floors at 4
    base = np.float32(23.424242)
    got = np.zeros((w // 16) * (h // 16), np.float32)
    L.ro_aq_steps(vp(y), w, h, base, vp(got))
    lns = []
    for by in range(h // 16):
        for bx in range(w // 16):
            blk = y[16 * by:16 * by + 16, 16 * bx:16 * bx + 16].astype(np.int64)
            s, s2 = int(blk.sum()), int((blk * blk).sum())
            var = np.float32(np.float32(s2 * 256 - s * s) / np.float32(65536.0))
            lns.append(np.float32(libm.logf(max(var, np.float32(4.0)))))
    acc = np.float32(0.0)
    for v in lns:
        acc = np.float32(acc + v)
    avg = np.float32(acc / np.float32(len(lns)))
    want = [np.float32(base * np.float32(libm.exp2f(np.float32(np.float32(np.float32(0.8) * np.float32(v - avg)) / np.float32(6.0)))))
            for v in lns]
    assert np.array_equal(got, np.array(want, np.float32))


# ---------------------------------------------------------------------------------------------
# Whole loop: the encoder's reconstruction is what the decoder outputs (residual.rs:1306-1309)
# ---------------------------------------------------------------------------------------------
def test_encoder_recon_equals_decoder_output():
    from riv2_amd import synth
    frames = synth.Sequence(synth.cif_params(seed=21, frames=6)).frames(0, 6)
    enc = oracle.OracleEncoder(oracle.config(352, 288))
    recons = []
    for i, f in enumerate(frames):
        enc.encode(f, i * 33)
        recons.append(enc.last_yuv(2))
    data = enc.finish()
    dec = oracle.OracleDecoder(data)
    for i in range(len(frames)):
        dec.decode()
        for a, b in zip(dec.last_yuv(), recons[i]):
            assert np.array_equal(a, b), i


# ---------------------------------------------------------------------------------------------
# The project's own BT.709 limited-range conversion (parity with the reference's `yuv` crate is
# unpinned): closed points of the standard
# ---------------------------------------------------------------------------------------------
def test_bt709_limited_closed_points():
    w, h = 16, 16
    for rgb, want in (((0, 0, 0), (16, 128, 128)), ((255, 255, 255), (235, 128, 128)),
                      ((255, 0, 0), (63, 102, 240)), ((0, 255, 0), (173, 42, 26)), ((0, 0, 255), (32, 240, 118))):
        img = np.tile(np.array(rgb, np.uint8), (h, w, 1))
        y = np.zeros((h, w), np.uint8)
        u = np.zeros((h // 2, w // 2), np.uint8)
        v = np.zeros((h // 2, w // 2), np.uint8)
        L.ro_rgb_to_yuv420(vp(img), w, h, vp(y), vp(u), vp(v))
        assert (int(y[0, 0]), int(u[0, 0]), int(v[0, 0])) == want, rgb
