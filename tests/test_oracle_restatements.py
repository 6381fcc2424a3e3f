"""Independent restatements that pin the oracle's remaining functions, each written from the
reference's text (its SPEC and source, cited per function) in a different form from the oracle's
C++ code, and checked against the oracle on inputs chosen so that a plausible misreading (a term
or context dropped, an order or index swapped, a rounding or sign changed) fails:

* the entropy coders' symbol layouts: BinProb, the dual-state rANS writer with its 16386-symbol
  batches, the residual coefficient coder and the MV coder (with its candidate history), as bytes;
* the mode decision (Zero / Nearest / Near / TopRight / TopLeft / Temporal / NEW, the NEW base by
  estimated bits, the reconstructed vector) and the NEW-delta bias search, brute force;
* RdoContext::decide against its solved closed form (the skip/code crossover SAD);
* the deblocking filter against a restatement of the VP8 normal macroblock-edge filter;
* IDCT16 and IDCT8 against closed forms of DC-only and single-coefficient inputs;
* RGB<->YUV420 against real-valued BT.709 limited range within one code value.

CPU only."""
import ctypes
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle

L = oracle.lib()
P = ctypes.c_void_p
L.ro_mode_decision.restype = None
L.ro_mode_decision.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P, P]
L.ro_encode_residual_blocks.restype = ctypes.c_size_t
L.ro_encode_residual_blocks.argtypes = [P, P, ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t]
L.ro_mv_encode_sequence.restype = ctypes.c_size_t
L.ro_mv_encode_sequence.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, ctypes.c_size_t, P]
L.ro_idct16.restype = None
L.ro_idct16.argtypes = [P, P]
L.ro_idct8.restype = None
L.ro_idct8.argtypes = [P, P, P]


def vp(a):
    return a.ctypes.data_as(ctypes.c_void_p)


# =============================================================================================
# Entropy coding: BinProb and rANS (SPEC.md 13.1-13.2; reitero_video_common/src/rans.rs:25-298)
# =============================================================================================
class Prob:
    """A pair of counters (zeros, ones), each in [1, 255], starting at (1, 1)."""

    def __init__(self):
        self.z, self.o = 1, 1

    def p_zero(self):
        return max(1, min(255, (self.z * 256) // (self.z + self.o)))

    def observe(self, bit):
        seen, other = (self.o, self.z) if bit else (self.z, self.o)
        if seen < 255:
            seen += 1
        elif other != 1:
            # rans.rs:73-84 halves the other counter rounding up ((other + 1) >> 1); SPEC 13.2's
            # "other >> 1" differs for odd counts, and the code is authoritative
            seen, other = 129, (other + 1) >> 1
        if bit:
            self.o, self.z = seen, other
        else:
            self.z, self.o = seen, other


class SymbolWriter:
    """Collects (start, freq) symbols in coding order; bytes() runs the rANS batches."""

    BATCH = 16386

    def __init__(self):
        self.syms = []

    def bit(self, b, ctx):
        p = ctx.p_zero()
        ctx.observe(b)
        self.syms.append((p, 256 - p) if b else (0, p))

    def bypass(self, b):
        self.syms.append((128, 128) if b else (0, 128))

    @staticmethod
    def _batch(batch):
        """One batch from fresh states: symbols last to first, the first of an odd batch alone in
        state 0 before the states swap, then pairs (state 0 the earlier of the two in reverse);
        words go to the front of the output, the final states last."""
        words = []   # emitted in order; the output is their reverse after the states
        st = [1 << 16, 1 << 16]

        def put(k, sym):
            start, freq = sym
            x = st[k]
            if x >= freq << 24:
                words.append(x & 0xffff)
                x >>= 16
            st[k] = ((x // freq) << 8) + (x % freq) + start

        rev = batch[::-1]
        i = 0
        if len(rev) % 2:
            put(0, rev[0])
            i = 1
            st.reverse()
        while i < len(rev):
            put(0, rev[i])
            put(1, rev[i + 1])
            i += 2
        out = [st[1] & 0xffff, st[1] >> 16, st[0] & 0xffff, st[0] >> 16] + words[::-1]
        return b"".join(int(w).to_bytes(2, "little") for w in out)

    def bytes(self):
        n = len(self.syms)
        if n == 0:
            return self._batch([])
        return b"".join(self._batch(self.syms[b:b + self.BATCH]) for b in range(0, n, self.BATCH))


def test_binprob_odd_counter_halving():
    """The overflow rule on an odd opposite count: (255, 99) + 0 -> (129, 50) (the code's rounding)."""
    p = L.ro_binprob_observe((255 << 8) | 99, 0)
    assert ((p >> 8) & 255, p & 255) == (129, 50)
    q = Prob(); q.z, q.o = 255, 99
    q.observe(0)
    assert (q.z, q.o) == (129, 50)


@pytest.mark.parametrize("n", [0, 1, 2, 7, 16385, 16386, 16387, 2 * 16386 + 5])
def test_rans_writer_batches(n):
    """Adaptive and bypass symbols across the batch boundary, bytes equal to the oracle's writer."""
    rng = np.random.default_rng(100 + n)
    ctx_ids = rng.integers(-1, 6, n).astype(np.int8)
    bits = (rng.random(n) < np.where(ctx_ids < 0, 0.5, (ctx_ids + 1) / 8.0)).astype(np.uint8)
    out = np.zeros(4 * n + 64, np.uint8)
    size = L.ro_rans_encode(vp(bits), vp(ctx_ids), n, vp(out), out.size)
    w = SymbolWriter()
    ctxs = [Prob() for _ in range(6)]
    for b, c in zip(bits.tolist(), ctx_ids.tolist()):
        if c < 0:
            w.bypass(bool(b))
        else:
            w.bit(bool(b), ctxs[c])
    assert out[:size].tobytes() == w.bytes()


# =============================================================================================
# Residual coefficient coder (SPEC.md 5.4; reitero_residual/src/rans.rs:189-327)
# =============================================================================================
def class_code(w, v, ctxs):
    """Class k of v (0, 1, then [2^(k-1), 2^k - 1]) as truncated unary over ctxs, then the
    offset within the class MSB first as bypass bits."""
    k = 0 if v == 0 else v.bit_length()
    assert k < len(ctxs)
    for i in range(k):
        w.bit(True, ctxs[i])
    if k < len(ctxs) - 1:
        w.bit(False, ctxs[k])
    for b in range(k - 2, -1, -1):
        w.bypass(bool(((v - (1 << (k - 1))) >> b) & 1))


class ResidualModel:
    def __init__(self, dc_pred):
        self.dc_pred = dc_pred
        self.sig = {"Y": [Prob() for _ in range(256)], "C": [Prob() for _ in range(64)]}
        self.eob = {"Y": [Prob() for _ in range(9)], "C": [Prob() for _ in range(7)]}
        self.mag = {"Y": [[Prob() for _ in range(13)] for _ in range(4)], "C": [[Prob() for _ in range(13)] for _ in range(3)]}
        self.prev_dc = [0, 0, 0]

    @staticmethod
    def band(kind, i):
        if i == 0:
            return 0
        if kind == "Y":
            return 1 if i < 16 else (2 if i < 64 else 3)
        return 1 if i < 8 else 2

    def plane(self, w, q, pidx):
        kind = "Y" if pidx == 0 else "C"
        vals = [int(x) for x in q]
        if self.dc_pred:   # the DC as a delta to the previous block's DC of the plane; position 0 always coded
            vals[0], self.prev_dc[pidx] = vals[0] - self.prev_dc[pidx], vals[0]
            last = max([i for i in range(1, len(vals)) if vals[i]], default=0)
        else:
            last = max([i for i in range(len(vals)) if vals[i]], default=0)
        class_code(w, last, self.eob[kind])
        for i in range(last + 1):
            v = vals[i]
            w.bit(v != 0, self.sig[kind][i])
            if v:
                w.bypass(v < 0)
                class_code(w, abs(v) - 1, self.mag[kind][self.band(kind, i)])


def residual_bytes(coeffs, coded, dc_pred):
    if not coded.any():
        return b""
    w, m = SymbolWriter(), ResidualModel(dc_pred)
    for b in range(len(coded)):
        if coded[b]:
            m.plane(w, coeffs[b, :256], 0)
            m.plane(w, coeffs[b, 256:320], 1)
            m.plane(w, coeffs[b, 320:], 2)
    return w.bytes()


def random_blocks(rng, nb, style):
    c = np.zeros((nb, 384), np.int16)
    for b in range(nb):
        kind = style if style != "mixed" else ["sparse", "dense", "edge", "empty"][b % 4]
        if kind == "sparse":
            idx = rng.choice(384, size=rng.integers(1, 8), replace=False)
            c[b, idx] = rng.choice([-3, -2, -1, 1, 2, 3, 9, -17], size=idx.size)
        elif kind == "dense":
            m = rng.random(384) < 0.4
            c[b, m] = (rng.integers(1, 40, m.sum()) * rng.choice([-1, 1], m.sum())).astype(np.int16)
        elif kind == "edge":   # last positions of every plane, class limits of EOB and magnitude
            c[b, [255, 319, 383, 0, 256, 320]] = [4096, -4096, 2047, 1, -1, 2048]
    return c


@pytest.mark.parametrize("dc_pred,style,nb", [(False, "mixed", 40), (True, "mixed", 40), (False, "dense", 60),
                                               (True, "sparse", 30), (False, "edge", 5)])
def test_residual_coder_bytes(dc_pred, style, nb):
    rng = np.random.default_rng(nb * 7 + int(dc_pred))
    coeffs = random_blocks(rng, nb, style)
    coded = (rng.random(nb) < 0.8).astype(np.uint8)
    coded[0] = 1
    out = np.zeros(1 << 20, np.uint8)
    size = L.ro_encode_residual_blocks(vp(coeffs), vp(coded), nb, int(dc_pred), vp(out), out.size)
    assert out[:size].tobytes() == residual_bytes(coeffs, coded, dc_pred)


def test_residual_coder_no_coded_block_is_empty():
    coeffs = np.zeros((4, 384), np.int16)
    coded = np.zeros(4, np.uint8)
    assert L.ro_encode_residual_blocks(vp(coeffs), vp(coded), 4, 0, vp(np.zeros(8, np.uint8)), 8) == 0


# =============================================================================================
# MV coder (SPEC.md 4.5; reitero_residual/src/mv_rans/encoder.rs:60-413, mv_predictor.rs:276-308)
# =============================================================================================
ZERO, NEAREST, NEAR, TOPRIGHT, TOPLEFT, NEW, TEMPORAL = range(7)


def mode_context(modes, bw, bx, by):
    score = 0
    for m in ([modes[-1]] if bx > 0 else []) + ([modes[len(modes) - bw]] if by > 0 else []):
        score += 2 if m == ZERO else (0 if m == NEW else 1)
    return min(score, 3)


def mv_class(mag):
    return 0 if mag == 0 else min(mag.bit_length(), 10)


class MvModel:
    def __init__(self):
        c = lambda n: [Prob() for _ in range(n)]
        self.skip = Prob()
        self.mode = {k: c(4) for k in ("new", "nonzero", "near", "temporal", "tr", "tl")}
        self.base = {k: c(4) for k in ("nearest", "near", "tr", "tl")}
        self.cls = {"x": c(4), "y": c(4)}
        self.sign = {"x": Prob(), "y": Prob()}
        self.magb = {"x": c(10), "y": c(10)}
        self.frac = {"x": Prob(), "y": Prob()}
        self.hit = c(3)
        self.idx = c(4)
        self.cands, self.counts, self.total = [], {}, 0

    def component(self, w, axis, d):
        mag = abs(d)
        k = mv_class(mag)
        lo, hi, depth = 0, 10, 0
        while lo < hi:   # binary search over the classes 0..10
            mid = (lo + hi) // 2
            w.bit(k > mid, self.cls[axis][depth])
            lo, hi = (mid + 1, hi) if k > mid else (lo, mid)
            depth += 1
        if k == 0:
            return
        w.bit(d < 0, self.sign[axis])
        off = mag - (1 << (k - 1))
        for b in range(max(0, k - 1)):   # LSB first
            w.bit(bool((off >> b) & 1), self.magb[axis][b])

    def frame(self, w, blocks, bw, bh):
        modes, hits = [], []
        for i, (mode, nbase, dx, dy, sx, sy, skip) in enumerate(blocks):
            bx, by = i % bw, i // bw
            w.bit(bool(skip), self.skip)
            ctx = mode_context(modes, bw, bx, by)
            chain = [("new", mode == NEW)]
            if mode != NEW:
                chain.append(("nonzero", mode != ZERO))
                if mode != ZERO:
                    chain.append(("near", mode == NEAR))
                    if mode != NEAR:
                        chain.append(("temporal", mode == TEMPORAL))
                        if mode != TEMPORAL:
                            chain.append(("tr", mode == TOPRIGHT))
                            if mode != TOPRIGHT:
                                chain.append(("tl", mode == TOPLEFT))
            for name, b in chain:
                w.bit(b, self.mode[name][ctx])
            hit = False
            if mode == NEW:
                for k, name in enumerate(("nearest", "near", "tr", "tl")):
                    w.bit(nbase == k, self.base[name][ctx])
                    if nbase == k:
                        break
                key = (dx, dy, sx, sy)
                via_candidate = False
                if self.cands:
                    hit = key in self.cands
                    lh = hits[i - 1] if bx > 0 else False
                    th = hits[i - bw] if by > 0 else False
                    w.bit(hit, self.hit[int(lh) + int(th)])
                    if hit:
                        index, lo, hi, depth = self.cands.index(key), 0, 16, 0
                        while hi - lo > 1:
                            mid = (lo + hi) // 2
                            w.bit(index >= mid, self.idx[depth])
                            lo, hi = (mid, hi) if index >= mid else (lo, mid)
                            depth += 1
                        via_candidate = True
                if not via_candidate:
                    self.component(w, "x", dx)
                    self.component(w, "y", dy)
                    w.bit(sx != 0, self.frac["x"])
                    w.bit(sy != 0, self.frac["y"])
            hits.append(hit)
            modes.append(mode)
        # history: decay by 7/8, count this frame's NEW keys, keep the frequent ones
        self.counts = {k: v * 7 // 8 for k, v in self.counts.items()}
        self.total = self.total * 7 // 8
        for (mode, _, dx, dy, sx, sy, _) in blocks:
            if mode == NEW:
                self.counts[(dx, dy, sx, sy)] = self.counts.get((dx, dy, sx, sy), 0) + 1
                self.total += 1
        self.counts = {k: v for k, v in self.counts.items() if v > 0}
        if self.total < 100:
            self.cands = []
        else:
            order = sorted(self.counts.items(), key=lambda kv: (-kv[1], kv[0]))
            self.cands = [k for k, v in order if v > self.total // 100][:16]


def random_mv_frames(rng, nframes, bw, bh, pool_size):
    pool = [(int(rng.integers(-6, 7)), int(rng.integers(-6, 7)), int(rng.integers(0, 2)), int(rng.integers(0, 2)))
            for _ in range(pool_size)]
    frames = []
    for _ in range(nframes):
        blocks = []
        for _ in range(bw * bh):
            mode = int(rng.choice([ZERO, NEAREST, NEAR, TOPRIGHT, TOPLEFT, NEW, TEMPORAL], p=[.1, .1, .1, .05, .05, .5, .1]))
            if mode == NEW:
                if rng.random() < 0.7:
                    dx, dy, sx, sy = pool[int(rng.integers(len(pool)))]
                else:
                    dx, dy = (int(v) for v in rng.choice([-128, -100, -37, -1, 0, 1, 64, 127], 2))
                    sx, sy = int(rng.integers(0, 2)), int(rng.integers(0, 2))
                blocks.append((mode, int(rng.integers(0, 5)), dx, dy, sx, sy, int(rng.random() < 0.3)))
            else:
                blocks.append((mode, 0, 0, 0, 0, 0, int(rng.random() < 0.3)))
        frames.append(blocks)
    return frames


@pytest.mark.parametrize("seed,pool", [(1, 3), (2, 12), (3, 40)])
def test_mv_coder_bytes_with_candidate_history(seed, pool):
    """Five P-frames of one stream: contexts and the candidate history carry across frames (at least
    100 NEW blocks per frame, so the history is in use from the second frame on)."""
    bw, bh, nf = 20, 12, 5
    rng = np.random.default_rng(seed)
    frames = random_mv_frames(rng, nf, bw, bh, pool)
    arr = np.array([list(b) for f in frames for b in f], np.int8)
    out = np.zeros(1 << 20, np.uint8)
    sizes = np.zeros(nf, np.uint32)
    L.ro_mv_encode_sequence(vp(arr), nf, bw, bh, vp(out), out.size, vp(sizes))
    model = MvModel()
    pos = 0
    used_candidates = False
    for f in range(nf):
        w = SymbolWriter()
        used_candidates |= bool(model.cands)
        model.frame(w, frames[f], bw, bh)
        got = out[pos:pos + sizes[f]].tobytes()
        pos += int(sizes[f])
        assert got == w.bytes(), f"frame {f}"
    assert used_candidates


# =============================================================================================
# Mode decision and the NEW-delta bias (reitero_encode/src/encoder.rs:128-159, :447-733,
# :984-1048; reitero_residual/src/mv_predictor.rs:59-272)
# =============================================================================================
def neighbours(coded, prev, bw, bx, by):
    """Raw left / top / top-right / top-left of the reconstructed prefix and the temporal vector."""
    i = by * bw + bx
    get = lambda j: tuple(coded[j])
    left = get(i - 1) if bx > 0 else None
    top = get(i - bw) if by > 0 else None
    tr = get(i - bw + 1) if by > 0 and bx + 1 < bw else None
    tl = get(i - bw - 1) if by > 0 and bx > 0 else None
    temporal = tuple(prev[i]) if prev is not None else None
    return left, top, tr, tl, temporal


def predictors(coded, prev, bw, bx, by):
    left, top, tr, tl, temporal = neighbours(coded, prev, bw, bx, by)
    uniq = []
    for c in (left, top, tr, tl, temporal):
        if c is not None and c not in uniq:
            uniq.append(c)
    uniq += [(0, 0, 0, 0)] * 2
    return uniq[0], uniq[1], (temporal if temporal is not None else (0, 0, 0, 0)), tr, tl


def est_bits(d):
    return 4 if d == 0 else 4 + mv_class(abs(d))


def clamp8(v):
    return max(-128, min(127, v))


def decide_modes_ref(search, prev, bw, bh):
    coded, recs = [], []
    for i, mv in enumerate(search):
        bx, by = i % bw, i // bw
        nearest, near, temporal, tr, tl = predictors(coded, prev, bw, bx, by)
        hp = lambda v: (2 * v[0] + v[2], 2 * v[1] + v[3])
        if tuple(mv) == (0, 0, 0, 0):
            mode = ZERO
        elif hp(mv) == hp(nearest):
            mode = NEAREST
        elif hp(mv) == hp(near):
            mode = NEAR
        elif tr is not None and hp(mv) == hp(tr):
            mode = TOPRIGHT
        elif tl is not None and hp(mv) == hp(tl):
            mode = TOPLEFT
        elif hp(mv) == hp(temporal):
            mode = TEMPORAL
        else:
            mode = NEW
        base_idx, raw = 0, (0, 0)
        rec = tuple(mv)
        if mode == NEW:
            bases = [nearest, near, tr, tl, temporal]
            best = None
            for k, b in enumerate(bases):
                if b is None:
                    continue
                cost = [1, 2, 3, 4, 4][k] + est_bits(clamp8(mv[0] - b[0])) + est_bits(clamp8(mv[1] - b[1]))
                if best is None or cost < best[0]:
                    best = (cost, k, b)
            _, base_idx, b = best
            raw = (clamp8(mv[0] - b[0]), clamp8(mv[1] - b[1]))
            rec = (clamp8(b[0] + raw[0]), clamp8(b[1] + raw[1]), mv[2], mv[3])
        elif mode != ZERO:
            rec = {NEAREST: nearest, NEAR: near, TOPRIGHT: tr, TOPLEFT: tl, TEMPORAL: temporal}[mode]
        coded.append(rec)
        recs.append(tuple(mv) + tuple(rec) + (mode, base_idx) + raw)
    return recs


def bias_ref(recs):
    deltas = [(r[10], r[11]) for r in recs if r[8] == NEW]
    if not deltas:
        return 0, 0
    out = []
    for axis in (0, 1):
        best = None
        for b in range(-24, 25):
            adj = [clamp8(d[axis] - b) for d in deltas]
            key = (-sum(a == 0 for a in adj), sum(abs(a) for a in adj))   # most zeros, then the smallest L1, the first bias on ties
            if best is None or key < best[0]:
                best = (key, b)
        out.append(best[1])
    return tuple(out)


def canonical_field(rng, n, pool):
    """Canonical vectors (sub-pel in {0, +1}; -128 only with sub-pel 0) drawn mostly from a small pool."""
    out = []
    for _ in range(n):
        if rng.random() < 0.75:
            out.append(pool[int(rng.integers(len(pool)))])
        else:
            dx, dy = (int(v) for v in rng.integers(-128, 128, 2))
            sx, sy = int(rng.integers(0, 2)) if dx != -128 else 0, int(rng.integers(0, 2)) if dy != -128 else 0
            out.append((dx, dy, sx, sy))
    return out


@pytest.mark.parametrize("seed", range(6))
def test_mode_decision_and_bias_brute_force(seed):
    rng = np.random.default_rng(seed)
    bw, bh = [(7, 5), (1, 6), (9, 1), (11, 8), (3, 3), (16, 9)][seed]
    pool = [(0, 0, 0, 0), (1, 0, 0, 0), (0, -1, 1, 0), (3, 2, 1, 1), (-2, 5, 0, 1), (127, -128, 0, 0)]
    search = canonical_field(rng, bw * bh, pool)
    prev = canonical_field(rng, bw * bh, pool) if seed % 2 else None
    s4 = np.array(search, np.int8)
    p4 = np.array(prev, np.int8) if prev is not None else None
    out = np.zeros((bw * bh, 12), np.int8)
    bias = np.zeros(2, np.int32)
    L.ro_mode_decision(vp(s4), vp(p4) if p4 is not None else None, bw, bh, vp(out), vp(bias))
    want = decide_modes_ref(search, prev, bw, bh)
    for i, r in enumerate(want):
        assert tuple(int(v) for v in out[i]) == r, f"block {i}: {out[i].tolist()} vs {r}"
    assert tuple(int(v) for v in bias) == bias_ref(want)
    if bw * bh >= 30:
        assert len({r[8] for r in want}) >= 4   # several modes exercised


# =============================================================================================
# RdoContext::decide (reitero_encode/src/rdo.rs:23-178): skip iff
# SAD + lambda*skip_bits <= SAD*retention + lambda*(keep_bits + max(0, b0 + b1*SAD)), solved for SAD
# =============================================================================================
def crossover(quality, mult):
    q = Fraction(float(L.ro_quant_step(quality)))
    lam = Fraction(L.ro_lambda(quality, mult))
    nq = Fraction(quality, 100)
    b1, b0 = Fraction(1, 2) / q, 5 + Fraction(56, 100) * q
    dr = Fraction(15, 100) + (q - 1) / 111 * Fraction(55, 100)
    sk, keep = Fraction(55, 100) - nq * Fraction(15, 100), Fraction(120, 100) + nq * Fraction(35, 100)
    slope, rhs = 1 - dr - lam * b1, lam * (keep + b0 - sk)
    return slope, rhs


@pytest.mark.parametrize("quality,mult", [(80, 0.49), (70, 0.49), (95, 1.0), (60, 2.5)])
def test_rdo_crossover_closed_form(quality, mult):
    slope, rhs = crossover(quality, mult)
    assert slope > 0
    cut = rhs / slope   # skip for SAD <= cut
    assert abs(cut - round(cut)) > Fraction(1, 10 ** 6)   # away from a tie, where float rounding decides
    s_star = math.floor(cut)
    for thr in (3, 255):
        for sad in sorted({0, 1, max(0, s_star - 1), s_star, s_star + 1, s_star + 2, 768 * thr, 768 * thr + 1}):
            want = sad <= cut and sad <= 768 * thr
            assert bool(L.ro_rdo_skip(quality, mult, thr, sad, 1)) == want, (sad, thr)
            assert not L.ro_rdo_skip(quality, mult, thr, sad, 0)   # fractional vectors never skip
        assert not L.ro_rdo_skip(quality, mult, 0, 0, 1)            # skip threshold 0 disables skipping
    if quality == 80 and mult == 0.49:
        assert s_star == 12


def test_rdo_guard_binds_when_every_cost_says_skip():
    """A large lambda makes the coded cost grow faster than the skip cost (negative slope), so only
    the SAD guard 768 * skip_threshold decides."""
    for quality, mult in ((80, 100.0), (30, 0.49)):
        slope, _ = crossover(quality, mult)
        assert slope < 0
        for thr in (1, 2):
            g = 768 * thr
            assert L.ro_rdo_skip(quality, mult, thr, g, 1) and not L.ro_rdo_skip(quality, mult, thr, g + 1, 1)


# =============================================================================================
# Deblocking: the VP8 normal macroblock-edge filter (RFC 6386 sections 15.2-15.3) with the
# thresholds of SPEC.md 14.1 and the order and masks of 14.2-14.3 (deblock.rs:40-202)
# =============================================================================================
def filter_line(px, E, I, hev_t):
    p3, p2, p1, p0, q0, q1, q2, q3 = px
    if 2 * abs(p0 - q0) + abs(q1 - p1) // 2 > E:
        return px
    if max(abs(p3 - p2), abs(p2 - p1), abs(p1 - p0), abs(q1 - q0), abs(q2 - q1), abs(q3 - q2)) > I:
        return px
    c = lambda v: max(-128, min(127, v))
    P3, P2, P1, P0, Q0, Q1, Q2, Q3 = (v - 128 for v in px)
    w = c(c(P1 - Q1) + 3 * (Q0 - P0))
    if abs(p1 - p0) > hev_t or abs(q1 - q0) > hev_t:
        f1, f2 = c(w + 4) >> 3, c(w + 3) >> 3
        Q0, P0 = c(Q0 - f1), c(P0 + f2)
    else:
        a = c((27 * w + 63) >> 7); Q0, P0 = c(Q0 - a), c(P0 + a)
        a = c((18 * w + 63) >> 7); Q1, P1 = c(Q1 - a), c(P1 + a)
        a = c((9 * w + 63) >> 7); Q2, P2 = c(Q2 - a), c(P2 + a)
    return [max(0, min(255, v + 128)) for v in (P3, P2, P1, P0, Q0, Q1, Q2, Q3)]


def deblock_plane_ref(plane, grid, E, I, hev_t, edge_on):
    a = plane.astype(np.int32).copy()
    h, w = a.shape
    for x in range(grid, w, grid):          # every vertical edge, every row
        for y in range(h):
            if edge_on(x // grid - 1, y // grid, x // grid, y // grid):
                a[y, x - 4:x + 4] = filter_line(a[y, x - 4:x + 4].tolist(), E, I, hev_t)
    for y in range(grid, h, grid):          # then every horizontal edge, every column
        for x in range(w):
            if edge_on(x // grid, y // grid - 1, x // grid, y // grid):
                a[y - 4:y + 4, x] = filter_line(a[y - 4:y + 4, x].tolist(), E, I, hev_t)
    return a.astype(np.uint8)


@pytest.mark.parametrize("level,intra", [(14, False), (11, True), (45, False), (20, False), (1, True), (63, True)])
def test_deblock_against_rfc6386_restatement(level, intra):
    rng = np.random.default_rng(level * 3 + intra)
    w, h = 64, 48
    # smooth ramps with steps at block edges, plus noise: every filter branch (off, hev, wide) occurs
    base = (np.add.outer(np.arange(h) * 2, np.arange(w) * 3) % 200).astype(np.int32)
    steps = rng.integers(-20, 21, (h // 16 + 1, w // 16 + 1))
    Y = np.clip(base + np.kron(steps, np.ones((16, 16), np.int32))[:h, :w] + rng.integers(-3, 4, (h, w)), 0, 255).astype(np.uint8)
    U = np.clip(Y[::2, ::2].astype(np.int32) // 2 + 60 + rng.integers(-4, 5, (h // 2, w // 2)), 0, 255).astype(np.uint8)
    V = np.clip(255 - U.astype(np.int32) + rng.integers(-6, 7, U.shape), 0, 255).astype(np.uint8)
    mask = None if intra else (rng.random((h // 16) * (w // 16)) < 0.5).astype(np.uint8)
    gy, gu, gv = Y.copy(), U.copy(), V.copy()
    L.ro_deblock(vp(gy), vp(gu), vp(gv), w, h, level, int(intra), vp(mask) if mask is not None else None)
    I = max(level, 1)
    E = (level + 2) * 2 + I
    hev_t = (2 if level >= 40 else 1 if level >= 15 else 0) if intra else (3 if level >= 40 else 2 if level >= 20 else 1 if level >= 15 else 0)
    bw = w // 16
    flag = (lambda bx, by: True) if mask is None else (lambda bx, by: bool(mask[by * bw + bx]))
    on = lambda ax, ay, bx, by: flag(ax, ay) or flag(bx, by)
    assert np.array_equal(gy, deblock_plane_ref(Y, 16, E, I, hev_t, on))
    assert np.array_equal(gu, deblock_plane_ref(U, 8, E, I, hev_t, on))
    assert np.array_equal(gv, deblock_plane_ref(V, 8, E, I, hev_t, on))


# =============================================================================================
# Inverse transforms: closed forms (dct16.rs:196-412, fast_dct8.rs:123-174)
# =============================================================================================
def basis16(k, n):
    ck = math.sqrt(1 / 16) if k == 0 else math.sqrt(2 / 16)
    return round((1 << 18) * ck * math.cos((2 * n + 1) * k * math.pi / 32))


def r18(v):
    return (v + (1 << 17)) >> 18


@pytest.mark.parametrize("v0,u0,D", [(0, 0, 1000), (0, 0, -7), (0, 5, 333), (3, 0, -250), (7, 9, 64), (15, 15, 4000), (2, 13, -1)])
def test_idct16_single_coefficient_closed_form(v0, u0, D):
    """One coefficient D at row v0, column u0 (row-major, d[v*16+u]): rows first, then columns,
    each a 16-point product with the basis and (+2^17) >> 18:
    out[y][x] = R(R(D * B[u0][x]) * B[v0][y])."""
    d = np.zeros(256, np.int32)
    d[v0 * 16 + u0] = D
    out = np.zeros(256, np.int16)
    L.ro_idct16(vp(d), vp(out))
    want = np.array([[r18(r18(D * basis16(u0, x)) * basis16(v0, y)) for x in range(16)] for y in range(16)], np.int16)
    assert np.array_equal(out.reshape(16, 16), want)


@pytest.mark.parametrize("c0,q", [(1, 1), (-5, 23), (100, 23), (37, 7), (-1, 1), (511, 64)])
def test_idct8_dc_closed_form(c0, q):
    """DC only: every output sample is floor((c0*q + 4) / 8), the orthonormal 8x8 DC gain 1/8
    rounded half up."""
    c = np.zeros(64, np.int16)
    c[0] = c0
    qt = np.full(64, q, np.uint16)
    out = np.zeros(64, np.int16)
    L.ro_idct8(vp(c), vp(qt), vp(out))
    assert np.all(out == (c0 * q + 4) // 8)


@pytest.mark.parametrize("r,cc", [(0, 1), (1, 0), (0, 7), (5, 0), (3, 6), (7, 7)])
def test_idct8_single_coefficient_within_one(r, cc):
    """A single AC coefficient at row r, column cc: the output is the separable cosine product
    within one code value (the fixed-point AAN constants round), oriented rows by r, columns by cc."""
    D = 400
    c = np.zeros(64, np.int16)
    c[r * 8 + cc] = D
    out = np.zeros(64, np.int16)
    L.ro_idct8(vp(c), vp(np.ones(64, np.uint16)), vp(out))
    a = lambda k: math.sqrt(1 / 8) if k == 0 else math.sqrt(2 / 8)
    exact = np.array([[D * a(r) * a(cc) * math.cos((2 * y + 1) * r * math.pi / 16) * math.cos((2 * x + 1) * cc * math.pi / 16)
                       for x in range(8)] for y in range(8)])
    assert np.max(np.abs(out.reshape(8, 8) - exact)) <= 1.0


# =============================================================================================
# RGB <-> YUV420 (yuv.rs:41-110; the third-party `yuv` crate itself is unavailable): real-valued
# BT.709 limited range, 2x2 box-averaged chroma, within one code value
# =============================================================================================
KR, KB = 0.2126, 0.0722
KG = 1 - KR - KB


def yuv_real(r, g, b):
    yp = KR * r + KG * g + KB * b
    return 16 + 219 / 255 * yp, 128 + 224 / 255 * (b - yp) / (2 * (1 - KB)), 128 + 224 / 255 * (r - yp) / (2 * (1 - KR))


def test_rgb_to_yuv420_within_one_of_bt709():
    vals = np.arange(0, 256, 15)
    cols = np.array(np.meshgrid(vals, vals, vals, indexing="ij")).reshape(3, -1).T   # 5832 colours
    rng = np.random.default_rng(3)
    # each colour fills a 2x2 block; a second frame mixes four random colours per block
    n = len(cols)
    w, h = 2 * 54, 2 * (n // 54)
    for mixed in (False, True):
        img = np.zeros((h, w, 3), np.uint8)
        for k in range(n):
            by, bx = divmod(k, 54)
            blk = cols[rng.integers(0, n, 4)] if mixed else np.repeat(cols[k][None], 4, 0)
            img[2 * by:2 * by + 2, 2 * bx:2 * bx + 2] = blk.reshape(2, 2, 3)
        Y = np.zeros((h, w), np.uint8); U = np.zeros((h // 2, w // 2), np.uint8); V = np.zeros_like(U)
        L.ro_rgb_to_yuv420(vp(np.ascontiguousarray(img)), w, h, vp(Y), vp(U), vp(V))
        f = img.astype(np.float64)
        ry, _, _ = yuv_real(f[..., 0], f[..., 1], f[..., 2])
        assert np.max(np.abs(Y - ry)) <= 1.0
        m = f.reshape(h // 2, 2, w // 2, 2, 3).mean(axis=(1, 3))
        _, ru, rv = yuv_real(m[..., 0], m[..., 1], m[..., 2])
        assert np.max(np.abs(U - ru)) <= 1.0 and np.max(np.abs(V - rv)) <= 1.0


def test_yuv420_to_rgb_within_one_of_bt709():
    rng = np.random.default_rng(5)
    h, w = 64, 96
    Y = rng.integers(16, 236, (h, w)).astype(np.uint8)
    U = rng.integers(16, 241, (h // 2, w // 2)).astype(np.uint8)
    V = rng.integers(16, 241, (h // 2, w // 2)).astype(np.uint8)
    rgb = np.zeros((h, w, 3), np.uint8)
    L.ro_yuv420_to_rgb(vp(Y), vp(U), vp(V), w, h, vp(rgb))
    M = np.array([[KR, KG, KB], [-KR / (2 * (1 - KB)), -KG / (2 * (1 - KB)), 0.5], [0.5, -KG / (2 * (1 - KR)), -KB / (2 * (1 - KR))]])
    Minv = np.linalg.inv(M)
    yp = (Y.astype(np.float64) - 16) * 255 / 219
    pb = (np.repeat(np.repeat(U, 2, 0), 2, 1).astype(np.float64) - 128) * 255 / 224   # nearest chroma
    pr = (np.repeat(np.repeat(V, 2, 0), 2, 1).astype(np.float64) - 128) * 255 / 224
    want = np.clip(np.einsum("ij,hwj->hwi", Minv, np.stack([yp, pb, pr], -1)), 0, 255)
    assert np.max(np.abs(rgb - want)) <= 1.0
