"""Learned form of the encoder: a BF16 MFMA MV-proposal network seeding the motion search.

The reference has no learned form; this module drives the network of riv2_amd/csrc/mvnet.hip
through the C ABI (include/riv_gpu.h, riv_mvnet_*):

* `MvNet`: parameters, one Adam step, gradients, raw outputs (layer-wise or fused kernels) and
  proposals;
* `collect_pairs`: (block, vector) training pairs from the plain GPU search, i.e. the encoder's
  own search vectors on synthetic sequences, with the frame pairs resident in HBM;
* `train`: minibatch training on those pairs;
* `compare`: plain and learned encodes of the same sequences at several qualities: bytes, PSNR-Y
  (the BT.601 luma of ri-quality, reitero_video_quality_test/src/main.rs:126-128 and
  :1034-1056), Bjontegaard rate difference, the fallback rate and the fraction of blocks whose
  learned SAD is at most the plain one.

The bitstream of the learned form is an ordinary RIV v5 stream: the plain decoder decodes it.
"""
import os
from ctypes import byref, c_float, c_void_p
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch   # before the backend initialises the HIP runtime, so that torch finds the device

from . import _lib as L
from .api import Encoder, EncoderConfig, Frame, KeyframeInterval, _check

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS_PATH = os.path.join(HERE, "mvnet_weights.npz")


class MvNet:
    """The proposal network on one device (riv_mvnet_create ... riv_mvnet_destroy)."""

    def __init__(self, device: int = 0, seed: int = 0):
        self._h = c_void_p()
        self.device = device
        _check(L.lib().riv_mvnet_create(device, seed, byref(self._h)))
        self._sources = []   # keeps the device tensors the sources point into alive

    @staticmethod
    def param_count() -> int:
        return int(L.lib().riv_mvnet_param_count())

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.param_count(), np.float32)
        _check(L.lib().riv_mvnet_get_params(self._h, p.ctypes.data, p.size))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, np.float32)
        _check(L.lib().riv_mvnet_set_params(self._h, p.ctypes.data, p.size))

    def save(self, path: str = WEIGHTS_PATH, **meta):
        np.savez(path, params=self.get_params(), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path: str = WEIGHTS_PATH, device: int = 0) -> "MvNet":
        with np.load(path, allow_pickle=False) as z:
            params = z["params"]
        net = cls(device)
        net.set_params(params)
        return net

    def set_sources(self, sources):
        """sources: (cur, ref, prev) per frame pair; cur/ref uint8 (H, W) device tensors, prev an
        int8 (nb, 2) device tensor of the previous P-frame's vectors or None."""
        arr = (L.MvSourceC * max(1, len(sources)))()
        for i, (cur, ref, prev) in enumerate(sources):
            assert cur.is_cuda and ref.is_cuda and cur.dtype == ref.dtype and cur.shape == ref.shape and cur.is_contiguous()
            h, w = cur.shape
            arr[i].cur_y = cur.data_ptr(); arr[i].ref_y = ref.data_ptr()
            arr[i].prev_mv = prev.data_ptr() if prev is not None else None
            arr[i].prev_stride = 2
            arr[i].width = w; arr[i].height = h
        _check(L.lib().riv_mvnet_set_sources(self._h, arr, len(sources)))
        self._sources = list(sources)

    def train_step(self, samples: np.ndarray, targets: np.ndarray, lr: float, want_loss: bool = False) -> Optional[float]:
        s = np.ascontiguousarray(samples, np.int32)
        t = np.ascontiguousarray(targets, np.int8)
        loss = c_float()
        _check(L.lib().riv_mvnet_train_step(self._h, s.ctypes.data, t.ctypes.data, len(s), lr, byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def gradients(self, samples: np.ndarray, targets: np.ndarray):
        s = np.ascontiguousarray(samples, np.int32)
        t = np.ascontiguousarray(targets, np.int8)
        g = np.zeros(self.param_count(), np.float32)
        loss = c_float()
        _check(L.lib().riv_mvnet_gradients(self._h, s.ctypes.data, t.ctypes.data, len(s), g.ctypes.data, byref(loss)))
        return g, loss.value

    def forward(self, samples: np.ndarray, fused: bool = False) -> np.ndarray:
        s = np.ascontiguousarray(samples, np.int32)
        y = np.zeros((len(s), 16), np.float32)
        _check(L.lib().riv_mvnet_forward(self._h, s.ctypes.data, len(s), 1 if fused else 0, y.ctypes.data))
        return y

    def propose(self, samples: np.ndarray, confidence: float, sr: int) -> np.ndarray:
        s = np.ascontiguousarray(samples, np.int32)
        p = np.zeros((len(s), 4), np.uint32)
        _check(L.lib().riv_mvnet_propose(self._h, s.ctypes.data, len(s), confidence, sr, p.ctypes.data))
        return p

    def close(self):
        if self._h:
            L.lib().riv_mvnet_destroy(self._h)
            self._h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def unpack_proposals(p: np.ndarray):
    """(n, 4, 2) int candidate vectors, (n,) confident flags, (n,) logits from propose()."""
    b = p[:, :2].astype(np.uint32).view(np.uint8).reshape(-1, 4, 2).view(np.int8).astype(np.int32)
    return b, p[:, 2] != 0, p[:, 3].view(np.float32)


def hd_config(quality: int = 80, intra_quality: int = 85, depth: int = 0, width: int = 1920, height: int = 1080) -> EncoderConfig:
    """The 1080p P-frame workload's settings: q85/80, sr 12, skip 3, one keyframe."""
    cfg = EncoderConfig(width, height, 30).with_intra_quality(intra_quality).with_inter_quality(quality) \
        .with_keyframe_interval(KeyframeInterval.fixed(1000))
    cfg._c.pipeline_depth = depth
    return cfg


@dataclass
class PairSet:
    """Training pairs: samples[i] = (source, block), targets[i] = the plain search's (x, y)."""
    sources: List[tuple] = field(default_factory=list)
    samples: np.ndarray = None
    targets: np.ndarray = None


def collect_pairs(seeds, pframes: int = 4, device: int = 0, log=None, params=None) -> PairSet:
    """Encodes 1 I + `pframes` P-frames of each seed's synthetic sequence (params(seed, frames);
    default the 1080p workload's generator) with the plain GPU encoder; every P-frame block gives
    one pair, with the frame pair resident on the device."""
    from . import synth
    if params is None:
        params = lambda seed, n: synth.hd_params(seed=seed, frames=n)
    dev = torch.device("cuda", device)
    ps = PairSet()
    samples, targets = [], []
    for seed in seeds:
        frames = synth.Sequence(params(seed, pframes + 1)).frames(0, pframes + 1)
        h, w = frames.shape[1:3]
        enc = Encoder(hd_config(width=w, height=h))
        enc.encode_frame(Frame(frames[0], w, h, 0))
        ref = torch.from_numpy(enc.debug_yuv(2)[0]).to(dev)
        prev = None
        for k in range(1, pframes + 1):
            enc.encode_frame(Frame(frames[k], w, h, 33 * k))
            cur = torch.from_numpy(enc.debug_yuv(0)[0]).to(dev)
            rec = enc.debug_blocks()[0]
            si = len(ps.sources)
            ps.sources.append((cur, ref, prev))
            nb = rec.shape[0]
            samples.append(np.stack([np.full(nb, si, np.int32), np.arange(nb, dtype=np.int32)], 1))
            targets.append(rec[:, 0:2].copy())
            prev = torch.from_numpy(np.ascontiguousarray(rec[:, 4:6])).to(dev)
            ref = torch.from_numpy(enc.debug_yuv(2)[0]).to(dev)
        enc.close()
        if log and (seed % 50 == 0 or len(seeds) <= 100):
            log(f"seed {seed}: {pframes} P-frames")
    ps.samples = np.concatenate(samples)
    ps.targets = np.concatenate(targets)
    return ps


def merge_pairs(parts) -> PairSet:
    """One pair set from several, their source indices renumbered."""
    ps = PairSet()
    samples, targets = [], []
    for p in parts:
        s = p.samples.copy()
        s[:, 0] += len(ps.sources)
        ps.sources.extend(p.sources)
        samples.append(s)
        targets.append(p.targets)
    ps.samples = np.concatenate(samples)
    ps.targets = np.concatenate(targets)
    return ps


def train(net: MvNet, ps: PairSet, epochs: int = 2, batch: int = 2048, lr: float = 2e-3, seed: int = 0, log=None):
    """Adam over shuffled minibatches with a cosine-decayed rate; returns the logged losses."""
    net.set_sources(ps.sources)
    rng = np.random.default_rng(seed)
    n = len(ps.samples)
    per_epoch = n // batch
    total = max(1, per_epoch * epochs)
    losses, step = [], 0
    for ep in range(epochs):
        perm = rng.permutation(n)
        for i in range(per_epoch):
            idx = perm[i * batch:(i + 1) * batch]
            rate = lr * 0.5 * (1.0 + np.cos(np.pi * step / total))
            want = step % 100 == 0 or step == total - 1
            l = net.train_step(ps.samples[idx], ps.targets[idx], rate, want)
            if want:
                losses.append((step, l))
                if log:
                    log(f"epoch {ep} step {step}/{total} loss {l:.4f}")
            step += 1
    return losses


def proposal_accuracy(net: MvNet, ps: PairSet, confidence: float = 0.5, sr: int = 12, limit: int = 200_000):
    """Over (up to `limit` of) the pairs: the fraction whose nearest proposal is within one pixel of
    the plain vector, the confident fraction, and the first fraction among confident blocks."""
    net.set_sources(ps.sources)
    s, t = ps.samples[:limit], ps.targets[:limit].astype(np.int32)
    c, conf, _ = unpack_proposals(net.propose(s, confidence, sr))
    near = (np.abs(c - t[:, None, :]).max(axis=2) <= 1).any(axis=1)
    return float(near.mean()), float(conf.mean()), float(near[conf].mean()) if conf.any() else 0.0


def psnr_y(a: np.ndarray, b: np.ndarray) -> float:
    """PSNR of the BT.601 luma of two RGB24 frames (ri-quality's compute_psnr, y part)."""
    ta = torch.from_numpy(a).to("cuda", torch.float64)
    tb = torch.from_numpy(b).to("cuda", torch.float64)
    w = torch.tensor([0.299, 0.587, 0.114], dtype=torch.float64, device="cuda")
    mse = float((((ta - tb) * w).sum(-1) ** 2).mean())
    return float("inf") if mse == 0.0 else 10.0 * np.log10(255.0 * 255.0 / mse)


def bd_rate(r1, p1, r2, p2) -> float:
    """Bjontegaard rate difference of curve 2 against curve 1 (percent; cubic fits of log-rate
    against PSNR, integrated over the overlapping PSNR range)."""
    l1, l2 = np.log(np.asarray(r1, float)), np.log(np.asarray(r2, float))
    f1, f2 = np.polyfit(p1, l1, 3), np.polyfit(p2, l2, 3)
    lo, hi = max(min(p1), min(p2)), min(max(p1), max(p2))
    i1, i2 = np.polyint(f1), np.polyint(f2)
    a1 = (np.polyval(i1, hi) - np.polyval(i1, lo)) / (hi - lo)
    a2 = (np.polyval(i2, hi) - np.polyval(i2, lo)) / (hi - lo)
    return float((np.exp(a2 - a1) - 1.0) * 100.0)


def encode_sequence(frames: np.ndarray, quality: int, net: Optional[MvNet] = None, confidence: float = 0.5, passes: int = 0):
    """(bitstream, mean PSNR-Y of the decoded P-frames, learned counts, per-frame block SADs).
    With a network, `passes` > 0 stops each motion search after that many relaxation passes."""
    from .api import Decoder
    h, w = frames.shape[1:3]
    enc = Encoder(hd_config(quality, width=w, height=h))
    if net is not None:
        enc.set_mvnet(net, confidence)
        enc.set_learned_passes(passes)
    sads = []
    for k in range(len(frames)):
        enc.encode_frame(Frame(frames[k], w, h, 33 * k))
        if k > 0:
            sads.append(enc.debug_blocks()[1].copy())
    counts = enc.learned_stats()
    enc.finish()
    data = enc.output()
    enc.close()
    dec = Decoder(data)
    ps = []
    for k in range(len(frames)):
        f = dec.decode_frame()
        if k > 0:
            ps.append(psnr_y(frames[k], f.data))
    return data, float(np.mean(ps)), counts, sads


def compare(net: MvNet, seeds, frames: int = 10, qualities=(70, 75, 80, 85), confidence: float = 0.5, log=None, passes: int = 0,
            params=None):
    """Plain against learned on held-out sequences: totals per quality, BD-rate, fallback rate and
    the share of first-P-frame blocks (same reference in both forms) with learned SAD <= plain.
    params(seed, frames) gives a sequence's generator parameters (default: the 1080p workload's)."""
    from . import synth
    if params is None:
        params = lambda seed, n: synth.hd_params(seed=seed, frames=n)
    rows = []
    le_num = le_den = 0
    blocks = confident = taken = 0
    for seed in seeds:
        fr = synth.Sequence(params(seed, frames)).frames(0, frames)
        for q in qualities:
            pb, pp, _, psad = encode_sequence(fr, q)
            lb, lp, cnt, lsad = encode_sequence(fr, q, net, confidence, passes)
            rows.append((seed, q, len(pb), pp, len(lb), lp))
            le_num += int((lsad[0] <= psad[0]).sum()); le_den += len(psad[0])
            blocks += cnt[0]; confident += cnt[1]; taken += cnt[2]
            if log:
                log(f"seed {seed} q{q}: plain {len(pb)} B {pp:.3f} dB, learned {len(lb)} B {lp:.3f} dB, taken {cnt[2]}/{cnt[0]}")
    out = {"per_quality": {}}
    for q in qualities:
        sel = [r for r in rows if r[1] == q]
        out["per_quality"][q] = {
            "plain_bytes": sum(r[2] for r in sel), "plain_psnr_y": float(np.mean([r[3] for r in sel])),
            "learned_bytes": sum(r[4] for r in sel), "learned_psnr_y": float(np.mean([r[5] for r in sel])),
        }
    pq = out["per_quality"]
    out["bd_rate_percent"] = bd_rate([pq[q]["plain_bytes"] for q in qualities], [pq[q]["plain_psnr_y"] for q in qualities],
                                     [pq[q]["learned_bytes"] for q in qualities], [pq[q]["learned_psnr_y"] for q in qualities])
    out["fallback_rate"] = 1.0 - taken / max(1, blocks)
    out["confident_rate"] = confident / max(1, blocks)
    out["learned_sad_le_plain"] = le_num / max(1, le_den)
    return out
