"""Learned form: the BF16 MFMA proposal network against a float32 PyTorch restatement of the same
network (forward and gradients), the fused inference kernel against the layer-wise kernels,
training, proposals, and the learned encoder's streams (fallback-only equals the plain form byte
for byte; any stream decodes identically on the GPU decoder and the oracle decoder)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

import oracle
import riv2_amd
from riv2_amd import learned, synth

OFF = {}
_o = 0
for name, size in [("w1", 48 * 16), ("b1", 16), ("w2", 144 * 16), ("b2", 16), ("w3", 144 * 16), ("b3", 16),
                   ("w4", 272 * 128), ("b4", 128), ("w5", 128 * 16), ("b5", 16)]:
    OFF[name] = (_o, size)
    _o += size
NPARAM = _o


def unflat(p):
    return {k: p[o:o + s] for k, (o, s) in OFF.items()}


def bf(t):
    return t.to(torch.bfloat16).to(torch.float32)


def conv_weight(w, cin):
    """[K][16] with k = (3ky + kx) * cin + ci -> torch's [16][cin][3][3]."""
    w = w[:9 * cin].reshape(3, 3, cin, 16)
    return w.permute(3, 2, 0, 1).contiguous()


def inputs(cur, ref, prev, bw, bh, blocks):
    """The network's input (B, 4, 32, 32) and predictor inputs (B, 16), as float32."""
    H, W = cur.shape
    xs, ps = [], []
    offs = [(0, 0), (-1, 0), (1, 0), (0, -1), (0, 1), (-1, -1), (1, -1)]
    for b in blocks:
        bx, by = b % bw, b // bw
        ys = np.clip(np.arange(by * 16 - 8, by * 16 + 24), 0, H - 1)
        xs_ = np.clip(np.arange(bx * 16 - 8, bx * 16 + 24), 0, W - 1)
        c = cur[np.ix_(ys, xs_)].astype(np.float32)
        r = ref[np.ix_(ys, xs_)].astype(np.float32)
        xs.append(np.stack([(c - 128) / 64, (r - 128) / 64, (c - r) / 32, np.zeros_like(c)]))
        p = np.zeros(16, np.float32)
        if prev is not None:
            for v, (ox, oy) in enumerate(offs):
                x, y = min(max(bx + ox, 0), bw - 1), min(max(by + oy, 0), bh - 1)
                p[2 * v:2 * v + 2] = prev[y * bw + x] / 8.0
        ps.append(p)
    return torch.from_numpy(np.stack(xs)), torch.from_numpy(np.stack(ps))


def ref_forward(params, x, preds):
    """The network in float32 with the kernels' bf16 rounding points: bf16 weights and inputs,
    every activation rounded to bf16 after its ReLU."""
    P = {k: torch.from_numpy(v.copy()) for k, v in unflat(params).items()}
    w1 = bf(conv_weight(P["w1"].reshape(48, 16), 4))
    w2 = bf(conv_weight(P["w2"].reshape(144, 16), 16))
    w3 = bf(conv_weight(P["w3"].reshape(144, 16), 16))
    a = bf(torch.relu(torch.nn.functional.conv2d(bf(x), w1, P["b1"], stride=2, padding=1)))
    a = bf(torch.relu(torch.nn.functional.conv2d(a, w2, P["b2"], stride=2, padding=1)))
    a = bf(torch.relu(torch.nn.functional.conv2d(a, w3, P["b3"], stride=2, padding=1)))
    feat = torch.cat([a.permute(0, 2, 3, 1).reshape(len(x), 256), bf(preds)], 1)
    h = bf(torch.relu(feat @ bf(P["w4"].reshape(272, 128)) + P["b4"]))
    return h @ bf(P["w5"].reshape(128, 16)) + P["b5"]


def ref_loss(y, t):
    """The training loss (k_net_loss) in float32 autograd form."""
    tt = t.float() / 8.0
    c = y[:, :8].reshape(-1, 4, 2)
    d = torch.nn.functional.smooth_l1_loss(c, tt[:, None, :].expand_as(c), reduction="none", beta=0.125).sum(2)
    best = d.argmin(1)
    reg = d.gather(1, best[:, None])[:, 0] + 0.02 * d.sum(1)
    cb = c[torch.arange(len(c)), best]
    label = ((cb - tt).abs() <= 0.125).all(1).float().detach()
    z = y[:, 8]
    return (reg + torch.nn.functional.softplus(z) - label * z).mean()


def planes(seed, w=128, h=96, shift=(3, -2)):
    rng = np.random.default_rng(seed)
    ref = rng.integers(0, 256, (h, w), dtype=np.uint8)
    cur = np.roll(np.roll(ref, shift[1], 0), shift[0], 1)
    cur = np.clip(cur.astype(int) + rng.integers(-6, 7, cur.shape), 0, 255).astype(np.uint8)
    prev = rng.integers(-12, 13, ((w // 16) * (h // 16), 2)).astype(np.int8)
    return cur, ref, prev


def make_net(seed=3):
    net = learned.MvNet(0, seed)
    assert net.param_count() == NPARAM
    return net


def with_source(net, cur, ref, prev):
    dev = torch.device("cuda", 0)
    src = (torch.from_numpy(cur).to(dev), torch.from_numpy(ref).to(dev),
           torch.from_numpy(np.ascontiguousarray(prev)).to(dev) if prev is not None else None)
    net.set_sources([src])
    return src


def test_layerwise_forward_matches_float32_reference():
    print("[learned] layer-wise forward vs torch", flush=True)
    net = make_net()
    cur, ref, prev = planes(1)
    _src = with_source(net, cur, ref, prev)
    bw, bh = 128 // 16, 96 // 16
    blocks = np.arange(bw * bh)
    samples = np.stack([np.zeros_like(blocks), blocks], 1).astype(np.int32)
    y = net.forward(samples, fused=False)
    x, p = inputs(cur, ref, prev, bw, bh, blocks)
    want = ref_forward(net.get_params(), x, p).numpy()
    assert np.abs(y[:, 9:]).max() == 0.0
    np.testing.assert_allclose(y[:, :9], want[:, :9], rtol=0, atol=2e-2 * max(1.0, np.abs(want[:, :9]).max()))


def test_fused_inference_equals_layerwise():
    print("[learned] fused vs layer-wise", flush=True)
    net = make_net(5)
    cur, ref, prev = planes(2, 160, 112)
    _src = with_source(net, cur, ref, prev)
    nb = (160 // 16) * (112 // 16)
    # an uneven sample list: repeats, every block, a count that is not a multiple of 16
    blocks = np.concatenate([np.arange(nb), [0, nb - 1, 7]]).astype(np.int32)
    samples = np.stack([np.zeros_like(blocks), blocks], 1).astype(np.int32)
    a = net.forward(samples, fused=False)
    b = net.forward(samples, fused=True)
    np.testing.assert_allclose(b[:, :9], a[:, :9], rtol=0, atol=1e-3 * max(1.0, np.abs(a[:, :9]).max()))


def test_fused_inference_follows_new_weights():
    """The fused kernel reads a fragment-order copy of the weights, remade after they change: after
    new parameters and after a training step it still equals the layer-wise kernels, and the
    proposals follow the fused forward."""
    print("[learned] fused after weight changes", flush=True)
    net = make_net(31)
    cur, ref, prev = planes(6, 128, 96)
    _src = with_source(net, cur, ref, prev)
    nb = (128 // 16) * (96 // 16)
    samples = np.stack([np.zeros(nb), np.arange(nb)], 1).astype(np.int32)

    def check(label):
        a = net.forward(samples, fused=False)
        b = net.forward(samples, fused=True)
        np.testing.assert_allclose(b[:, :9], a[:, :9], rtol=0, atol=1e-3 * max(1.0, np.abs(a[:, :9]).max()), err_msg=label)
        c, conf, logit = learned.unpack_proposals(net.propose(samples, 0.5, 12))
        np.testing.assert_array_equal(logit, b[:, 8], err_msg=label)
        np.testing.assert_array_equal(c, np.clip(np.rint(8.0 * b[:, :8].reshape(-1, 4, 2)), -12, 12).astype(int), err_msg=label)
        return b

    y0 = check("fresh network")
    other = make_net(37)
    net.set_params(other.get_params())
    other.close()
    y1 = check("after set_params")
    assert not np.array_equal(y0[:, :9], y1[:, :9]), "new parameters did not change the fused output"
    targets = np.tile(np.array([[3, -2]], np.int8), (nb, 1))
    net.train_step(samples, targets, 1e-2)
    y2 = check("after a training step")
    assert not np.array_equal(y1[:, :9], y2[:, :9]), "a training step did not change the fused output"


def test_new_parameters_leave_frames_in_flight_alone():
    """Pipelined encoder: frames submitted before set_params run on the old parameters (their
    inference is queued on the encoder's stream; the upload waits for it), so the stream equals an
    encode that never saw the new parameters."""
    print("[learned] set_params with frames in flight", flush=True)
    fr = cif_frames(27, 10)
    a, b = make_net(41), make_net(43)
    want, _, _ = encode(fr, a, 1e-6, 8)
    net = make_net(41)
    h, w = fr.shape[1:3]
    enc = riv2_amd.Encoder(learned.hd_config(80, width=w, height=h, depth=8))
    enc.set_mvnet(net, 1e-6)
    for k in range(len(fr)):
        enc.encode_frame(riv2_amd.Frame(fr[k], w, h, 33 * k))
    net.set_params(b.get_params())   # every frame is submitted; some are still on the GPU
    enc.finish()
    out = enc.output()
    enc.close()
    assert out == want, "frames already submitted saw the new parameters"
    for n in (a, b, net):
        n.close()


def test_gradients_match_autograd():
    print("[learned] gradients vs autograd", flush=True)
    net = make_net(7)
    cur, ref, prev = planes(3)
    _src = with_source(net, cur, ref, prev)
    bw, bh = 128 // 16, 96 // 16
    rng = np.random.default_rng(0)
    blocks = rng.integers(0, bw * bh, 40).astype(np.int32)
    samples = np.stack([np.zeros_like(blocks), blocks], 1).astype(np.int32)
    targets = rng.integers(-12, 13, (40, 2)).astype(np.int8)
    g, loss = net.gradients(samples, targets)
    params = net.get_params()
    x, p = inputs(cur, ref, prev, bw, bh, blocks)
    leaf = torch.from_numpy(params.copy()).requires_grad_(True)
    # autograd through the same rounding points (bf16 rounding passes the gradient straight through)
    P = {k: leaf[o:o + s] for k, (o, s) in OFF.items()}
    ste = lambda t: t + (bf(t) - t).detach()
    w1 = ste(conv_weight(P["w1"].reshape(48, 16), 4))
    w2 = ste(conv_weight(P["w2"].reshape(144, 16), 16))
    w3 = ste(conv_weight(P["w3"].reshape(144, 16), 16))
    a = ste(torch.relu(torch.nn.functional.conv2d(bf(x), w1, P["b1"], stride=2, padding=1)))
    a = ste(torch.relu(torch.nn.functional.conv2d(a, w2, P["b2"], stride=2, padding=1)))
    a = ste(torch.relu(torch.nn.functional.conv2d(a, w3, P["b3"], stride=2, padding=1)))
    feat = torch.cat([a.permute(0, 2, 3, 1).reshape(40, 256), bf(p)], 1)
    h = ste(torch.relu(feat @ ste(P["w4"].reshape(272, 128)) + P["b4"]))
    y = h @ ste(P["w5"].reshape(128, 16)) + P["b5"]
    l = ref_loss(y, torch.from_numpy(targets.astype(np.int32)))
    l.backward()
    want = leaf.grad.numpy()
    lv = float(l.detach())
    assert abs(loss - lv) <= 1e-3 * max(1.0, abs(lv))
    for k, (o, s) in OFF.items():
        gk, wk = g[o:o + s], want[o:o + s]
        if k == "w5" or k == "b5":   # the unused head columns stay out of training
            gk, wk = gk.reshape(-1, 16)[:, :9], wk.reshape(-1, 16)[:, :9]
        if k == "w1":
            gk, wk = gk[:36 * 16], wk[:36 * 16]
        scale = max(np.abs(wk).max(), 1e-6)
        # the backward pass carries bf16 gradients between layers
        assert np.abs(gk - wk).max() <= 0.05 * scale, f"{k}: {np.abs(gk - wk).max()} vs scale {scale}"


def test_training_learns_a_global_shift():
    print("[learned] training", flush=True)
    net = make_net(11)
    dev = torch.device("cuda", 0)
    srcs, samples, targets = [], [], []
    rng = np.random.default_rng(4)
    for i in range(8):
        sx, sy = int(rng.integers(-6, 7)), int(rng.integers(-6, 7))
        base = synth.Sequence(synth.custom_params(192, 128, 1, seed=40 + i)).frames(0, 1)[0][:, :, 1]
        cur = np.ascontiguousarray(np.roll(np.roll(base, -sy, 0), -sx, 1))
        srcs.append((torch.from_numpy(cur).to(dev), torch.from_numpy(np.ascontiguousarray(base)).to(dev), None))
        nb = (192 // 16) * (128 // 16)
        samples.append(np.stack([np.full(nb, i), np.arange(nb)], 1))
        targets.append(np.tile(np.array([[sx, sy]], np.int8), (nb, 1)))
    net.set_sources(srcs)
    s, t = np.concatenate(samples).astype(np.int32), np.concatenate(targets)
    first = net.train_step(s, t, 0.0, True)
    losses = []
    for step in range(300):
        idx = rng.integers(0, len(s), 256)
        l = net.train_step(s[idx], t[idx], 2e-3 * 0.5 * (1.0 + np.cos(np.pi * step / 300)), step % 50 == 49)
        if l is not None:
            losses.append(l)
    last = net.train_step(s, t, 0.0, True)
    print(f"loss {first:.4f} -> {last:.4f} ({[round(x, 3) for x in losses]})")
    assert np.isfinite(last) and last < 0.5 * first


def test_proposals_stay_in_range_and_follow_the_threshold():
    print("[learned] proposals", flush=True)
    net = make_net(13)
    cur, ref, prev = planes(5)
    _src = with_source(net, cur, ref, prev)
    nb = (128 // 16) * (96 // 16)
    samples = np.stack([np.zeros(nb), np.arange(nb)], 1).astype(np.int32)
    y = net.forward(samples, fused=True)
    for sr in (4, 12):
        c, conf, logit = learned.unpack_proposals(net.propose(samples, 0.5, sr))
        assert np.abs(c).max() <= sr
        np.testing.assert_array_equal(c, np.clip(np.rint(8.0 * y[:, :8].reshape(-1, 4, 2)), -sr, sr).astype(int))
        np.testing.assert_array_equal(logit, y[:, 8])
        np.testing.assert_array_equal(conf, y[:, 8] >= 0.0)
    _, never, _ = learned.unpack_proposals(net.propose(samples, 1.0, 12))
    assert not never.any()


def cif_frames(seed, n):
    return synth.Sequence(synth.cif_params(seed=seed, frames=n)).frames(0, n)


def encode(frames, net=None, confidence=0.5, depth=0):
    h, w = frames.shape[1:3]
    enc = riv2_amd.Encoder(learned.hd_config(80, width=w, height=h, depth=depth))
    if net is not None:
        enc.set_mvnet(net, confidence)
    recons = []
    for k in range(len(frames)):
        enc.encode_frame(riv2_amd.Frame(frames[k], w, h, 33 * k))
        if depth == 0:
            recons.append(enc.debug_yuv(2))
    counts = enc.learned_stats()
    enc.finish()
    return enc.output(), recons, counts


def test_never_confident_is_the_plain_form():
    print("[learned] confidence 1 == plain", flush=True)
    net = make_net(17)
    fr = cif_frames(21, 6)
    plain, _, _ = encode(fr)
    got, _, counts = encode(fr, net, 1.0)
    assert got == plain
    assert counts[0] == 5 * 396 and counts[1] == 0 and counts[2] == 0


@pytest.mark.parametrize("depth", [0, 8])
def test_learned_stream_decodes_identically(depth):
    print(f"[learned] decode, depth {depth}", flush=True)
    net = make_net(19)
    fr = cif_frames(23, 8)
    data, recons, counts = encode(fr, net, 1e-6, depth)
    assert counts[0] == 7 * 396 and counts[1] == counts[0] and 0 < counts[2] <= counts[1]
    dec, rdec = riv2_amd.Decoder(data), oracle.OracleDecoder(data)
    for k in range(len(fr)):
        a = dec.decode_frame().data
        b, _ = rdec.decode()
        assert np.array_equal(a, b), f"frame {k}: GPU and oracle decoders differ"
        if recons:
            y, u, v = dec.debug_yuv()
            ry, ru, rv = recons[k]
            assert np.array_equal(y, ry) and np.array_equal(u, ru) and np.array_equal(v, rv), f"frame {k}: encoder recon != decoder"


def test_learned_streams_in_a_batch_equal_their_solo_encodes():
    """A batch step with learned and plain streams: each learned stream's bytes equal the same
    stream encoded alone in the learned form, the plain stream's equal the oracle's."""
    print("[learned] batch", flush=True)
    net = make_net(23)
    specs = [(160, 96, 41, True), (96, 64, 42, False), (200, 120, 43, True)]
    seqs = [synth.Sequence(synth.stress_params(w, h, 6, seed=s)).frames() for w, h, s, _ in specs]
    encs = []
    for w, h, _, use in specs:
        e = riv2_amd.Encoder(learned.hd_config(80, width=w, height=h))
        if use:
            e.set_mvnet(net, 1e-6)
        encs.append(e)
    for i in range(6):
        riv2_amd.encode_batch_step(encs, [s[i] for s in seqs], [33 * i] * len(encs))
    counts = encs[0].learned_stats()
    outs = []
    for e in encs:
        e.finish()
        outs.append(e.output())
        e.close()
    assert counts[0] == 5 * sum(((w + 15) // 16) * ((h + 15) // 16) for w, h, _, _ in specs) and counts[2] > 0
    for (w, h, s, use), seq, got in zip(specs, seqs, outs):
        if use:
            solo, _, _ = encode(seq, net, 1e-6)
            assert got == solo, f"{w}x{h}: batch learned stream differs from its solo encode"
        else:
            ref = oracle.OracleEncoder(oracle.config(w, h))
            for i, f in enumerate(seq):
                ref.encode(f, 33 * i)
            assert got == ref.finish(), f"{w}x{h}: plain stream in a learned batch differs from the oracle"


def test_fixed_passes_apply_only_to_learned_launches():
    """learned_passes > 0 stops a learned encoder's search after that many passes (the pass count
    reported equals it); a plain stream in the same batch keeps the whole launch relaxing to the
    fixed point, so it still equals the oracle, and a plain encoder ignores the setting."""
    print("[learned] fixed passes", flush=True)
    net = make_net(29)
    w, h = 200, 120
    seq = synth.Sequence(synth.custom_params(w, h, 5, seed=44, global_motion=6.0, n_rects=3, rect_min=16, rect_max=48,
                                             rect_velocity=9.0)).frames()
    ref = oracle.OracleEncoder(oracle.config(w, h))
    for i, f in enumerate(seq):
        ref.encode(f, 33 * i)
    want_plain = ref.finish()
    # solo learned encoder: every P-frame runs exactly two passes; the stream decodes on the oracle
    e = riv2_amd.Encoder(learned.hd_config(80, width=w, height=h))
    e.set_mvnet(net, 1e-6)
    e.set_learned_passes(2)
    for i, f in enumerate(seq):
        st = e.encode_frame_with_stats(riv2_amd.Frame(f, w, h, 33 * i))
        if i > 0:
            assert st.me_passes == 2
    e.finish()
    data = e.output()
    e.close()
    dec, rdec = riv2_amd.Decoder(data), oracle.OracleDecoder(data)
    for k in range(len(seq)):
        assert np.array_equal(dec.decode_frame().data, rdec.decode()[0]), f"frame {k}"
    # a plain encoder ignores the setting
    p = riv2_amd.Encoder(learned.hd_config(80, width=w, height=h))
    p.set_learned_passes(2)
    for i, f in enumerate(seq):
        p.encode_frame(riv2_amd.Frame(f, w, h, 33 * i))
    p.finish()
    assert p.output() == want_plain
    p.close()
    # mixed batch: the learned stream's setting does not truncate the plain stream's search
    encs = [riv2_amd.Encoder(learned.hd_config(80, width=w, height=h)) for _ in range(2)]
    encs[0].set_mvnet(net, 1e-6)
    encs[0].set_learned_passes(2)
    for i in range(len(seq)):
        riv2_amd.encode_batch_step(encs, [seq[i], seq[i]], [33 * i, 33 * i])
    outs = []
    for x in encs:
        x.finish()
        outs.append(x.output())
        x.close()
    assert outs[1] == want_plain
