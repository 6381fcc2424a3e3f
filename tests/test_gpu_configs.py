"""GPU parity on the workloads the throughput figures are quoted on.

* 1080p (1920x1080, 1 I + 49 P, q85/80, sr 12): every frame record, every reconstruction, the
  whole stream and every decoded RGB frame against the oracle's committed digests
  (tests/golden/hd_config2.json).
* Batched (64 streams, 16 each of 352x288, 640x360, 1280x720, 1920x1080, 1 I + 20 P, one batch
  step per frame of every stream): every stream and its decoded RGB against the oracle's solo
  encodes' digests (tests/golden/batch_config3.json).
* Learned form at 1080p with the committed trained network: every frame of the encoder's
  reconstruction equals the oracle decoder's output on the learned stream, and the GPU decoder
  equals the oracle decoder.
* Learned + batched stress (64 streams of high-motion content, 1 I + 49 P): the same check for
  every frame of every stream.

The oracle decoder is the reference decoder's restatement (oracle/), so "equals the oracle
decoder" is decoding kept bit-exact to the reference decoder."""
import hashlib
import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
import riv2_amd
from riv2_amd import learned, synth

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(a if isinstance(a, (bytes, bytearray)) else np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def split_frames(stream):
    """Frame records of a .riv stream (reitero_video_common/src/format.rs:171-205), after the
    36-byte header."""
    out, p = [], 36
    while p < len(stream):
        ft = stream[p + 8]
        q = p + 10
        if ft == 1:
            q += 4 + int.from_bytes(stream[q:q + 4], "little")
        else:
            q += 3
            q += 4 + int.from_bytes(stream[q:q + 4], "little")
            q += 4 + int.from_bytes(stream[q:q + 4], "little")
        out.append(stream[p:q])
        p = q
    return out


def _cfg(w, h, fps=30, depth=0):
    return riv2_amd.EncoderConfig(w, h, fps).with_intra_quality(85).with_inter_quality(80).with_search_range(12) \
        .with_skip_threshold(3).with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(1000)).with_pipeline_depth(depth)


@pytest.mark.timeout(600)
def test_hd_50_frames_against_fixture():
    with open(os.path.join(GOLDEN, "hd_config2.json")) as f:
        want = json.load(f)
    wl = want["workload"]
    w, h, n = wl["width"], wl["height"], wl["frames"]
    frames = synth.Sequence(synth.hd_params(seed=wl["seed"], frames=n)).frames(0, n)
    enc = riv2_amd.Encoder(_cfg(w, h, 50, depth=0))
    for i in range(n):
        print(f"[config hd] frame {i}", flush=True)
        st = enc.encode_frame_with_stats(riv2_amd.Frame(frames[i], w, h, i * 20))
        fw = want["frames"][i]
        assert st.total_bytes == fw["frame_bytes"], f"frame {i} size"
        assert sha(*enc.debug_yuv(2)) == fw["recon_sha256"], f"frame {i} reconstruction"
    enc.finish()
    data = enc.output()
    recs = split_frames(data)
    assert len(recs) == n
    for i, r in enumerate(recs):
        assert sha(r) == want["frames"][i]["frame_sha256"], f"frame {i} record"
    assert len(data) == want["stream_bytes"] and sha(data) == want["stream_sha256"], "stream"
    # the pipelined encoder writes the same stream
    penc = riv2_amd.Encoder(_cfg(w, h, 50, depth=16))
    for i in range(n):
        penc.encode_frame(riv2_amd.Frame(frames[i], w, h, i * 20))
    penc.finish()
    assert penc.output() == data, "pipelined stream"
    penc.close()
    dec = riv2_amd.Decoder(data)
    for i in range(n):
        assert sha(dec.decode_frame().data) == want["frames"][i]["decoded_rgb_sha256"], f"frame {i} decoded RGB"
    dec.close()


def _batch_frames(w, h, seed, frames):
    return synth.Sequence(synth.custom_params(w, h, frames, seed=seed, global_motion=3.0, n_rects=4, rect_min=32, rect_max=96,
                                              rect_velocity=6.0)).frames()


@pytest.mark.timeout(900)
def test_batched_64_streams_against_fixture():
    import torch
    with open(os.path.join(GOLDEN, "batch_config3.json")) as f:
        want = json.load(f)
    n = want["workload"]["frames"]
    streams = want["streams"]
    print(f"[config batch] generating {len(streams)} streams", flush=True)
    dev = [torch.from_numpy(_batch_frames(s["width"], s["height"], s["seed"], n)).to("cuda:0") for s in streams]
    torch.cuda.synchronize()
    encs = [riv2_amd.Encoder(_cfg(s["width"], s["height"], depth=2)) for s in streams]
    for i in range(n):
        print(f"[config batch] step {i}", flush=True)
        riv2_amd.encode_batch_step(encs, [d[i].data_ptr() for d in dev], [i * 33] * len(encs), on_device=True)
    outs = []
    for e in encs:
        e.finish()
        outs.append(e.output())
        e.close()
    del dev
    bad = [k for k, (s, o) in enumerate(zip(streams, outs)) if len(o) != s["stream_bytes"] or sha(o) != s["stream_sha256"]]
    assert not bad, f"streams differ from the oracle's solo encodes: {bad}"
    for k, (s, o) in enumerate(zip(streams, outs)):
        dec = riv2_amd.Decoder(o)
        hh = hashlib.sha256()
        for _ in range(n):
            hh.update(np.ascontiguousarray(dec.decode_frame().data).tobytes())
        dec.close()
        assert hh.hexdigest() == s["decoded_rgb_sha256"], f"stream {k} decoded RGB"


def _oracle_decode_digests(stream, nframes):
    """Per-frame (YUV reconstruction, RGB) digests of the oracle decoder's output."""
    rdec = oracle.OracleDecoder(stream)
    out = []
    for _ in range(nframes):
        rgb, _ = rdec.decode()
        out.append((sha(*rdec.last_yuv()), sha(rgb)))
    return out


@pytest.mark.timeout(600)
@pytest.mark.parametrize("passes", [0, 4])
def test_learned_1080p_trained_network_decodes_bit_exact(passes):
    if not os.path.exists(learned.WEIGHTS_PATH):
        pytest.skip("no trained weights committed")
    n = 12
    w, h = 1920, 1080
    frames = synth.Sequence(synth.hd_params(seed=1003, frames=50)).frames(0, n)
    net = learned.MvNet.load(learned.WEIGHTS_PATH)
    enc = riv2_amd.Encoder(_cfg(w, h, 50, depth=0))
    enc.set_mvnet(net, 0.5)
    enc.set_learned_passes(passes)
    recon = []
    for i in range(n):
        print(f"[config learned passes {passes}] frame {i}", flush=True)
        st = enc.encode_frame_with_stats(riv2_amd.Frame(frames[i], w, h, i * 20))
        if passes and i > 0:
            assert st.me_passes == passes
        recon.append(sha(*enc.debug_yuv(2)))
    counts = enc.learned_stats()
    enc.finish()
    data = enc.output()
    enc.close()
    net.close()
    assert counts[2] > 0, "no block took the network's proposal"
    want = _oracle_decode_digests(data, n)
    dec = riv2_amd.Decoder(data)
    for i in range(n):
        assert recon[i] == want[i][0], f"frame {i}: encoder reconstruction differs from the oracle decoder"
        assert sha(dec.decode_frame().data) == want[i][1], f"frame {i}: GPU decoder differs from the oracle decoder"
    dec.close()


@pytest.mark.timeout(1200)
@pytest.mark.parametrize("passes", [0, 4])
def test_learned_batched_stress_64_streams_decode_bit_exact(passes):
    if not os.path.exists(learned.WEIGHTS_PATH):
        pytest.skip("no trained weights committed")
    import torch
    sizes = [(352, 288), (640, 360), (1280, 720), (1920, 1080)]
    per_size, n = 16, 50
    jobs = [(w, h, 5000 + 100 * k + j) for k, (w, h) in enumerate(sizes) for j in range(per_size)]
    print(f"[config stress] generating {len(jobs)} streams", flush=True)
    dev = [torch.from_numpy(synth.Sequence(synth.stress_params(w, h, n, seed=s)).frames()).to("cuda:0") for w, h, s in jobs]
    torch.cuda.synchronize()
    net = learned.MvNet.load(learned.WEIGHTS_PATH)
    encs = [riv2_amd.Encoder(_cfg(w, h, depth=2)) for w, h, _ in jobs]
    for e in encs:
        e.set_mvnet(net, 0.5)
        e.set_learned_passes(passes)
    recon = [[] for _ in jobs]
    for i in range(n):
        print(f"[config stress passes {passes}] step {i}", flush=True)
        riv2_amd.encode_batch_step(encs, [d[i].data_ptr() for d in dev], [i * 33] * len(encs), on_device=True)
        for k, e in enumerate(encs):
            recon[k].append(sha(*e.debug_yuv(2)))
    counts = encs[0].learned_stats()
    outs = []
    for e in encs:
        e.finish()
        outs.append(e.output())
        e.close()
    net.close()
    del dev
    assert counts[2] > 0, "no block took the network's proposal"
    print("[config stress] oracle decoder on every stream", flush=True)
    with ThreadPoolExecutor(16) as ex:   # the oracle's ctypes calls release the GIL
        want = list(ex.map(lambda o: _oracle_decode_digests(o, n), outs))
    for k, o in enumerate(outs):
        for i in range(n):
            assert recon[k][i] == want[k][i][0], f"stream {k} frame {i}: encoder reconstruction differs from the oracle decoder"
        dec = riv2_amd.Decoder(o)
        for i in range(n):
            assert sha(dec.decode_frame().data) == want[k][i][1], f"stream {k} frame {i}: GPU decoder differs"
        dec.close()
