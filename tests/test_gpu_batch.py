"""Batched encoding: one frame per stream per step, the motion searches of all streams in one
launch. Every stream must come out byte-identical to the oracle encoding it alone, for streams of
different sizes, with intra and inter frames mixed within a step, in synchronous and pipelined
mode."""
import numpy as np
import pytest

import oracle
import riv2_amd
from riv2_amd import synth

pytestmark = pytest.mark.gpu

# (width, height, keyframe interval, seed)
STREAMS = [(64, 48, 1000, 21), (200, 120, 1000, 22), (352, 288, 2, 23), (96, 64, 1000, 24)]
FRAMES = 4


def _frames(w, h, seed):
    return synth.Sequence(synth.custom_params(w, h, FRAMES, seed=seed, global_motion=5.0, n_rects=2, rect_min=12,
                                              rect_max=40, rect_velocity=7.0)).frames()


def _cfg(w, h, kf, depth):
    return riv2_amd.EncoderConfig(w, h, 30).with_intra_quality(85).with_inter_quality(80).with_search_range(12) \
        .with_skip_threshold(3).with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(kf)).with_pipeline_depth(depth)


def _oracle(frames, w, h, kf):
    ref = oracle.OracleEncoder(oracle.config(w, h, keyframe_interval=kf))
    for i, f in enumerate(frames):
        ref.encode(f, i * 33)
    return ref.finish()


@pytest.mark.parametrize("depth", [0, 2])
def test_batch_streams_match_solo_oracle(depth):
    seqs = [_frames(w, h, seed) for w, h, _, seed in STREAMS]
    encs = [riv2_amd.Encoder(_cfg(w, h, kf, depth)) for w, h, kf, _ in STREAMS]
    for i in range(FRAMES):
        print(f"[batch depth {depth}] step {i}", flush=True)
        riv2_amd.encode_batch_step(encs, [s[i] for s in seqs], [i * 33] * len(encs))
    outs = []
    for e in encs:
        e.finish()
        outs.append(e.output())
        e.close()
    for k, ((w, h, kf, _), seq, got) in enumerate(zip(STREAMS, seqs, outs)):
        want = _oracle(seq, w, h, kf)
        assert got == want, f"stream {k} ({w}x{h}) differs from the oracle"
        dec = riv2_amd.Decoder(got)
        rdec = oracle.OracleDecoder(want)
        for i in range(FRAMES):
            a = dec.decode_frame().data
            b, _ = rdec.decode()
            assert np.array_equal(a, b), f"stream {k} frame {i}: decoded RGB differs"
        dec.close()


def test_batch_then_solo_frames_on_the_same_encoders():
    # batch steps and single-frame calls interleave on the same encoders (the batch stream and each
    # encoder's own stream are ordered against each other)
    seqs = [_frames(w, h, seed) for w, h, _, seed in STREAMS[:2]]
    encs = [riv2_amd.Encoder(_cfg(w, h, kf, 0)) for w, h, kf, _ in STREAMS[:2]]
    for i in range(FRAMES):
        print(f"[mixed] step {i}", flush=True)
        if i % 2 == 0:
            riv2_amd.encode_batch_step(encs, [s[i] for s in seqs], [i * 33] * 2)
        else:
            for e, s, (w, h, _, _) in zip(encs, seqs, STREAMS[:2]):
                e.encode_frame(riv2_amd.Frame(s[i], w, h, i * 33))
    for e, seq, (w, h, kf, _) in zip(encs, seqs, STREAMS[:2]):
        e.finish()
        assert e.output() == _oracle(seq, w, h, kf)
        e.close()


def test_batch_rejects_mismatched_encoders():
    a = riv2_amd.Encoder(_cfg(64, 48, 1000, 0))
    b = riv2_amd.Encoder(_cfg(64, 48, 1000, 0).with_search_range(8))
    f = np.zeros((48, 64, 3), np.uint8)
    with pytest.raises(riv2_amd.EncodeError):
        riv2_amd.encode_batch_step([a, b], [f, f])
    with pytest.raises(riv2_amd.EncodeError):
        riv2_amd.encode_batch_step([a, a], [f, f])
    a.close()
    b.close()


def test_autotuner_settles_without_changing_any_stream():
    """The batch step's launch-shape autotuner tries every candidate shape on the first steps of a
    batch (three times each), keeps the fastest, and no stream's bytes change on the way."""
    streams = [(64, 48, 1000, 31), (200, 120, 1000, 32), (96, 64, 1000, 33)]
    nframes = 30
    seqs = [synth.Sequence(synth.custom_params(w, h, nframes, seed=seed, global_motion=5.0, n_rects=2, rect_min=12,
                                               rect_max=40, rect_velocity=7.0)).frames() for w, h, _, seed in streams]
    encs = [riv2_amd.Encoder(_cfg(w, h, kf, 0)) for w, h, kf, _ in streams]
    seen = set()
    settled_at = None
    for i in range(nframes):
        riv2_amd.encode_batch_step(encs, [s[i] for s in seqs], [i * 33] * len(encs))
        if i > 0:
            t = encs[0].batch_tuning()
            seen.add((t["waves"], t["groups"]))
            if t["settled"] and settled_at is None:
                settled_at = i
                print(f"[autotune] settled at step {i}: {t}", flush=True)
    t = encs[0].batch_tuning()
    assert t["settled"] and 0 <= t["chosen"] < t["candidates"] and t["candidates"] > 1
    assert len(seen) == t["candidates"], "every candidate shape was launched"
    outs = []
    for e in encs:
        e.finish()
        outs.append(e.output())
        e.close()
    for k, ((w, h, kf, _), seq, got) in enumerate(zip(streams, seqs, outs)):
        assert got == _oracle(seq, w, h, kf), f"stream {k} ({w}x{h}) differs from the oracle"
