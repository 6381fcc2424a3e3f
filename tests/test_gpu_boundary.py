"""The drop-in boundary as the reference's callers use it: the VideoWriter / VideoReader traits
through the C vtables (riv_encoder_create / riv_decoder_create; reitero_encode/src/writer.rs:5-18,
reitero_decode/src/reader.rs:4-13, decoder.rs:112-171), decode_frame_null and
set_skip_residuals (decoder.rs:463-499), Encoder::config (encoder.rs:974-976), and two encoder
handles driven from two threads at once (the header's one-handle-per-thread model), whose
persistent motion searches must not share the GPU. Every stream is compared with the oracle."""
import threading

import numpy as np
import pytest

import oracle
import riv2_amd
from riv2_amd import synth

pytestmark = pytest.mark.gpu

W, H, N = 352, 288, 6


def _frames(seed=5, n=N, w=W, h=H):
    return synth.Sequence(synth.custom_params(w, h, n, seed=seed)).frames()


def _cfg(depth, w=W, h=H, kf=1000):
    return riv2_amd.EncoderConfig(w, h, 30).with_intra_quality(85).with_inter_quality(80) \
        .with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(kf)).with_pipeline_depth(depth)


def _oracle_stream(frames, w=W, h=H, kf=1000):
    ref = oracle.OracleEncoder(oracle.config(w, h, keyframe_interval=kf))
    for i, f in enumerate(frames):
        ref.encode(f, i * 33)
    return ref.finish()


class BytesReader:
    """A VideoReader over bytes, handing out at most `chunk` bytes per read (short reads)."""

    def __init__(self, data, chunk=4093):
        self.data, self.pos, self.chunk = bytes(data), 0, chunk

    def read(self, n):
        out = self.data[self.pos:self.pos + min(n, self.chunk)]
        self.pos += len(out)
        return out

    def position(self):
        return self.pos

    def seek(self, p):
        self.pos = p


@pytest.mark.parametrize("depth", [0, 4])
def test_writer_vtable_stream_equals_oracle(depth):
    frames = _frames()
    w = riv2_amd.VecWriter()
    enc = riv2_amd.Encoder(_cfg(depth), writer=w)
    cfg = enc.get_config()
    assert (cfg.display_width, cfg.display_height, cfg.storage_height, cfg.pipeline_depth) == (W, H, 288, depth)
    for i, f in enumerate(frames):
        print(f"[boundary writer depth {depth}] frame {i}", flush=True)
        enc.encode_frame(riv2_amd.Frame(f, W, H, i * 33))
    enc.finish()
    enc.close()
    assert w.into_inner() == _oracle_stream(frames)


def test_reader_vtable_decode_equals_oracle_decoder():
    data = _oracle_stream(_frames())
    dec = riv2_amd.Decoder(reader=BytesReader(data))
    rdec = oracle.OracleDecoder(data)
    assert dec.header().frame_count == N
    for i in range(N):
        print(f"[boundary reader] frame {i}", flush=True)
        assert dec.has_more_frames()
        a = dec.decode_frame()
        b, info = rdec.decode()
        assert np.array_equal(a.data, b), f"frame {i}"
        assert (a.frame_index, a.frame_type, a.timestamp) == (info.frame_index, info.frame_type, info.timestamp)
    assert not dec.has_more_frames()
    with pytest.raises(riv2_amd.DecodeError) as ei:
        dec.decode_frame()
    assert ei.value.kind == "EndOfStream"
    dec.close()


@pytest.mark.parametrize("use_reader", [False, True])
def test_decode_frame_null_then_decode(use_reader):
    data = _oracle_stream(_frames())
    dec = riv2_amd.Decoder(reader=BytesReader(data)) if use_reader else riv2_amd.Decoder(data)
    rdec = oracle.OracleDecoder(data)
    for i in range(3):   # decoded and discarded: later frames still predict from them
        dec.decode_frame_null()
        rdec.decode()
    for i in range(3, N):
        a = dec.decode_frame().data
        b, _ = rdec.decode()
        assert np.array_equal(a, b), f"frame {i} after decode_frame_null"
    with pytest.raises(riv2_amd.DecodeError):
        dec.decode_frame_null()
    dec.close()


@pytest.mark.parametrize("use_reader", [False, True])
def test_skip_residuals_equals_oracle(use_reader):
    data = _oracle_stream(_frames())
    dec = riv2_amd.Decoder(reader=BytesReader(data)) if use_reader else riv2_amd.Decoder(data)
    rdec = oracle.OracleDecoder(data)
    for i in range(N):
        skip = i in (0, 2, 3)   # switched on and off mid-stream; intra and inter frames both
        dec.set_skip_residuals(skip)
        rdec.set_skip_residuals(skip)
        a = dec.decode_frame()
        b, _ = rdec.decode()
        assert np.array_equal(a.data, b), f"frame {i} (skip_residuals={skip})"
        assert np.array_equal(dec.debug_yuv()[0], rdec.last_yuv()[0]), f"frame {i} luma"
    dec.close()


@pytest.mark.parametrize("depth", [0, 3])
def test_two_encoders_on_two_threads(depth):
    # different sizes, so the two persistent searches have different grids and pass counts
    jobs = [(352, 288, 31), (640, 368, 32)]
    seqs = [_frames(seed, 8, w, h) for w, h, seed in jobs]
    outs, errs = [None, None], []

    def run(k):
        try:
            w, h, _ = jobs[k]
            enc = riv2_amd.Encoder(_cfg(depth, w, h))
            for i, f in enumerate(seqs[k]):
                enc.encode_frame(riv2_amd.Frame(f, w, h, i * 33))
            enc.finish()
            outs[k] = enc.output()
            enc.close()
        except Exception as ex:   # reported from the main thread
            errs.append(ex)

    for rep in range(2):
        print(f"[boundary threads depth {depth}] round {rep}", flush=True)
        th = [threading.Thread(target=run, args=(k,)) for k in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        assert not errs, errs
        for k, (w, h, _) in enumerate(jobs):
            assert outs[k] == _oracle_stream(seqs[k], w, h), f"thread {k} ({w}x{h}) round {rep}"


@pytest.mark.parametrize("pinned", [False, True])
@pytest.mark.parametrize("depth", [0, 4])
def test_host_frame_memory_is_free_when_the_call_returns(pinned, depth):
    """riv_encode_frame copies the host frame before it returns (the copy runs on a stream of its
    own, double-buffered on the GPU): a caller that refills ONE buffer, pageable or pinned, for
    every frame gets the oracle's stream. Also through a batch step with host frames."""
    import torch
    frames = _frames(seed=9, n=8)
    want = _oracle_stream(frames)
    if pinned:
        t = torch.empty(frames[0].size, dtype=torch.uint8, pin_memory=True)
        buf = t.numpy().reshape(frames[0].shape)
    else:
        buf = np.empty_like(frames[0])
    enc = riv2_amd.Encoder(_cfg(depth))
    for i, f in enumerate(frames):
        print(f"[boundary host memory pinned={pinned} depth {depth}] frame {i}", flush=True)
        buf[...] = f
        enc.encode_frame(riv2_amd.Frame(buf, W, H, i * 33))
        buf[...] = 0   # the next frame's data, or garbage, must not reach this frame
    enc.finish()
    assert enc.output() == want
    enc.close()
    # batch steps of two streams from two reused buffers
    bufs = [np.empty_like(frames[0]) for _ in range(2)]
    if pinned:
        ts = [torch.empty(frames[0].size, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        bufs = [x.numpy().reshape(frames[0].shape) for x in ts]
    encs = [riv2_amd.Encoder(_cfg(depth)) for _ in range(2)]
    for i, f in enumerate(frames):
        for b in bufs:
            b[...] = f
        riv2_amd.encode_batch_step(encs, bufs, [i * 33] * 2)
        for b in bufs:
            b[...] = 0
    for e in encs:
        e.finish()
        assert e.output() == want
        e.close()


@pytest.mark.parametrize("use_reader", [False, True])
def test_decoder_current_frame_and_phase_timings(use_reader):
    """current_frame counts the frames returned; drain_timings returns the phase sums and resets
    them (decoder.rs:474-489); the device phases are measured only with profiling on."""
    data = _oracle_stream(_frames())
    dec = riv2_amd.Decoder(reader=BytesReader(data)) if use_reader else riv2_amd.Decoder(data)
    assert dec.current_frame() == 0 and not dec.skip_residuals()
    dec.decode_frame()
    dec.decode_frame()
    assert dec.current_frame() == 2
    t = dec.drain_timings()
    assert t["read_bits_ns"] > 0 and t["parse_frame_ns"] > 0 and t["mv_decode_ns"] > 0 and t["residual_rans_ns"] > 0
    assert t["build_pred_ns"] == t["residual_apply_ns"] == t["yuv_to_rgb_ns"] == 0
    assert t["residual_ns"] == t["residual_rans_ns"]
    dec.set_profiling(True)
    dec.decode_frame()
    t2 = dec.drain_timings()
    assert t2["build_pred_ns"] > 0 and t2["residual_apply_ns"] > 0 and t2["yuv_to_rgb_ns"] > 0
    assert t2["residual_ns"] == t2["residual_rans_ns"] + t2["residual_apply_ns"]
    if use_reader:   # frames are decoded in the calls (an in-memory stream's look-ahead runs on)
        assert all(v == 0 for v in dec.drain_timings().values())
    dec.set_skip_residuals(True)
    assert dec.skip_residuals()
    dec.close()
