"""Full-size checks at 1920x1080 (storage 1920x1088), the workload the throughput is quoted on:
the pipelined GPU encoder's bitstream equals the oracle's, every decoded frame equals the encoder's
reconstruction (an encode -> decode round trip), and the GPU decoder equals the oracle decoder."""
import numpy as np
import pytest

import oracle
import riv2_amd
from riv2_amd import synth

pytestmark = pytest.mark.gpu

W, H, N = 1920, 1080, 4


def test_hd_bitstream_and_round_trip():
    frames = synth.Sequence(synth.hd_params(seed=2, frames=N)).frames(0, N)
    # synchronous encoder: per-frame reconstructions
    cfg = riv2_amd.EncoderConfig(W, H, 50).with_intra_quality(85).with_inter_quality(80).with_search_range(12) \
        .with_skip_threshold(3).with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(1000))
    enc = riv2_amd.Encoder(cfg.with_pipeline_depth(0))
    recons = []
    for i in range(N):
        print(f"[hd] encode {i}", flush=True)
        enc.encode_frame(riv2_amd.Frame(frames[i], W, H, i * 20))
        recons.append([p.copy() for p in enc.debug_yuv(2)])
    enc.finish()
    data = enc.output()
    # pipelined encoder: same stream
    penc = riv2_amd.Encoder(cfg.with_pipeline_depth(8))
    for i in range(N):
        penc.encode_frame(riv2_amd.Frame(frames[i], W, H, i * 20))
    penc.finish()
    assert penc.output() == data, "pipelined stream differs from synchronous stream"
    ref = oracle.OracleEncoder(oracle.config(W, H, fps=50))
    for i in range(N):
        print(f"[hd] oracle {i}", flush=True)
        ref.encode(frames[i], i * 20)
    assert ref.finish() == data, "stream differs from the oracle"
    dec, rdec = riv2_amd.Decoder(data), oracle.OracleDecoder(data)
    for i in range(N):
        print(f"[hd] decode {i}", flush=True)
        a = dec.decode_frame()
        for p, q in zip(dec.debug_yuv(), recons[i]):
            assert np.array_equal(p, q), f"frame {i}: decoder output differs from the encoder reconstruction"
        b, _ = rdec.decode()
        assert np.array_equal(a.data, b), f"frame {i}: decoded RGB differs from the oracle decoder"


@pytest.fixture(scope="module")
def hd_oracle_stream():
    frames = synth.Sequence(synth.hd_params(seed=5, frames=N)).frames(0, N)
    ref = oracle.OracleEncoder(oracle.config(W, H, fps=50))
    for i in range(N):
        ref.encode(frames[i], i * 20)
    return frames, ref.finish()


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_hd_residual_coding_modes_write_the_oracle_stream(hd_oracle_stream, mode):
    """A P-frame's residual symbols modelled on the GPU (mode 0, the default: the host only
    rANS-codes them), or modelled and coded on the host serially (1), split across its workers (2)
    or split when workers are idle (3): every form writes the oracle's bytes."""
    frames, want = hd_oracle_stream
    cfg = riv2_amd.EncoderConfig(W, H, 50).with_intra_quality(85).with_inter_quality(80).with_search_range(12) \
        .with_skip_threshold(3).with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(1000))
    for depth in (0, 8):
        print(f"[hd split] mode {mode} depth {depth}", flush=True)
        enc = riv2_amd.Encoder(cfg.with_pipeline_depth(depth))
        enc.debug_split_coding(mode)
        on_gpu = []
        for i in range(N):
            st = enc.encode_frame_with_stats(riv2_amd.Frame(frames[i], W, H, i * 20))
            on_gpu.append(st.residual_symbols_on_gpu)
        if depth:
            enc.flush()
            on_gpu = [st.residual_symbols_on_gpu for st in enc.pop_stats()]
        enc.finish()
        assert enc.output() == want, f"mode {mode}, pipeline depth {depth}: stream differs from the oracle"
        assert on_gpu == [False] + [mode == 0] * (N - 1), on_gpu


@pytest.mark.parametrize("cap", [1, 600_000])
def test_hd_frames_beyond_the_gpu_symbol_buffer_are_coded_on_the_host(hd_oracle_stream, cap):
    """A P-frame with more residual symbols than the GPU modelling may hold is coded on the host
    from the same coefficients: with a 1-symbol buffer every P-frame is, with 600,000 symbols the
    heavier frames are; the stream is the oracle's either way."""
    frames, want = hd_oracle_stream
    cfg = riv2_amd.EncoderConfig(W, H, 50).with_intra_quality(85).with_inter_quality(80).with_search_range(12) \
        .with_skip_threshold(3).with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(1000))
    for depth in (0, 8):
        print(f"[hd model cap] cap {cap} depth {depth}", flush=True)
        enc = riv2_amd.Encoder(cfg.with_pipeline_depth(depth))
        enc.debug_model_cap(cap)
        for i in range(N):
            enc.encode_frame(riv2_amd.Frame(frames[i], W, H, i * 20))
        enc.flush()
        on_gpu = [st.residual_symbols_on_gpu for st in enc.pop_stats()] if depth else None
        enc.finish()
        assert enc.output() == want, f"cap {cap}, pipeline depth {depth}: stream differs from the oracle"
        if depth:
            print("on GPU", on_gpu, flush=True)
            assert not on_gpu[0] and (cap > 1 or not any(on_gpu)), on_gpu
