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


@pytest.mark.parametrize("mode", [1, 2])
def test_hd_serial_and_split_residual_coding_write_the_oracle_stream(hd_oracle_stream, mode):
    """The host codes a P-frame's residuals serially or split across its workers (position-range
    modelling tasks, then rANS block groups side by side); both forms write the oracle's bytes."""
    frames, want = hd_oracle_stream
    cfg = riv2_amd.EncoderConfig(W, H, 50).with_intra_quality(85).with_inter_quality(80).with_search_range(12) \
        .with_skip_threshold(3).with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(1000))
    for depth in (0, 8):
        print(f"[hd split] mode {mode} depth {depth}", flush=True)
        enc = riv2_amd.Encoder(cfg.with_pipeline_depth(depth))
        enc.debug_split_coding(mode)
        for i in range(N):
            enc.encode_frame(riv2_amd.Frame(frames[i], W, H, i * 20))
        enc.finish()
        assert enc.output() == want, f"mode {mode}, pipeline depth {depth}: stream differs from the oracle"
