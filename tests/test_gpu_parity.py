"""GPU encoder/decoder parity against the oracle, stage by stage and end to end (bit-exact)."""
import numpy as np
import pytest

import oracle
import riv2_amd
from riv2_amd import synth

pytestmark = pytest.mark.gpu


def _cfg(w, h, sr=12, q_intra=85, q_inter=80, skip=3):
    return riv2_amd.EncoderConfig(w, h, 30).with_intra_quality(q_intra).with_inter_quality(q_inter) \
        .with_search_range(sr).with_skip_threshold(skip).with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(1000))


def _run_pair(frames, w, h, sr=12, q_intra=85, q_inter=80, skip=3, label="", mv_extent=None):
    enc = riv2_amd.Encoder(_cfg(w, h, sr, q_intra, q_inter, skip))
    ref = oracle.OracleEncoder(oracle.config(w, h, intra_quality=q_intra, inter_quality=q_inter, search_range=sr, skip_threshold=skip))
    for i, f in enumerate(frames):
        print(f"[{label}] frame {i}", flush=True)
        st = enc.encode_frame_with_stats(riv2_amd.Frame(f, w, h, i * 33))
        ref.encode(f, i * 33)
        for k, name in ((0, "cur"),) + (((1, "pred"),) if i > 0 else ()) + ((2, "recon"),):
            g, o = enc.debug_yuv(k), ref.last_yuv(k)
            for pl, (a, b) in enumerate(zip(g, o)):
                assert np.array_equal(a, b), f"[{label}] frame {i}: {name} plane {pl} differs at {np.argwhere(a != b)[:5].tolist()}"
        if i > 0:
            grec, gsad, grdo, gopt = enc.debug_blocks()
            orec, osad, ordo, oopt = ref.last_blocks()
            if mv_extent is not None:
                mv_extent.append(int(np.abs(orec[:, :2].astype(np.int8).astype(int)).max()))
            bad = np.argwhere((grec[:, :4] != orec[:, :4]).any(1))
            assert bad.size == 0, f"[{label}] frame {i}: search MV differs at blocks {bad[:5].ravel().tolist()} gpu={grec[bad[:3].ravel(), :4].tolist()} ref={orec[bad[:3].ravel(), :4].tolist()}"
            assert np.array_equal(gsad, osad), f"[{label}] frame {i}: best SAD differs"
            assert np.array_equal(grec, orec), f"[{label}] frame {i}: mode decision differs at {np.argwhere((grec != orec).any(1))[:5].ravel().tolist()}"
            assert np.array_equal(grdo, ordo), f"[{label}] frame {i}: RDO skip differs"
            assert np.array_equal(gopt, oopt), f"[{label}] frame {i}: optimised skip differs"
            ga, oa = enc.debug_aq(), ref.last_aq()
            assert np.array_equal(ga.view(np.uint32), oa.view(np.uint32)), f"[{label}] frame {i}: AQ steps differ"
        assert np.array_equal(enc.debug_coeffs(), ref.last_coeffs()), f"[{label}] frame {i}: coefficients differ"
        assert st.total_bytes == len(ref.last_frame_bytes()), f"[{label}] frame {i}: frame size differs"
    enc.finish()
    gb, ob = enc.output(), ref.finish()
    assert gb == ob, f"[{label}] bitstream differs"
    return gb


def _decode_pair(data, n, label=""):
    dec, rdec = riv2_amd.Decoder(data), oracle.OracleDecoder(data)
    for i in range(n):
        print(f"[{label}] decode {i}", flush=True)
        a = dec.decode_frame()
        b, info = rdec.decode()
        assert a.frame_index == i and a.frame_type == info.frame_type
        assert np.array_equal(a.data, b), f"[{label}] decoded frame {i} differs"
        gy, oy = dec.debug_yuv(), rdec.last_yuv()
        for p, q in zip(gy, oy):
            assert np.array_equal(p, q)
    assert not dec.has_more_frames()


def test_cif_sequence_bit_exact():
    frames = synth.Sequence(synth.cif_params(seed=1, frames=8)).frames()
    data = _run_pair(frames, 352, 288, label="cif")
    _decode_pair(data, len(frames), label="cif")


def test_padded_small_frame_bit_exact():
    # 200x120 -> storage 208x128: right/bottom padding, edge MBs on every side
    frames = synth.Sequence(synth.custom_params(200, 120, 5, seed=3, global_motion=4.0, n_rects=3, rect_min=16, rect_max=48, rect_velocity=7.0)).frames()
    data = _run_pair(frames, 200, 120, label="small")
    _decode_pair(data, len(frames), label="small")


def test_search_range_and_quality_variants():
    frames = synth.Sequence(synth.custom_params(320, 176, 4, seed=5, global_motion=6.0, n_rects=4, rect_min=24, rect_max=64, rect_velocity=9.0)).frames()
    for sr, qi, qp, sk in ((4, 60, 40, 3), (24, 95, 90, 0), (16, 30, 10, 12)):
        data = _run_pair(frames, 320, 176, sr=sr, q_intra=qi, q_inter=qp, skip=sk, label=f"sr{sr}")
        _decode_pair(data, len(frames), label=f"sr{sr}")


def test_search_ranges_beyond_32():
    """Search ranges 33-63 (the reference allows up to 127; this backend up to 63): content moving up
    to 40 px per frame, so the long vectors occur; every stage equals the oracle's."""
    frames = synth.Sequence(synth.custom_params(320, 176, 4, seed=8, global_motion=24.0, n_rects=4, rect_min=24, rect_max=64,
                                                rect_velocity=40.0)).frames()
    for sr in (40, 63):
        ext = []
        data = _run_pair(frames, 320, 176, sr=sr, label=f"sr{sr}", mv_extent=ext)
        _decode_pair(data, len(frames), label=f"sr{sr}")
        assert max(ext) > 32, f"sr{sr}: no vector beyond 32 px ({ext})"


def test_pipelined_encoder_same_bitstream():
    frames = synth.Sequence(synth.cif_params(seed=11, frames=12)).frames()
    outs = []
    for depth in (0, 1, 3, 8):
        cfg = _cfg(352, 288).with_pipeline_depth(depth)
        if depth == 0:
            cfg = cfg.with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(5))
        else:
            cfg = cfg.with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(5))
        enc = riv2_amd.Encoder(cfg)
        for i, f in enumerate(frames):
            print(f"[pipe{depth}] frame {i}", flush=True)
            enc.encode_frame(riv2_amd.Frame(f, 352, 288, i))
        enc.finish()
        stats = enc.pop_stats()
        if depth:
            assert [s.frame_index for s in stats] == list(range(len(frames)))
        outs.append(enc.output())
        enc.close()
    ref = oracle.OracleEncoder(oracle.config(352, 288, keyframe_interval=5))
    for i, f in enumerate(frames):
        ref.encode(f, i)
    want = ref.finish()
    for o in outs:
        assert o == want
