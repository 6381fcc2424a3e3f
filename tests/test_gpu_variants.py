"""Configuration paths of the reference encoder beyond the default settings, each bit-exact against
the oracle stage by stage and end to end:

* the motion search's early exits: me_zero_mv_threshold (skip the integer search when the zero
  vector's SAD is within it) and me_predictor_threshold (skip it when the best seed's SAD is)
  (reitero_encode/src/motion/scalar.rs:588-606);
* rdo_lambda_mult 0 (every MV cost zero, estimate_mv_cost's early return, scalar.rs:453-544) and
  1.5 (rdo.rs:82-86);
* inter_dead_zone 0.5 and 1.0 (the quantiser's dead zone, reitero_dct/src/dct16.rs:321-333);
* KeyframeInterval::AllIntra and ::Automatic (an I-frame every 30 frames), encoder.rs:265-291."""
import numpy as np
import pytest

import oracle
import riv2_amd
from riv2_amd import synth

pytestmark = pytest.mark.gpu

W, H = 352, 288


def _frames(n, seed=9, w=W, h=H, motion=5.0):
    return synth.Sequence(synth.custom_params(w, h, n, seed=seed, global_motion=motion, n_rects=4, rect_min=24, rect_max=80,
                                              rect_velocity=8.0)).frames()


def _run(frames, label, w=W, h=H, zthr=0, pthr=0, lam=0.49, dz=0.75, kf_mode=0, kf=1000, check_stages=True):
    kfi = {0: riv2_amd.KeyframeInterval.fixed(kf), 1: riv2_amd.KeyframeInterval.all_intra(),
           2: riv2_amd.KeyframeInterval.automatic()}[kf_mode]
    cfg = riv2_amd.EncoderConfig(w, h, 30).with_intra_quality(85).with_inter_quality(80).with_search_range(12) \
        .with_skip_threshold(3).with_keyframe_interval(kfi).with_me_zero_mv_threshold(zthr) \
        .with_me_predictor_threshold(pthr).with_rdo_lambda_mult(lam).with_inter_dead_zone(dz)
    enc = riv2_amd.Encoder(cfg)
    ref = oracle.OracleEncoder(oracle.config(w, h, zero_thr=zthr, pred_thr=pthr, lambda_mult=lam, dead_zone=dz,
                                             keyframe_mode=kf_mode, keyframe_interval=kf))
    types = []
    for i, f in enumerate(frames):
        print(f"[{label}] frame {i}", flush=True)
        st = enc.encode_frame_with_stats(riv2_amd.Frame(f, w, h, i * 33))
        ost = ref.encode(f, i * 33)
        types.append(st.frame_type)
        assert st.frame_type == ost.frame_type, f"[{label}] frame {i}: frame type"
        if check_stages and st.frame_type == 2:
            grec, gsad, grdo, gopt = enc.debug_blocks()
            orec, osad, ordo, oopt = ref.last_blocks()
            bad = np.argwhere((grec != orec).any(1)).ravel()
            assert bad.size == 0, f"[{label}] frame {i}: records differ at {bad[:5].tolist()}: {grec[bad[:2]].tolist()} vs {orec[bad[:2]].tolist()}"
            assert np.array_equal(gsad, osad), f"[{label}] frame {i}: SADs"
            assert np.array_equal(grdo, ordo) and np.array_equal(gopt, oopt), f"[{label}] frame {i}: skip flags"
            assert np.array_equal(enc.debug_aq().view(np.uint32), ref.last_aq().view(np.uint32)), f"[{label}] frame {i}: AQ"
        if check_stages:
            assert np.array_equal(enc.debug_coeffs(), ref.last_coeffs()), f"[{label}] frame {i}: coefficients"
            for a, b in zip(enc.debug_yuv(2), ref.last_yuv(2)):
                assert np.array_equal(a, b), f"[{label}] frame {i}: reconstruction"
    enc.finish()
    data = enc.output()
    assert data == ref.finish(), f"[{label}] stream"
    dec, rdec = riv2_amd.Decoder(data), oracle.OracleDecoder(data)
    for i in range(len(frames)):
        assert np.array_equal(dec.decode_frame().data, rdec.decode()[0]), f"[{label}] decoded frame {i}"
    dec.close()
    return types, enc


# thresholds that change 8-386 of the 396 blocks' vectors against the search without early exits
@pytest.mark.parametrize("zthr,pthr", [(2500, 0), (6000, 0), (0, 2500), (3000, 3000)])
def test_motion_search_early_exits(zthr, pthr):
    _run(_frames(5), f"thr {zthr}/{pthr}", zthr=zthr, pthr=pthr)
    # edge blocks on every side of a padded frame, where the early exits meet clamped windows
    _run(_frames(4, seed=12, w=200, h=120), f"thr {zthr}/{pthr} small", w=200, h=120, zthr=zthr, pthr=pthr)


@pytest.mark.parametrize("lam", [0.0, 1.5])
def test_lambda_multipliers(lam):
    _run(_frames(5), f"lambda x{lam}", lam=lam)


@pytest.mark.parametrize("dz", [0.5, 1.0])
def test_dead_zones(dz):
    _run(_frames(5), f"dead zone {dz}", dz=dz)


def test_all_intra():
    types, _ = _run(_frames(4), "all intra", kf_mode=1)
    assert types == [1, 1, 1, 1]


def test_automatic_keyframes():
    frames = _frames(33, motion=3.0)
    types, _ = _run(frames, "automatic", kf_mode=2, check_stages=False)
    assert [i for i, t in enumerate(types) if t == 1] == [0, 30]
