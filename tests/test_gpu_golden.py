"""The GPU encoder and decoder reproduce the committed golden fixture of the CIF reference case
(352x288, 30 frames, q85/80, seed 1) without running the oracle: per-frame digests of records,
SADs, skip flags, AQ steps, coefficients and reconstructions for frames 0-5, the stream digest and
the digest of the decoded RGB frames."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

import riv2_amd
from riv2_amd import synth

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "cif_config1.json")


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(a if isinstance(a, (bytes, bytearray)) else np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_gpu_matches_cif_fixture():
    with open(FIXTURE) as f:
        want = json.load(f)
    wl = want["workload"]
    w, h, n = wl["width"], wl["height"], wl["frames"]
    frames = synth.Sequence(synth.cif_params(seed=wl["seed"], frames=n)).frames(0, n)
    cfg = riv2_amd.EncoderConfig(w, h, 30).with_intra_quality(wl["intra_quality"]).with_inter_quality(wl["inter_quality"]) \
        .with_search_range(wl["search_range"]).with_skip_threshold(wl["skip_threshold"]) \
        .with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(wl["keyframe_interval"])).with_pipeline_depth(0)
    enc = riv2_amd.Encoder(cfg)
    frame_want = {fr["index"]: fr for fr in want["frames"]}
    for i in range(n):
        print(f"[golden] frame {i}", flush=True)
        st = enc.encode_frame_with_stats(riv2_amd.Frame(frames[i], w, h, i * 33))
        if i not in frame_want:
            continue
        fw = frame_want[i]
        assert st.total_bytes == fw["frame_bytes"], f"frame {i} size"
        assert sha(*enc.debug_yuv(2)) == fw["recon_sha256"], f"frame {i} reconstruction"
        assert sha(enc.debug_coeffs()) == fw["coeffs_sha256"], f"frame {i} coefficients"
        if i > 0:
            rec, sad, rdo, opt = enc.debug_blocks()
            assert sha(rec) == fw["records_sha256"], f"frame {i} MB records"
            assert sha(sad.astype(np.int64)) == fw["sads_sha256"], f"frame {i} SADs"
            assert sha(rdo) == fw["rdo_skip_sha256"] and sha(opt) == fw["opt_skip_sha256"], f"frame {i} skip flags"
            assert sha(enc.debug_aq()) == fw["aq_sha256"], f"frame {i} AQ steps"
            if i == 1:
                for mb in want["frame1_mbs"]:
                    b = mb["mb"]
                    assert [int(x) for x in rec[b]] == mb["record"] and int(sad[b]) == mb["sad"], f"MB {b}"
    enc.finish()
    data = enc.output()
    assert len(data) == want["stream_bytes"] and sha(data) == want["stream_sha256"], "stream"
    dec = riv2_amd.Decoder(data)
    hh = hashlib.sha256()
    for i in range(n):
        hh.update(np.ascontiguousarray(dec.decode_frame().data).tobytes())
    assert hh.hexdigest() == want["decoded_rgb_sha256"], "decoded RGB"
