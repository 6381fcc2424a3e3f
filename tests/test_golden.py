"""The oracle reproduces the committed golden fixture of the CIF reference case (a regression pin:
any change to the oracle's output on this workload shows up here). CPU only."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402


def test_oracle_reproduces_cif_fixture():
    with open(make_golden.PATH) as f:
        want = json.load(f)
    got = json.loads(json.dumps(make_golden.compute()))
    assert got["workload"] == want["workload"]
    for a, b in zip(got["frames"], want["frames"]):
        assert a == b, f"frame {b['index']}"
    assert got["frame1_mbs"] == want["frame1_mbs"]
    assert (got["stream_bytes"], got["stream_sha256"]) == (want["stream_bytes"], want["stream_sha256"])
    assert got["decoded_rgb_sha256"] == want["decoded_rgb_sha256"]


def test_oracle_reproduces_hd_fixture_prefix():
    """The 1080p fixture's first frames (any prefix of the sequence encodes to the same records)."""
    import numpy as np
    import oracle
    from riv2_amd import synth
    with open(make_golden.HD_PATH) as f:
        want = json.load(f)
    k = 3
    frames = synth.Sequence(synth.hd_params(seed=make_golden.HD_SEED, frames=make_golden.HD_N)).frames(0, k)
    enc = oracle.OracleEncoder(oracle.config(make_golden.HD_W, make_golden.HD_H, fps=50))
    for i in range(k):
        enc.encode(frames[i], i * 20)
        fw = want["frames"][i]
        fb = enc.last_frame_bytes()
        assert (len(fb), make_golden.sha(fb)) == (fw["frame_bytes"], fw["frame_sha256"]), f"frame {i}"
        assert make_golden.sha(*enc.last_yuv(2)) == fw["recon_sha256"], f"frame {i} reconstruction"
    assert np.all([fr["index"] == i for i, fr in enumerate(want["frames"])]) and len(want["frames"]) == make_golden.HD_N


def test_oracle_reproduces_batch_fixture_streams():
    """Two of the batched fixture's 64 streams (the first CIF and 640x360 ones), in full."""
    with open(make_golden.BATCH_PATH) as f:
        want = json.load(f)
    jobs = make_golden.batch_streams()
    assert len(want["streams"]) == len(jobs) == 64
    for k in (0, 16):
        assert make_golden._batch_one(jobs[k]) == want["streams"][k], f"stream {k}"
