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
