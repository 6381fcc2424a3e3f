"""Host-side pieces of the learned form and the stress generator, without a GPU: the Bjontegaard
rate difference against closed forms, proposal unpacking, the network's parameter count through
the C ABI, and the generator's zoom/rotation options (deterministic; off leaves the earlier
sequences untouched)."""
import hashlib

import numpy as np
import pytest

from riv2_amd import _lib as L
from riv2_amd import synth

learned = pytest.importorskip("riv2_amd.learned")


def test_bd_rate_closed_forms():
    psnr = [30.0, 31.0, 32.5, 34.0]
    rate = [1e5, 1.4e5, 2.1e5, 3.3e5]
    assert abs(learned.bd_rate(rate, psnr, rate, psnr)) < 1e-9
    # every rate scaled by k at the same PSNR: exactly (k - 1) * 100 percent
    for k in (0.9, 1.01, 1.25):
        assert abs(learned.bd_rate(rate, psnr, [r * k for r in rate], psnr) - (k - 1) * 100) < 1e-6
    # the same curve shifted up in PSNR costs less rate: negative difference
    assert learned.bd_rate(rate, psnr, rate, [p + 0.5 for p in psnr]) < 0


def test_unpack_proposals():
    cands = np.array([[(3, -2), (0, 0), (-12, 12), (5, 1)], [(-1, -1), (2, 7), (0, -3), (12, -12)]], np.int8)
    words = np.zeros((2, 4), np.uint32)
    words[:, :2] = cands.reshape(2, 8).view(np.uint8).view(np.uint32).reshape(2, 2)
    words[:, 2] = [1, 0]
    words[:, 3] = np.array([0.75, -2.5], np.float32).view(np.uint32)
    c, conf, logit = learned.unpack_proposals(words)
    assert np.array_equal(c, cands.astype(np.int32))
    assert conf.tolist() == [True, False]
    assert logit.tolist() == [0.75, -2.5]


def test_param_count_matches_the_layer_layout():
    # conv1 48x16 + 16, conv2 / conv3 144x16 + 16, fc 272x128 + 128, head 128x16 + 16
    want = 48 * 16 + 16 + 2 * (144 * 16 + 16) + 272 * 128 + 128 + 128 * 16 + 16
    assert L.lib().riv_mvnet_param_count() == want == 42432


def test_stress_content_is_deterministic_and_moves():
    a = synth.Sequence(synth.stress_params(160, 96, 27, seed=9)).frames()
    b = synth.Sequence(synth.stress_params(160, 96, 27, seed=9)).frames()
    assert np.array_equal(a, b)
    assert np.abs(a[1].astype(int) - a[0].astype(int)).mean() > 5


def test_zoom_and_rotation_off_keep_earlier_sequences():
    p0 = synth.custom_params(128, 96, 4, seed=3)
    p1 = synth.custom_params(128, 96, 4, seed=3, rect_zoom=0.0, rect_rotate_deg=0.0)
    f0, f1 = synth.Sequence(p0).frames(), synth.Sequence(p1).frames()
    assert np.array_equal(f0, f1)
    # turned on, the rectangles render through the rotating path (a different sequence), still
    # a function of the parameters alone
    p2 = synth.custom_params(128, 96, 4, seed=3, rect_zoom=0.005, rect_rotate_deg=1.0)
    f2 = synth.Sequence(p2).frames()
    assert not np.array_equal(f0[3], f2[3])
    assert hashlib.md5(f2.tobytes()).hexdigest() == hashlib.md5(synth.Sequence(p2).frames().tobytes()).hexdigest()
