"""The C-ABI library loads without a GPU and exports every function include/*.h declares."""
import ctypes
import os
import re

import riv2_amd
from riv2_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(inc, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(riv_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = declared_functions()
    assert len(names) > 30
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, f"declared but not exported: {missing}"


def test_ctypes_prototypes_cover_the_header():
    bound = {n for n, _, _ in _lib.PROTOTYPES}
    assert declared_functions() <= bound | {"riv_host_placeholder"}


def test_config_defaults_and_validation_on_cpu():
    # pure host functions: EncoderConfig::new defaults (config.rs:56-77) and validate (:129-164)
    c = riv2_amd.EncoderConfig(1920, 1080, 30)
    assert (c.storage_width, c.storage_height) == (1920, 1088)
    assert (c.intra_quality, c.inter_quality, c.search_range, c.skip_threshold) == (90, 35, 12, 3)
    assert abs(c.rdo_lambda_mult - 0.49) < 1e-12 and abs(c.inter_dead_zone - 0.75) < 1e-7
    assert c.keyframe_interval.mode == 0 and c.keyframe_interval.interval == 30
    c.validate()
    import pytest
    for field, value in (("search_range", 128), ("mv_deflate_level", 11), ("fps", 0)):
        bad = riv2_amd.EncoderConfig(64, 64, 30)
        setattr(bad, field, value)
        with pytest.raises(riv2_amd.EncodeError) as ei:
            bad.validate()
        assert ei.value.kind == "InvalidConfig"
    bad = riv2_amd.EncoderConfig(64, 64, 30).with_keyframe_interval(riv2_amd.KeyframeInterval.fixed(0))
    with pytest.raises(riv2_amd.EncodeError):
        bad.validate()
