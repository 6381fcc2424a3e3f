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


def test_ctypes_structs_match_the_c_header(tmp_path):
    """Every ctypes mirror of a header struct has the C compiler's size and field offsets
    (gcc on a probe that includes include/riv_gpu.h)."""
    import subprocess
    import shutil
    import pytest
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = {"riv_encoder_config": _lib.EncoderConfigC, "riv_video_header": _lib.VideoHeaderC,
               "riv_frame_stats": _lib.FrameStatsC, "riv_decoded_info": _lib.DecodedInfoC,
               "riv_mv_source": _lib.MvSourceC, "riv_writer_vtbl": _lib.WriterVtblC, "riv_reader_vtbl": _lib.ReaderVtblC,
               "riv_decode_timings": _lib.DecodeTimingsC}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "riv_gpu.h"', 'int main(void) {']
    for cname, py in structs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines.append("return 0; }")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        s, f, v = line.split()
        got[(s, f)] = int(v)
    for cname, py in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, f"{cname}.{f[0]}"


def test_stage_names_cover_the_header_stages():
    text = open(os.path.join(ROOT, "include", "riv_gpu.h")).read()
    n = int(re.search(r"#define RIV_NUM_STAGES (\d+)", text).group(1))
    assert len(riv2_amd.Encoder.STAGES) == n
