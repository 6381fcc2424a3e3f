"""The host's fast and parallel coding forms against their plain serial forms (CPU only): builds
host_coder_check.cpp with the encoder's host coder (riv2_amd/csrc/entropy.cpp) and runs it.

* code_symbol_stream in groups of 16 and of 8 blocks, by tasks, against one serial writer
  (rANS blocks of 16,386 symbols, each from fresh states: rans.rs:201-269);
* RansWriter::flush_parallel, one task per block, against flush();
* MvCoder::contexts, one table word per decision, against RansWriter::bit on Ctx states
  (BinProb, rans.rs:30-84), with skewed contexts whose counters saturate."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIP_INC = "/opt/rocm/include"


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.exists(os.path.join(HIP_INC, "hip", "hip_runtime.h")),
                    reason="needs g++ and the HIP headers (for the shared type definitions)")
def test_host_coder_forms_give_the_serial_bytes(tmp_path):
    exe = str(tmp_path / "host_coder_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I" + HIP_INC, "-D__HIP_PLATFORM_AMD__", "-o", exe,
                    os.path.join(HERE, "host_coder_check.cpp"), os.path.join(ROOT, "riv2_amd", "csrc", "entropy.cpp"), "-lpthread"],
                   check=True, capture_output=True, timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "host coder checks passed" in r.stdout
