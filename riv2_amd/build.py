"""Builds the in-tree shared libraries: riv2_amd/libriv_gpu.so (gfx950 kernels + C ABI),
riv2_amd/libriv_synth.so (sequence generator) and, when asked, oracle/liboracle_riv.so."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _make(directory: str, force: bool):
    jobs = str(min(16, os.cpu_count() or 1))
    if force:
        subprocess.run(["make", "-s", "-C", directory, "clean"], check=True)
    subprocess.run(["make", "-s", "-j", jobs, "-C", directory], check=True)


def build(force: bool = True, oracle: bool = True):
    _make(os.path.join(HERE, "csrc"), force)
    if oracle:
        _make(os.path.join(ROOT, "oracle"), force)


if __name__ == "__main__":
    build()
