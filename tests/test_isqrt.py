"""The Euclidean heuristic's integer square root (ddp_amd/csrc/isqrt.h),
built for the host with g++ and checked against r^2 <= x < (r+1)^2 in 128-bit
arithmetic; the device kernels compile the same header."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_isqrt_exact(tmp_path):
    exe = tmp_path / "isqrt_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "ddp_amd", "csrc"),
                    os.path.join(ROOT, "tests", "isqrt_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "failures 0" in out.stdout
