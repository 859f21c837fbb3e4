"""Host logic of the library built for the CPU with g++ from the same headers
the library compiles: the server's grid-mode routing policy
(ddp_amd/csrc/serve_route.h), its rules and its behaviour under load against a
queueing model of the server, and DA*'s insert grouping
(ddp_amd/csrc/host_sort.h)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def build_and_run(tmp_path, name, sanitize):
    exe = tmp_path / name
    flags = ["-fsanitize=address,undefined", "-fno-omit-frame-pointer"] if sanitize else []
    subprocess.run(["g++", "-O1", "-g", "-std=c++17", *flags, "-I", os.path.join(ROOT, "ddp_amd", "csrc"),
                    os.path.join(ROOT, "tests", name + ".cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "failures 0" in out.stdout, out.stdout
    return out.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_serve_route_policy(tmp_path):
    build_and_run(tmp_path, "route_check", sanitize=False)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_serve_route_under_load(tmp_path):
    # the policy against a queueing model of the server (Poisson arrivals,
    # stopped launches, relaunches behind the grid queue)
    build_and_run(tmp_path, "route_sim", sanitize=False)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_insert_grouping_radix(tmp_path):
    # under the address and undefined-behaviour sanitizers: the check also
    # guards the buffer bounds of the last pass
    build_and_run(tmp_path, "radix_check", sanitize=True)
