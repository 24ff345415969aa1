"""CPU unit test of the general kernels' scratch sizing (com-rs_amd/csrc/sizing.hpp, the
loop that gives a launch fewer lanes when the device refuses the whole request): builds
tests/sizing_check.cpp with g++ under the address and undefined-behaviour sanitizers and
runs it."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_scratch_sizing_on_the_host(tmp_path):
    exe = str(tmp_path / "sizing_check")
    b = subprocess.run(["g++", "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=undefined", "-o", exe, os.path.join(HERE, "sizing_check.cpp")],
                       capture_output=True, text=True, timeout=300)
    assert b.returncode == 0, b.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr[-4000:]
    assert "FAIL" not in r.stdout and "0 failure(s)" in r.stdout
