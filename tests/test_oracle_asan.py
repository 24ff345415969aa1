"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: `make -C oracle asan`
builds the oracle's sources with both sanitizers and leak detection into a driver that calls
every entry point of oracle.h (NR codes with filler bits, repetition and every redundancy
version up to BG1 Z=384, serial and threaded decoders in f64, FP32 and FP16, the tracer, the
learned decoder and its gradient, the link and the statistics) and checks that the run ends
clean."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(__file__), "..", "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_clean_under_asan_and_ubsan():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "asan"], capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "asan_check: ok" in out
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
