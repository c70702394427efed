"""The engine's host-only code (tap-table builders, Java float text, kernel constants,
rolling/load plan) built with g++ under AddressSanitizer + UndefinedBehaviorSanitizer and
libstdc++ bounds assertions, and run over wide parameter grids (tests/sanitize/host_check.cpp
lists what it checks besides running clean).  CPU only."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
EXE = os.path.join(REPO, "_build", "sanitize", "host_check")


@pytest.mark.skipif(shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"),
                    reason="needs g++ and the HIP headers")
def test_host_code_under_asan_ubsan():
    b = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "sanitize")],
                       capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "host_check: ok" in r.stdout
    assert "runtime error" not in out and "AddressSanitizer" not in out
