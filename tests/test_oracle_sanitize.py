"""The C++ oracle under AddressSanitizer and UndefinedBehaviorSanitizer
(host code only): a small reach with bed, banks, a hull under the lid,
actuator-disk forces, wall-model coefficients, inlet and outlet, stepped in
the FP64, FP32 and mixed instantiations with the boundary rows and the uref
reduction (oracle/sanitize_check.cpp). Any out-of-bounds access or undefined
operation fails the run."""
import os
import shutil
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.skipif(shutil.which("make") is None or shutil.which("g++") is None, reason="needs make and g++")
def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True, timeout=300)
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ORACLE, "_build", "sanitize_check")], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "sanitize ok" in r.stdout
