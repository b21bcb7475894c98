"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: a test program
(oracle/sanitize_check.cpp) solves every solver x state type x built-in right-hand
side x output mode (endpoint, saveat in both layouts, every step with room and
with overflow, the edge cases: NaN, dt below dtmin, maxiters, empty and reversed
spans) with one and several threads; any invalid access, leak or undefined
arithmetic conversion aborts it."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host compiler")
def test_oracle_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(ROOT, "oracle", "build", "sanitize_check")],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
    assert "0 failures" in r.stdout
