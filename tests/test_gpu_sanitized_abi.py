"""The library's host code under AddressSanitizer and UBSan with the kernels on the
GPU: simplediffeq.jl_amd/csrc/check/abi_check.cpp, linked against the
host-sanitized build (make -C simplediffeq.jl_amd/csrc asan), drives every C-ABI
entry point (host and device solves, the chunked pipeline, the staged schedule,
the allocation-free solve, the packed capacity protocol, graphs, hipRTC, the
cost-model kernels, the plan order, the autotuner) and the refusal paths, and
checks that the solves agree bit for bit."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "simplediffeq.jl_amd", "csrc", "build_asan", "abi_check")


def test_c_abi_under_host_sanitizers():
    if not os.path.exists(DRIVER):
        pytest.skip("host-sanitized build not present (make -C simplediffeq.jl_amd/csrc asan)")
    env = dict(os.environ, ASAN_OPTIONS="protect_shadow_gap=0:detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    print("running", DRIVER, flush=True)
    r = subprocess.run([DRIVER], capture_output=True, text=True, env=env, timeout=240)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "abi_check: 0 failures" in r.stdout
    assert "runtime error" not in out and "ERROR: AddressSanitizer" not in out
