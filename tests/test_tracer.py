"""The RHS tracer: the HIP C it emits computes exactly what the Python function
computes in binary64 (checked by compiling the emitted code for the CPU with
the same no-contraction flags), and it rejects what it cannot trace."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from simplediffeq_hip import tracer

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def lorenz(u, p, t):
    s, r, b = p[0], p[1], p[2]
    return [s * (u[1] - u[0]), u[0] * (r - u[2]) - u[1], u[0] * u[1] - b * u[2]]


def forced(u, p, t):
    # non-autonomous, transcendental, abs, constants (tracer.cos etc. are the
    # C library functions on plain floats)
    return [u[1], -p[0] * u[0] - p[1] * u[1] + p[2] * tracer.cos(t) - 0.25 * (u[0] * u[0] * u[0]),
            tracer.exp(-abs(u[2])) / 3.0 + tracer.sqrt(u[0] * u[0] + 1.0)]


A = np.random.default_rng(5).normal(0, 0.1, (4, 4))


def glv4(u, p, t):
    return [u[i] * (p[i] + sum(A[i, j] * u[j] for j in range(4))) for i in range(4)]


HARNESS = r'''
#include <cmath>
#include <cstdio>
#define __device__
using std::sin; using std::cos; using std::exp; using std::log; using std::sqrt;
using std::fabs; using std::tanh; using std::tan; using std::sinh; using std::cosh; using std::atan;
%s
int main() {
  double u[%d], p[%d], du[%d]; double t;
  while (scanf("%%la", &t) == 1) {
    for (int i = 0; i < %d; ++i) scanf("%%la", &u[i]);
    for (int i = 0; i < %d; ++i) scanf("%%la", &p[i]);
    %s<double>(du, u, p, nullptr, t);
    for (int i = 0; i < %d; ++i) printf("%%a ", du[i]);
    printf("\n");
  }
}
'''


@pytest.mark.parametrize("f,n,np_", [(lorenz, 3, 3), (forced, 3, 3), (glv4, 4, 4)])
def test_emitted_code_matches_python_bitwise(f, n, np_):
    src, fn = tracer.trace(f, n, np_, "traced_fn")
    rng = np.random.default_rng(0)
    cases = [(float(rng.uniform(0, 5)), rng.uniform(-2, 2, n), rng.uniform(-2, 2, np_))
             for _ in range(50)]
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "h.cpp")
        open(c, "w").write(HARNESS % (src, n, max(np_, 1), n, n, np_, fn, n))
        exe = os.path.join(d, "h")
        subprocess.run(["g++", "-O2", "-ffp-contract=off", "-std=c++17", c, "-o", exe],
                       check=True)
        inp = "\n".join(" ".join(x.hex() for x in [t] + list(u) + list(p)) for t, u, p in cases)
        out = subprocess.run([exe], input=inp, capture_output=True, text=True,
                             check=True).stdout.strip().splitlines()
    for (t, u, p), line in zip(cases, out):
        got = [float.fromhex(x) for x in line.split()]
        want = [float(v) for v in f(list(u), list(p), t)]
        assert got == want


def test_lorenz_trace_fixture():
    src, _ = tracer.trace(lorenz, 3, 3, "lorenz_traced")
    fx = os.path.join(ROOT, "tests", "golden", "lorenz_traced.hip")
    assert src == open(fx).read()


def test_integer_power_is_repeated_multiplication():
    # Julia's literal x^3 is x*x*x (Base.literal_pow), two roundings
    src, _ = tracer.trace(lambda u, p, t: [u[0] ** 3], 1, 0)
    assert "const T v0 = u[0] * u[0];" in src and "const T v1 = v0 * u[0];" in src


def test_untraceable_functions_raise():
    with pytest.raises(tracer.TraceError):
        tracer.trace(lambda u, p, t: [u[0] if u[0] > 0 else -u[0]], 1, 0)
    with pytest.raises(tracer.TraceError):
        tracer.trace(lambda u, p, t: [u[0], u[1]], 3, 0)
    with pytest.raises(IndexError):
        tracer.trace(lambda u, p, t: [p[3]], 1, 2)
    with pytest.raises(tracer.TraceError):
        tracer.trace(lambda u, p, t: [u[0] ** 0.5], 1, 0)


def test_scalar_state_and_parameter():
    """A scalar problem (u0 and p numbers, as the reference's Float64-scalar tests)
    traces to the same code as its one-component vector form."""
    src_s, _ = tracer.trace(lambda u, p, t: p * u + t, 1, 1, "g", scalar_u=True, scalar_p=True)
    src_v, _ = tracer.trace(lambda u, p, t: [p[0] * u[0] + t], 1, 1, "g")
    assert src_s == src_v
    with pytest.raises(tracer.TraceError):
        tracer.trace(lambda u, p, t: u, 2, 0, "g", scalar_u=True)
