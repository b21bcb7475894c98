"""Shared helpers of the test suite: run one ensemble through the HIP library
and through the CPU oracle on the same bytes, and compare."""
import numpy as np

from simplediffeq_hip import _abi, inputs, runtime

ALG_NAMES = {0: "Tsit5", 1: "ATsit5", 2: "Vern7", 3: "AVern7", 4: "Vern9", 5: "AVern9"}


def problem(rhs, n, seed=0):
    """Inputs (u0, p, p_shared) for a built-in RHS."""
    rng = np.random.default_rng(seed)
    if rhs == _abi.RHS_LORENZ:
        u0, p = inputs.lorenz_small(n, seed=seed)
        return u0, p, None
    if rhs == _abi.RHS_EXPDECAY3:
        u0 = rng.uniform(0.5, 2.0, (3, n))
        p = rng.uniform(-2.0, 0.5, (1, n))
        return u0, p, None
    if rhs == _abi.RHS_AFFINE1:
        u0 = rng.uniform(-1.0, 1.0, (1, n))
        p = np.stack([rng.uniform(-2.0, 0.2, n), rng.uniform(-1.0, 1.0, n)])
        return u0, p, None
    if rhs == _abi.RHS_LINEAR2:
        u0 = rng.uniform(-1.0, 1.0, (2, n))
        p = rng.uniform(-1.0, 0.6, (4, n))
        return u0, p, None
    if rhs == _abi.RHS_GLV16:
        u0, r, A = inputs.glv16(n, seed=seed)
        return u0, r, A
    raise ValueError(rhs)


def run_pair(ctx, desc, u0, p, p_shared=None, tspan=None, dt0=None, order=None, arith=0,
             pow_kind=0, backend="numpy"):
    """Solve on the GPU (via the C ABI) and with the oracle; return both buffers."""
    from oracle import pyoracle
    n_state = u0.shape[0]
    g = runtime.EnsembleBuffers(desc, n_state, u0, p, p_shared, tspan, dt0, order,
                                backend=backend, device="cuda:0" if backend == "torch" else None)
    st_g = ctx.solve(desc, g, allow=(_abi.E_DTMIN, _abi.E_SAVE_OVERFLOW))
    c = runtime.EnsembleBuffers(desc, n_state, u0, p, p_shared, tspan, dt0, None)
    st_c = pyoracle.solve(desc, c, arith=arith, pow_kind=pow_kind, nthreads=8)
    return st_g, g, st_c, c


def bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype == np.float64:
        return np.array_equal(a.view(np.int64), b.view(np.int64))
    if a.dtype == np.float32:
        return np.array_equal(a.view(np.int32), b.view(np.int32))
    return np.array_equal(a, b)


def assert_bit_identical(g, c, fields=("u_end", "t_end", "retcode", "naccept", "nreject")):
    for f in fields:
        a, b = g.numpy(f), c.numpy(f)
        if a is None and b is None:
            continue
        if not bits_equal(a, b):
            a = np.asarray(a)
            b = np.asarray(b)
            bad = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b)) if a.dtype.kind == "f"
                                else (a == b)))
            raise AssertionError("%s differs at %d places, first %s: gpu %r oracle %r" % (
                f, len(bad), bad[:3].tolist(), a[tuple(bad[0])], b[tuple(bad[0])]))
