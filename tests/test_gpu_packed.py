"""Packed save_everystep output (no per-trajectory cap): every trajectory's rows
back to back at steps_offset[i] .. steps_offset[i+1]-1 (a counting pass, an
offsets scan, then the saving pass writing in place).  Checked bit for bit
against the oracle's per-trajectory rows, for every algorithm and both
schedules, and, when the rows do not fit, through the truncation rule (rows in
trajectory order until the capacity is full: earlier trajectories complete, one
partial, later ones empty; steps_needed names the capacity that fits)."""
import numpy as np
import pytest

from helpers import bits_equal, problem
from simplediffeq_hip import _abi, inputs, runtime

pytestmark = pytest.mark.gpu

ALL_ALGS = list(range(8))


def _desc(alg, n, rhs=_abi.RHS_LORENZ, layout=_abi.SAVE_PACKED, max_saves=0, **kw):
    adaptive = alg in _abi.ADAPTIVE_ALGS
    return runtime.make_desc(alg, rhs=rhs, n_traj=n, tspan=(0.0, 1.0),
                             dt=0.1 if adaptive else 0.05, abstol=1e-6, reltol=1e-4,
                             save_everystep=True, max_saves=max_saves, save_layout=layout, **kw)


def _oracle_rows(desc_kw, alg, n, u0, p, ps=None, dtype=_abi.F64):
    """The oracle's save_everystep rows per trajectory (trajectory-major, enough rows)."""
    from oracle import pyoracle
    d = _desc(alg, n, layout=_abi.SAVE_TRAJ_MAJOR, max_saves=600, dtype=dtype, **desc_kw)
    c = runtime.EnsembleBuffers(d, u0.shape[0], u0, p, ps)
    assert pyoracle.solve(d, c, nthreads=8) == 0  # no trajectory needs more than 600 rows
    return c


def _check_packed_equals(g, c, n):
    k = c.numpy("n_steps")
    assert bits_equal(g.numpy("n_steps"), k)
    off = g.numpy("steps_offset")
    assert off[0] == 0 and np.array_equal(np.diff(off), k)
    for i in range(n):
        t, u = g.steps_of(i)
        assert bits_equal(t, c.t_steps[i, :k[i]]), i
        assert bits_equal(u, c.u_steps[i, :k[i]]), i


@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT],
                         ids=["one-launch", "persistent"])
@pytest.mark.parametrize("alg", ALL_ALGS, ids=lambda a: ["tsit5", "atsit5", "vern7", "avern7",
                                                          "vern9", "avern9", "rk4", "euler"][a])
def test_packed_bitexact(ctx, alg, sched):
    if alg not in _abi.ADAPTIVE_ALGS and sched == _abi.SCHED_PERSISTENT:
        pytest.skip("fixed-step solvers always run one launch")
    n = 777
    u0, p, _ = problem(_abi.RHS_LORENZ, n, seed=50 + alg)
    d = _desc(alg, n, schedule=sched, chunk=5)
    g = runtime.EnsembleBuffers(d, 3, u0, p, steps_capacity=n * 80)
    assert ctx.solve(d, g) == 0
    c = _oracle_rows({"schedule": sched, "chunk": 5}, alg, n, u0, p)
    _check_packed_equals(g, c, n)
    for f in ("u_end", "t_end", "retcode", "naccept", "nreject"):
        assert bits_equal(g.numpy(f), c.numpy(f)), f
    assert ctx.stats()["steps_needed"] == int(c.numpy("n_steps").sum())


def test_packed_device_buffers_and_binary32(ctx):
    n = 1000
    u0, p = inputs.lorenz_wide(n, seed=3)
    d = _desc(_abi.ALG_ATSIT5, n)
    host = runtime.EnsembleBuffers(d, 3, u0, p, steps_capacity=n * 80)
    ctx.solve(d, host)
    dev = runtime.EnsembleBuffers(d, 3, u0, p, steps_capacity=n * 80, backend="torch",
                                  device="cuda:0")
    ctx.solve(d, dev)
    for f in ("n_steps", "steps_offset", "u_end", "retcode"):
        assert bits_equal(dev.numpy(f), host.numpy(f)), f
    used = int(host.steps_offset[n])
    assert bits_equal(dev.numpy("t_steps")[:used], host.t_steps[:used])
    assert bits_equal(dev.numpy("u_steps")[:used], host.u_steps[:used])
    # binary32: the packed rows are the trajectory-major rows of the same kernel
    d32 = _desc(_abi.ALG_ATSIT5, n, dtype=_abi.F32)
    pk = runtime.EnsembleBuffers(d32, 3, u0, p, steps_capacity=n * 80)
    ctx.solve(d32, pk)
    dt = _desc(_abi.ALG_ATSIT5, n, dtype=_abi.F32, layout=_abi.SAVE_TRAJ_MAJOR, max_saves=600)
    tm = runtime.EnsembleBuffers(dt, 3, u0, p)
    ctx.solve(dt, tm)
    assert pk.u_steps.dtype == np.float32
    _check_packed_equals(pk, tm, n)


@pytest.mark.parametrize("frac", [0.5, 0.0], ids=["half", "none"])
def test_packed_overflow_keeps_prefixes(ctx, frac):
    n = 600
    u0, p, _ = problem(_abi.RHS_LORENZ, n, seed=61)
    c = _oracle_rows({}, _abi.ALG_ATSIT5, n, u0, p)
    full = c.numpy("n_steps").astype(np.int64)
    cap = int(full.sum() * frac)
    d = _desc(_abi.ALG_ATSIT5, n)
    g = runtime.EnsembleBuffers(d, 3, u0, p, steps_capacity=cap)
    assert ctx.solve(d, g, allow=(_abi.E_SAVE_OVERFLOW,)) == _abi.E_SAVE_OVERFLOW
    assert ctx.stats()["steps_needed"] == int(full.sum())
    k = g.numpy("n_steps").astype(np.int64)
    off = g.numpy("steps_offset")
    assert off[0] == 0 and np.array_equal(np.diff(off), k) and off[n] == cap
    cut = int(np.argmax(k < full))  # the first trajectory that lost rows
    assert (k[:cut] == full[:cut]).all() and (k[cut + 1:] == 0).all()
    assert off[cut] + full[cut] > cap
    ret = g.numpy("retcode")
    assert ((ret == _abi.RET_SAVE_OVERFLOW) == (k < full)).all()
    assert ((ret == _abi.RET_DEFAULT) == (k == full)).all()
    for i in range(n):
        t, u = g.steps_of(i)
        assert bits_equal(t, c.t_steps[i, :k[i]]) and bits_equal(u, c.u_steps[i, :k[i]]), i
    # the integration itself is unaffected
    assert bits_equal(g.numpy("u_end"), c.numpy("u_end"))


def test_packed_validation_and_empty(ctx):
    u0, p, _ = problem(_abi.RHS_LORENZ, 4, seed=1)
    d = runtime.make_desc(_abi.ALG_ATSIT5, n_traj=4, saveat=np.array([0.0, 0.5]),
                          save_layout=_abi.SAVE_PACKED)
    with pytest.raises(runtime.SolveError, match="packed layout holds save_everystep"):
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p))
    d = runtime.make_desc(_abi.ALG_ATSIT5, n_traj=4, save_layout=_abi.SAVE_PACKED)
    with pytest.raises(runtime.SolveError, match="packed layout holds save_everystep"):
        ctx.solve(d, runtime.EnsembleBuffers(d, 3, u0, p))
    d = _desc(_abi.ALG_ATSIT5, 4)
    b = runtime.EnsembleBuffers(d, 3, u0, p, steps_capacity=100)
    b.io.steps_offset = None
    with pytest.raises(runtime.SolveError, match="steps_offset is required"):
        ctx.solve(d, b)
    b.io.steps_capacity = -1
    with pytest.raises(runtime.SolveError, match="steps_capacity < 0"):
        ctx.solve(d, b)
    # capacity 0 with no row buffers: a probe for the capacity that fits
    b = runtime.EnsembleBuffers(d, 3, u0, p, steps_capacity=0)
    b.io.u_steps = b.io.t_steps = None
    assert ctx.solve(d, b, allow=(_abi.E_SAVE_OVERFLOW,)) == _abi.E_SAVE_OVERFLOW
    assert ctx.stats()["steps_needed"] > 4 and (b.n_steps == 0).all()
    d0 = _desc(_abi.ALG_ATSIT5, 0)
    e = runtime.EnsembleBuffers(d0, 3, np.zeros((3, 0)), np.zeros((3, 0)), steps_capacity=8)
    e.steps_offset[0] = 7
    assert ctx.solve(d0, e) == 0 and e.steps_offset[0] == 0
    with pytest.raises(runtime.SolveError, match="packed save layout"):
        dg = _desc(_abi.ALG_ATSIT5, 4)
        gb = runtime.EnsembleBuffers(dg, 3, u0, p, steps_capacity=64, backend="torch",
                                     device="cuda:0")
        ctx.graph_capture(dg, gb)
