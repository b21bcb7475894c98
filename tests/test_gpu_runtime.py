"""Runtime features on the GPU: persistent queue and hand-out order, graph
replay, autotuner, device-pointer validation, and the binary32 precision modes."""
import numpy as np
import pytest

from helpers import bits_equal
from simplediffeq_hip import _abi, inputs, runtime

pytestmark = pytest.mark.gpu

DEV = "cuda:0"


def _dev_bufs(d, u0, p, **kw):
    return runtime.EnsembleBuffers(d, u0.shape[0], u0, p, backend="torch", device=DEV, **kw)


def _desc(alg=_abi.ALG_ATSIT5, n=5000, **kw):
    base = dict(rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 5.0), dt=0.1, abstol=1e-6, reltol=1e-4)
    base.update(kw)
    return runtime.make_desc(alg, **base)


def test_order_permutation_leaves_results_unchanged(ctx):
    n = 5000
    u0, p = inputs.lorenz_wide(n, seed=1)
    ref = _dev_bufs(_desc(n=n, schedule=_abi.SCHED_ONE_LAUNCH), u0, p)
    ctx.solve(ref.desc, ref)
    perm = np.random.default_rng(0).permutation(n).astype(np.int64)
    for sched in (_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT):
        for chunk in (1, 7, 64):
            d = _desc(n=n, schedule=sched, chunk=chunk)
            b = _dev_bufs(d, u0, p, order=perm)
            ctx.solve(d, b)
            for f in ("u_end", "t_end", "naccept", "nreject", "retcode"):
                assert bits_equal(b.numpy(f), ref.numpy(f)), (sched, chunk, f)


def test_async_solves_on_two_streams_share_the_context_safely(ctx):
    """Device-memory solves return once queued; a second call on another stream
    must not overwrite the first one's staged saveat times or work queue."""
    import torch
    n = 200000
    u0, p = inputs.lorenz_wide(n, seed=5)
    sa = [np.linspace(0.0, 5.0, 11), np.linspace(0.5, 4.5, 7)]
    descs = [_desc(n=n, saveat=sa[k], schedule=_abi.SCHED_PERSISTENT) for k in range(2)]
    refs = []
    for d in descs:
        b = _dev_bufs(d, u0, p)
        ctx.solve(d, b)
        refs.append(b)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    outs = [_dev_bufs(d, u0, p) for d in descs]
    torch.cuda.synchronize()
    ctx.solve(descs[0], outs[0], stream=s1.cuda_stream, sync=False)
    ctx.solve(descs[1], outs[1], stream=s2.cuda_stream, sync=False)
    ctx.synchronize()
    torch.cuda.synchronize()
    for k in range(2):
        for f in ("u_end", "u_saveat", "naccept", "nreject", "retcode"):
            assert bits_equal(outs[k].numpy(f), refs[k].numpy(f)), (k, f)


def test_compile_rhs_rejects_bad_names_and_sources(ctx):
    src = ("template <class T> __device__ void decay(T* du, const T* u, const T* p,"
           " const T* __restrict__ ps, double t) { du[0] = p[0] * u[0]; }")
    for bad in ("", "1decay", "de cay", "x" * 201):
        with pytest.raises(runtime.SolveError, match="C identifier"):
            ctx.compile_rhs(src, bad, 1, 1)
    with pytest.raises(runtime.SolveError):
        ctx.compile_rhs(src.replace("du[0] =", "du[0] = = "), "decay", 1, 1)
    h = ctx.compile_rhs(src, "decay", 1, 1)
    d = runtime.make_desc(_abi.ALG_TSIT5, rhs=h, n_traj=4, tspan=(0.0, 1.0), dt=0.125)
    b = runtime.EnsembleBuffers(d, 1, np.ones((1, 4)), np.full((1, 4), -1.0))
    ctx.solve(d, b)
    assert np.allclose(b.u_end[0], np.exp(-1.0), rtol=1e-9)


def test_persistent_grid_smaller_than_needed_still_covers_all(ctx):
    n = 3000
    u0, p = inputs.lorenz_wide(n, seed=1)
    d = _desc(n=n, schedule=_abi.SCHED_PERSISTENT, block=64, grid=3, chunk=5)
    b = _dev_bufs(d, u0, p)
    ctx.solve(d, b)
    ref = _dev_bufs(_desc(n=n, schedule=_abi.SCHED_ONE_LAUNCH), u0, p)
    ctx.solve(ref.desc, ref)
    assert bits_equal(b.numpy("u_end"), ref.numpy("u_end"))
    assert (b.numpy("retcode") == 0).all()


def test_graph_replay_matches_eager_solve(ctx):
    torch = pytest.importorskip("torch")
    n = 4096
    u0, p, ts = inputs.lorenz_uneven(n, seed=2)
    d = _desc(_abi.ALG_AVERN7, n=n, abstol=1e-8, reltol=1e-8)
    eager = _dev_bufs(d, u0, p, tspan=ts)
    ctx.solve(d, eager)
    g = _dev_bufs(d, u0, p, tspan=ts)
    h = ctx.graph_capture(d, g)
    try:
        for rep in range(3):
            g.u_end.fill_(0.0)
            g.naccept.fill_(-1)
            ctx.graph_launch(h)
            ctx.synchronize()
            for f in ("u_end", "t_end", "naccept", "nreject", "retcode"):
                assert bits_equal(g.numpy(f), eager.numpy(f)), (rep, f)
            # eager work between replays
            ctx.solve(d, eager)
            torch.cuda.synchronize()
    finally:
        ctx.graph_destroy(h)


def test_autotune_then_solve_is_identical(ctx):
    n = 20000
    u0, p = inputs.lorenz_wide(n, seed=1)
    d = _desc(n=n)
    b = _dev_bufs(d, u0, p)
    best = ctx.autotune(d, b)
    assert best["kernel_ms"] > 0 and best["block"] in (64, 128, 256)
    ctx.solve(d, b)
    ref = _dev_bufs(_desc(n=n, schedule=_abi.SCHED_ONE_LAUNCH), u0, p)
    ctx.solve(ref.desc, ref)
    assert bits_equal(b.numpy("u_end"), ref.numpy("u_end"))


def test_invalid_device_buffers_are_refused(ctx):
    torch = pytest.importorskip("torch")
    n = 1000
    u0, p = inputs.lorenz_wide(n, seed=1)
    d = _desc(n=n)
    # Both refusals happen on the host before any launch: a host pointer passed
    # as device memory, and an ensemble size no allocation here can hold.
    b = _dev_bufs(d, u0, p)
    host = np.empty((3, n))
    b.io.u_end = host.ctypes.data
    with pytest.raises(runtime.SolveError, match="u_end"):
        ctx.solve(d, b)
    b = _dev_bufs(d, u0, p)
    d_big = _desc(n=n)
    d_big.n_traj = 10 ** 9
    with pytest.raises(runtime.SolveError, match="u0"):
        ctx.solve(d_big, b)
    b = _dev_bufs(d, u0, p, order=np.zeros(n, dtype=np.int64))
    with pytest.raises(runtime.SolveError, match="permutation"):
        ctx.solve(d, b)
    dg = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_GLV16, n_traj=8, tspan=(0.0, 1.0))
    bg = runtime.EnsembleBuffers(dg, 16, np.ones((16, 8)), np.ones((16, 8)), backend="torch",
                                 device=DEV)
    with pytest.raises(runtime.SolveError, match="shared"):
        ctx.solve(dg, bg)
    d2 = runtime.make_desc(_abi.ALG_TSIT5, rhs=_abi.RHS_LORENZ, n_traj=n, dt=0.0)
    with pytest.raises(runtime.SolveError, match="zero"):
        ctx.solve(d2, _dev_bufs(d2, u0, p))


def _rel_err(a, ref):
    a = np.asarray(a, dtype=np.float64)
    return np.max(np.abs(a - ref), axis=0) / np.max(np.abs(ref), axis=0)


def test_binary32_modes_against_binary64(ctx):
    # binary32 is a precision mode: error against the binary64 result at a tight
    # tolerance, with the compensated update no worse than the plain one.
    n = 4096
    u0, p = inputs.lorenz_wide(n, seed=1)
    kw = dict(n=n, tspan=(0.0, 10.0), abstol=1e-7, reltol=1e-5)
    ref = _dev_bufs(_desc(_abi.ALG_AVERN9, n=n, tspan=(0.0, 10.0), abstol=1e-13, reltol=1e-13),
                    u0, p)
    ctx.solve(ref.desc, ref)
    uref = ref.numpy("u_end")
    f64 = _dev_bufs(_desc(**kw), u0, p)
    ctx.solve(f64.desc, f64)
    plain = _dev_bufs(_desc(dtype=_abi.F32, **kw), u0, p)
    ctx.solve(plain.desc, plain)
    comp = _dev_bufs(_desc(dtype=_abi.F32, precision=_abi.PREC_KAHAN, **kw), u0, p)
    ctx.solve(comp.desc, comp)
    e64 = np.median(_rel_err(f64.numpy("u_end"), uref))
    e32 = np.median(_rel_err(plain.numpy("u_end"), uref))
    ek = np.median(_rel_err(comp.numpy("u_end"), uref))
    assert e64 < 1e-4 and ek < 1e-4
    assert ek <= e32 * 1.05
    assert (comp.numpy("retcode") == 0).all() and (comp.numpy("t_end") == 10.0).all()


def test_solve_on_another_stream_waits_for_the_buffer_conversion(ctx):
    """EnsembleBuffers converts binary64 device inputs to binary32 on torch's current
    stream; a solve on another stream must start after that conversion (the buffers
    carry an event the solve's stream waits on).  2^22 trajectories, so that the
    conversion takes long enough to overlap an unordered launch; maxiters bounds the
    run should inputs ever be read half-written."""
    import torch
    n = 1 << 22
    u0, p = inputs.lorenz_wide(n, seed=21)
    dev = torch.device(DEV)
    u0d, pd = torch.as_tensor(u0, device=dev), torch.as_tensor(p, device=dev)
    d = _desc(n=n, dtype=_abi.F32, schedule=_abi.SCHED_ONE_LAUNCH, maxiters=20000)
    ref = _dev_bufs(d, u0d, pd)
    ctx.solve(d, ref)
    side = torch.cuda.Stream(device=dev)
    torch.cuda.synchronize()
    for _ in range(3):
        b = _dev_bufs(d, u0d, pd)  # conversion kernels queued on the current stream
        ctx.solve(d, b, stream=side.cuda_stream, sync=True)
        for f in ("u_end", "t_end", "naccept", "nreject", "retcode"):
            assert bits_equal(b.numpy(f), ref.numpy(f)), f
