"""GLV-16 (16 species, interaction matrix shared by every trajectory) on the GPU
against the CPU oracle, in its three lane layouts: one lane per trajectory; the
group layout (16 adjacent lanes per trajectory, one component each, state
exchanged through LDS); and, for the binary64 adaptive solvers, the matrix-core
layout (4 lanes per trajectory 16 apart, 4 components each, A u as four
v_mfma_f64_16x16x4f64 per wave; binary32 ATsit5 on v_mfma_f32_16x16x4f32).  binary64
is bit for bit against the oracle (endpoint and saveat, both schedules); binary32 is
bit for bit between the one-lane, group and matrix-core layouts, which evaluate the
same operations in the same order."""
import numpy as np
import pytest

from helpers import assert_bit_identical, bits_equal, problem, run_pair
from simplediffeq_hip import _abi, runtime

pytestmark = pytest.mark.gpu

FIELDS = ("u_end", "t_end", "retcode", "naccept", "nreject")


def _desc(alg, n, saveat, lanes, **kw):
    adaptive = alg in (_abi.ALG_ATSIT5, _abi.ALG_AVERN9)
    sv = np.linspace(0.0, 5.0, 17) if saveat else None
    return runtime.make_desc(alg, rhs=_abi.RHS_GLV16, n_traj=n, tspan=(0.0, 5.0),
                             dt=0.1 if adaptive else 0.05, abstol=1e-9, reltol=1e-9,
                             saveat=sv, lanes=lanes, **kw)


@pytest.mark.parametrize("layout", [_abi.SAVE_POINT_MAJOR, _abi.SAVE_TRAJ_MAJOR],
                         ids=["point-major", "traj-major"])
@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT])
@pytest.mark.parametrize("lanes", [1, 16, 4])
@pytest.mark.parametrize("alg,saveat", [(_abi.ALG_AVERN9, False), (_abi.ALG_AVERN9, True),
                                        (_abi.ALG_ATSIT5, False), (_abi.ALG_ATSIT5, True),
                                        (_abi.ALG_TSIT5, False)],
                         ids=["avern9", "avern9-saveat", "atsit5", "atsit5-saveat", "tsit5"])
def test_glv16_binary64_bitexact(ctx, alg, saveat, lanes, sched, layout):
    if alg == _abi.ALG_TSIT5 and sched == _abi.SCHED_PERSISTENT:
        pytest.skip("fixed-step solvers always run one launch")
    if alg == _abi.ALG_TSIT5 and lanes == 4:
        pytest.skip("the matrix-core layout has adaptive kernels only")
    if layout == _abi.SAVE_TRAJ_MAJOR and not saveat:
        pytest.skip("the save layout only concerns saveat output")
    n = 203  # not a multiple of the 4 (group layout) or 16 (matrix-core) trajectories of a wave
    u0, r, A = problem(_abi.RHS_GLV16, n, seed=3)
    d = _desc(alg, n, saveat, lanes, schedule=sched, chunk=3, save_layout=layout)
    st_g, g, st_c, c = run_pair(ctx, d, u0, r, A)
    assert st_g == 0 and st_c == 0
    assert ctx.stats()["lanes"] == lanes
    assert_bit_identical(g, c, FIELDS + (("u_saveat",) if saveat else ()))
    if alg != _abi.ALG_TSIT5:
        assert (g.numpy("naccept") > 5).all()


@pytest.mark.parametrize("lanes", [1, 16, 4])
def test_glv16_refill_batches_bitexact(ctx, lanes):
    """Persistent refill batches (a wave starts new trajectories once `refill` of its
    slots are free), up to the wave's whole slot count, in every lane layout: the
    oracle's bits, saveat included."""
    n = 203
    u0, r, A = problem(_abi.RHS_GLV16, n, seed=3)
    for k in (2, 64):
        d = _desc(_abi.ALG_AVERN9, n, True, lanes, schedule=_abi.SCHED_PERSISTENT, chunk=3,
                  refill=k)
        st_g, g, st_c, c = run_pair(ctx, d, u0, r, A)
        assert st_g == 0 and st_c == 0
        assert ctx.stats()["refill"] == min(k, 64 // lanes)
        assert_bit_identical(g, c, FIELDS + ("u_saveat",))


@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT])
@pytest.mark.parametrize("kahan,saveat,layout",
                         [(False, False, _abi.SAVE_POINT_MAJOR),
                          (False, True, _abi.SAVE_POINT_MAJOR),
                          (False, True, _abi.SAVE_TRAJ_MAJOR),
                          (True, False, _abi.SAVE_POINT_MAJOR)],
                         ids=["plain", "plain-saveat", "plain-saveat-traj-major", "kahan"])
def test_glv16_binary32_layouts_identical(ctx, kahan, saveat, layout, sched):
    n = 301
    u0, r, A = problem(_abi.RHS_GLV16, n, seed=4)
    prec = _abi.PREC_KAHAN if kahan else _abi.PREC_PLAIN
    out = {}
    for lanes in (1, 16, 4):
        d = _desc(_abi.ALG_ATSIT5, n, saveat, lanes, dtype=_abi.F32, precision=prec,
                  schedule=sched, chunk=3, save_layout=layout)
        d.abstol, d.reltol = 1e-6, 1e-6
        b = runtime.EnsembleBuffers(d, 16, u0, r, A)
        ctx.solve(d, b)
        assert ctx.stats()["lanes"] == lanes
        out[lanes] = b
    for f in FIELDS + (("u_saveat",) if saveat else ()):
        assert bits_equal(out[1].numpy(f), out[16].numpy(f)), f
        assert bits_equal(out[1].numpy(f), out[4].numpy(f)), f
    # and close to the binary64 oracle at a tight tolerance
    from oracle import pyoracle
    d = _desc(_abi.ALG_AVERN9, n, False, 1)
    d.abstol, d.reltol = 1e-12, 1e-12
    ref = runtime.EnsembleBuffers(d, 16, u0, r, A)
    assert pyoracle.solve(d, ref, nthreads=8) == 0
    err = np.abs(out[16].numpy("u_end").astype(np.float64) - ref.numpy("u_end"))
    assert np.max(err / np.abs(ref.numpy("u_end"))) < 1e-3


def test_glv16_layout_selection(ctx):
    n = 64
    u0, r, A = problem(_abi.RHS_GLV16, n, seed=5)
    # auto, binary64: the matrix-core layout where it is built, else the group layout
    d = _desc(_abi.ALG_AVERN9, n, False, 0)
    ctx.solve(d, runtime.EnsembleBuffers(d, 16, u0, r, A))
    assert ctx.stats()["lanes"] == 4
    d = _desc(_abi.ALG_TSIT5, n, False, 0)
    ctx.solve(d, runtime.EnsembleBuffers(d, 16, u0, r, A))
    assert ctx.stats()["lanes"] == 16
    # binary32 defaults to one lane per trajectory
    d = _desc(_abi.ALG_ATSIT5, n, False, 0, dtype=_abi.F32)
    ctx.solve(d, runtime.EnsembleBuffers(d, 16, u0, r, A))
    assert ctx.stats()["lanes"] == 1
    # the matrix-core layout on request: binary64 AVern9 and ATsit5, binary32 ATsit5
    for kw in ({"alg": _abi.ALG_AVERN9}, {"alg": _abi.ALG_ATSIT5, "dtype": _abi.F32}):
        alg = kw.pop("alg")
        d = _desc(alg, n, False, 4, **kw)
        ctx.solve(d, runtime.EnsembleBuffers(d, 16, u0, r, A))
        assert ctx.stats()["lanes"] == 4
    for kw in ({"alg": _abi.ALG_TSIT5}, {"alg": _abi.ALG_AVERN9, "dtype": _abi.F32}):
        alg = kw.pop("alg")
        d = _desc(alg, n, False, 4, **kw)
        with pytest.raises(runtime.SolveError, match="no matrix-core kernel"):
            ctx.solve(d, runtime.EnsembleBuffers(d, 16, u0, r, A))
    # an unsupported group size, and a group layout for a right-hand side without one
    d = _desc(_abi.ALG_AVERN9, n, False, 8)
    with pytest.raises(runtime.SolveError, match="lanes per trajectory"):
        ctx.solve(d, runtime.EnsembleBuffers(d, 16, u0, r, A))
    dl = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=8, lanes=16)
    ul, pl, _ = problem(_abi.RHS_LORENZ, 8)
    with pytest.raises(runtime.SolveError, match="lane per trajectory"):
        ctx.solve(dl, runtime.EnsembleBuffers(dl, 3, ul, pl))


@pytest.mark.parametrize("lanes", [16, 4], ids=["group", "matrix-core"])
@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT])
def test_glv16_spans_dt0_order_and_retcodes(ctx, lanes, sched):
    """Per-trajectory spans (handed out longest first), per-trajectory dt0, a caller's
    order, and trajectories that stop early (maxiters, dt < dtmin) in the group and
    matrix-core layouts: every output bit for bit the oracle's."""
    n = 333
    u0, r, A = problem(_abi.RHS_GLV16, n, seed=21)
    rng = np.random.default_rng(21)
    ts = np.stack([rng.uniform(0.0, 1.0, n), rng.uniform(1.5, 6.0, n)])
    dt0 = rng.uniform(0.01, 0.3, n)
    dt0[::37] = 1e-16  # below AVern9's dtmin: ends at once with RET_DTMIN
    d = runtime.make_desc(_abi.ALG_AVERN9, rhs=_abi.RHS_GLV16, n_traj=n, dt=0.1, abstol=1e-9,
                          reltol=1e-9, lanes=lanes, schedule=sched, chunk=5, maxiters=12)
    st_g, g, st_c, c = run_pair(ctx, d, u0, r, A, tspan=ts, dt0=dt0)
    assert st_g == st_c == _abi.E_DTMIN
    assert ctx.stats()["lanes"] == lanes
    assert_bit_identical(g, c, FIELDS)
    ret = g.numpy("retcode")
    assert (ret[::37] == _abi.RET_DTLESSTHANMIN).all() and (ret == _abi.RET_MAXITERS).any()
    # a caller's hand-out order changes nothing
    order = rng.permutation(n).astype(np.int64)
    d2 = runtime.make_desc(_abi.ALG_AVERN9, rhs=_abi.RHS_GLV16, n_traj=n, dt=0.1, abstol=1e-9,
                           reltol=1e-9, lanes=lanes, schedule=sched, chunk=5, maxiters=12)
    h = runtime.EnsembleBuffers(d2, 16, u0, r, A, ts, dt0, order)
    assert ctx.solve(d2, h, allow=(_abi.E_DTMIN,)) == _abi.E_DTMIN
    assert_bit_identical(h, c, FIELDS)
