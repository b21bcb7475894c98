"""GPU integrators vs the CPU oracle on identical inputs, through the C ABI.

In binary64 the kernels implement the same operation order as the oracle's
fused mode and the same pow, so the comparison is bit for bit (endpoint state,
final time, retcode, accepted and rejected step counts, saveat output).
"""
import numpy as np
import pytest

from helpers import ALG_NAMES, assert_bit_identical, problem, run_pair
from simplediffeq_hip import _abi, inputs, runtime

pytestmark = pytest.mark.gpu

FIXED = (_abi.ALG_TSIT5, _abi.ALG_VERN7, _abi.ALG_VERN9)
ADAPT = (_abi.ALG_ATSIT5, _abi.ALG_AVERN7, _abi.ALG_AVERN9)


def test_lorenz_small_fixed_tsit5_bitexact(ctx):
    """Lorenz 1,024 trajectories, rho ~ U(0,21), GPUSimpleTsit5 dt = 0.01 on (0, 1)."""
    u0, p = inputs.lorenz_small(1024, seed=0)
    d = runtime.make_desc(_abi.ALG_TSIT5, rhs=_abi.RHS_LORENZ, n_traj=1024, tspan=(0.0, 1.0),
                          dt=0.01)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p)
    assert st_g == 0 and st_c == 0
    assert (g.naccept == 100).all()
    assert_bit_identical(g, c)


def test_lorenz_small_adaptive_atsit5_bitexact(ctx):
    """Same ensemble, GPUSimpleATsit5 abstol 1e-6, reltol 1e-3, dt0 = 0.1."""
    u0, p = inputs.lorenz_small(1024, seed=0)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=1024, tspan=(0.0, 1.0),
                          dt=0.1, abstol=1e-6, reltol=1e-3)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p)
    assert st_g == 0 and st_c == 0
    assert_bit_identical(g, c)


@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT])
@pytest.mark.parametrize("alg", list(range(6)), ids=lambda a: ALG_NAMES[a])
@pytest.mark.parametrize("rhs", [_abi.RHS_LORENZ, _abi.RHS_EXPDECAY3, _abi.RHS_AFFINE1,
                                 _abi.RHS_LINEAR2], ids=["lorenz", "expdecay3", "affine1",
                                                         "linear2"])
def test_endpoint_bitexact(ctx, rhs, alg, sched):
    n = 777
    u0, p, ps = problem(rhs, n, seed=alg + 10 * rhs)
    adaptive = alg in ADAPT
    if alg in FIXED and sched == _abi.SCHED_PERSISTENT:
        pytest.skip("fixed-step solvers always use one lane per trajectory")
    d = runtime.make_desc(alg, rhs=rhs, n_traj=n, tspan=(0.0, 2.0),
                          dt=0.1 if adaptive else 0.02, abstol=1e-7, reltol=1e-6,
                          schedule=sched, chunk=16)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p, ps)
    assert st_g == st_c
    assert_bit_identical(g, c)


LAYOUTS = [_abi.SAVE_POINT_MAJOR, _abi.SAVE_TRAJ_MAJOR]


@pytest.mark.parametrize("layout", LAYOUTS, ids=["point-major", "traj-major"])
@pytest.mark.parametrize("quirks", [True, False])
@pytest.mark.parametrize("alg", list(range(6)), ids=lambda a: ALG_NAMES[a])
@pytest.mark.parametrize("rhs", [_abi.RHS_LORENZ, _abi.RHS_AFFINE1], ids=["lorenz", "affine1"])
def test_saveat_bitexact(ctx, rhs, alg, quirks, layout):
    n = 300
    u0, p, ps = problem(rhs, n, seed=3 + alg)
    adaptive = alg in ADAPT
    saveat = np.concatenate([[0.0, 0.05, 0.4], np.linspace(0.5, 1.5, 23), [1.5, 1.5]])
    d = runtime.make_desc(alg, rhs=rhs, n_traj=n, tspan=(0.0, 1.5), dt=0.1 if adaptive else 0.01,
                          abstol=1e-8, reltol=1e-7, saveat=saveat, quirks=quirks,
                          save_layout=layout)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p, ps)
    assert st_g == st_c == 0
    assert_bit_identical(g, c, fields=("u_end", "t_end", "retcode", "naccept", "nreject",
                                       "u_saveat"))


@pytest.mark.parametrize("layout", LAYOUTS, ids=["point-major", "traj-major"])
@pytest.mark.parametrize("alg", list(range(6)), ids=lambda a: ALG_NAMES[a])
def test_save_everystep_bitexact(ctx, alg, layout):
    n = 200
    u0, p, _ = problem(_abi.RHS_LORENZ, n, seed=7)
    adaptive = alg in ADAPT
    d = runtime.make_desc(alg, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 1.0),
                          dt=0.1 if adaptive else 0.05, abstol=1e-6, reltol=1e-4,
                          save_everystep=True, max_saves=64, save_layout=layout)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p)
    assert st_g == st_c
    assert_bit_identical(g, c, fields=("u_end", "t_end", "retcode", "naccept", "nreject",
                                       "n_steps"))
    ns = g.n_steps
    tm = layout == _abi.SAVE_TRAJ_MAJOR
    for i in range(n):
        k = ns[i]
        if tm:
            assert np.array_equal(g.t_steps[i, :k], c.t_steps[i, :k])
            assert np.array_equal(g.u_steps[i, :k], c.u_steps[i, :k])
        else:
            assert np.array_equal(g.t_steps[:k, i], c.t_steps[:k, i])
            assert np.array_equal(g.u_steps[:k, :, i], c.u_steps[:k, :, i])


def test_device_memory_path_matches_host(ctx):
    torch = pytest.importorskip("torch")
    n = 4096
    u0, p = inputs.lorenz_wide(n, seed=1)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 10.0),
                          dt=float(np.float32(0.1)), abstol=1e-6, reltol=1e-3)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p, backend="torch")
    assert st_g == 0 and st_c == 0
    assert_bit_identical(g, c)


@pytest.mark.parametrize("alg", [_abi.ALG_ATSIT5, _abi.ALG_AVERN7, _abi.ALG_AVERN9],
                         ids=["atsit5", "avern7", "avern9"])
def test_controller_range_edges_bitexact(ctx, alg):
    """u' = lambda u with lambda spanning the error estimate's whole range, mixed in
    every wave: EEst exactly 0 (lambda = 0), subnormal or underflowing error sums
    (|lambda| ~ 1e-160 .. 1e-200), sums near the edges of the range-checked controller
    ([2^-700, 2^700]: |lambda| ~ 1e-100 .. 1e-110 puts it on either side of the lower
    edge), ordinary ones, and exploding ones that overflow and stop with a NaN error
    estimate.  The controller then takes its range-checked and its IEEE path in the
    same wave; every output bit for bit the oracle's."""
    n = 640
    rng = np.random.default_rng(31)
    lam = np.concatenate([
        np.zeros(64),                                   # EEst = 0: dt grows by qmax
        -np.exp(rng.uniform(np.log(1e-200), np.log(1e-160), 128)),
        -np.exp(rng.uniform(np.log(1e-110), np.log(1e-100), 64)),  # s near 2^-700
        -rng.uniform(0.1, 3.0, 256),                    # ordinary
        np.exp(rng.uniform(np.log(1e100), np.log(1e160), 64)),     # explode: NaN stop
        rng.uniform(1e300, 1e305, 64),                  # overflow at once: NaN stop
    ])
    perm = rng.permutation(n)
    p = lam[perm][None, :]
    u0 = rng.uniform(0.5, 2.0, (3, n))
    d = runtime.make_desc(alg, rhs=_abi.RHS_EXPDECAY3, n_traj=n, tspan=(0.0, 1.0), dt=0.1,
                          abstol=1e-6, reltol=1e-3, maxiters=200)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p)
    assert st_g == st_c
    assert_bit_identical(g, c)
    ret = c.numpy("retcode")[np.argsort(perm)]
    assert (ret[:512] == _abi.RET_DEFAULT).all() and (ret[512:] == _abi.RET_NAN).all()


@pytest.mark.parametrize("tols", [(1e-6, 1e-3), (1e-8, 0.125), (1e-300, 1e-3), (1e-6, 0.5),
                                  (0.0, 1e-3)],
                         ids=["bare", "bare-reltol-edge", "tiny-abstol", "large-reltol",
                              "abstol-0"])
@pytest.mark.parametrize("alg", [_abi.ALG_ATSIT5, _abi.ALG_AVERN7, _abi.ALG_AVERN9],
                         ids=["atsit5", "avern7", "avern9"])
@pytest.mark.parametrize("sched", [_abi.SCHED_ONE_LAUNCH, _abi.SCHED_PERSISTENT,
                                   _abi.SCHED_STAGED], ids=["one", "persistent", "staged"])
def test_error_norm_scales_bitexact(ctx, alg, tols, sched):
    """u' = lambda u with the three components of one trajectory at very different
    magnitudes (1e-250 .. 1e300), so single terms of the error norm square to +0, to a
    subnormal, or are divided by scales far above 2^60, while the sum stays in the
    range-checked controller's range.  The binary64 kernels divide the norm's terms by
    bare Newton iterations when abstol is in [2^-420, 2^1020] and reltol in [0, 2^-3]
    (the first two tolerance pairs) and by the IEEE sequence otherwise; every output bit
    for bit the oracle's either way, in every schedule (the staged one parks the ATsit5
    trajectories that need more than 32 attempts, up to 215 of them, and resumes them in a
    second launch)."""
    abstol, reltol = tols
    n = 768
    rng = np.random.default_rng(47)
    mags = np.array([1e-250, 1e-170, 1e-160, 1e-100, 1.0, 1e100, 1e250, 1e300])
    u0 = (mags[rng.integers(0, len(mags), (3, n))] * rng.uniform(0.5, 2.0, (3, n))
          * rng.choice([-1.0, 1.0], (3, n)))
    p = -rng.uniform(0.1, 3.0, (1, n))
    d = runtime.make_desc(alg, rhs=_abi.RHS_EXPDECAY3, n_traj=n, tspan=(0.0, 10.0), dt=0.1,
                          abstol=abstol, reltol=reltol, maxiters=1000, schedule=sched)
    st_g, g, st_c, c = run_pair(ctx, d, u0, p)
    assert st_g == st_c
    assert_bit_identical(g, c)
    assert (c.numpy("retcode") == _abi.RET_DEFAULT).all()
