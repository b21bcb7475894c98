"""The reference's solve surface, for an ensemble on the GPU.

Mirrors what a SimpleDiffEq.jl user writes:

    prob = ODEProblem(f, u0, tspan, p)                    # SciMLBase ODEProblem
    sol = solve(prob, GPUSimpleATsit5(), dt=..., abstol=..., reltol=...)
    eprob = EnsembleProblem(prob, prob_func=lambda prob, i, repeat: remake(prob, p=...))
    esol = solve(eprob, GPUSimpleATsit5(), EnsembleHIP(), trajectories=N, ...)

with the reference's keyword defaults (gpuatsit5.jl:55-61, 205-212:
dt = 0.1f0, abstol = 1f-6, reltol = 1f-3, save_everystep = true,
saveat = nothing) and its results: per trajectory an ODESolution with t, u
and retcode "Default", or the error "dt<dtmin" the reference throws
(gpuatsit5.jl:256).  The algorithm structs are the eight GPUSimple* types.

Every solve here runs on the HIP library (one trajectory per GPU lane); there
is no CPU fallback.  Right-hand sides: the library's built-ins (rhs.lorenz,
rhs.glv16(A), ...) use kernels compiled ahead of time; any other Python
function f(u, p, t) is traced into HIP C (tracer.py) and compiled with hipRTC.
"""
import copy
import inspect
import time

import numpy as np

from . import _abi, runtime, tracer


# ---------------------------------------------------------------------------
# Algorithms (gpuatsit5.jl:52,200; gpuvern7.jl:52,295; gpuvern9.jl:52,406;
# gpurk4.jl:50; gpueuler.jl:50)
# ---------------------------------------------------------------------------
class _Alg:
    code = None
    adaptive = False
    dt_required = False

    def __repr__(self):
        return type(self).__name__ + "()"


class GPUSimpleTsit5(_Alg):
    code = _abi.ALG_TSIT5


class GPUSimpleATsit5(_Alg):
    code = _abi.ALG_ATSIT5
    adaptive = True


class GPUSimpleVern7(_Alg):
    code = _abi.ALG_VERN7


class GPUSimpleAVern7(_Alg):
    code = _abi.ALG_AVERN7
    adaptive = True


class GPUSimpleVern9(_Alg):
    code = _abi.ALG_VERN9


class GPUSimpleAVern9(_Alg):
    code = _abi.ALG_AVERN9
    adaptive = True


class GPUSimpleRK4(_Alg):
    """Classic fourth-order Runge-Kutta on the grid tspan[0]:dt:tspan[1]
    (gpurk4.jl:53-98).  dt is required; the solution holds every grid point,
    whatever save_everystep or saveat say (the reference ignores them)."""
    code = _abi.ALG_RK4
    dt_required = True


class GPUSimpleEuler(_Alg):
    """Explicit Euler on the grid tspan[0]:dt:tspan[1] (gpueuler.jl:53-90); dt
    is required and every grid point is returned, as GPUSimpleRK4."""
    code = _abi.ALG_EULER
    dt_required = True


def isadaptive(alg):
    return alg.adaptive


class DtLessThanMin(RuntimeError):
    """The reference's error("dt<dtmin") (gpuatsit5.jl:256)."""


# ---------------------------------------------------------------------------
# Problems and solutions
# ---------------------------------------------------------------------------
class ODEProblem:
    """Out-of-place problem du = f(u, p, t) on tspan (SciMLBase ODEProblem)."""

    def __init__(self, f, u0, tspan, p=None):
        self.f = f
        # A scalar u0 / p (the reference's Float64-scalar problems) is carried as one
        # component; f sees a scalar and solutions hold scalar states.
        self.scalar_u = np.ndim(u0) == 0
        self.scalar_p = p is not None and np.ndim(p) == 0
        self.u0 = np.atleast_1d(np.asarray(u0))
        self.tspan = (float(tspan[0]), float(tspan[1]))
        self.p = None if p is None else np.atleast_1d(np.asarray(p))

    def isinplace(self):
        if isinstance(self.f, BuiltinRHS):
            return False
        try:
            return len(inspect.signature(self.f).parameters) == 4
        except (TypeError, ValueError):
            return False


def remake(prob, u0=None, tspan=None, p=None):
    return ODEProblem(prob.f, prob.u0 if u0 is None else u0,
                      prob.tspan if tspan is None else tspan, prob.p if p is None else p)


class ODESolution:
    """t, u (list of states) and retcode of one trajectory."""

    def __init__(self, t, u, retcode, naccept, nreject, alg):
        self.t = t
        self.u = u
        self.retcode = retcode
        self.stats = {"naccept": int(naccept), "nreject": int(nreject)}
        self.alg = alg

    def __len__(self):
        return len(self.t)


class EnsembleProblem:
    """SciMLBase EnsembleProblem: prob_func(prob, i, repeat) makes trajectory i
    (i is 1-based, as in Julia); output_func(sol, i) -> (out, rerun) and
    reduction(u, batch, I) -> (u, converged) post-process on the host.

    For large ensembles the trajectory data can instead be given directly as
    arrays: u0s (n_state, N), ps (n_param, N), tspans (2, N)."""

    def __init__(self, prob, prob_func=None, output_func=None, reduction=None, u_init=None,
                 safetycopy=True, u0s=None, ps=None, tspans=None):
        self.prob = prob
        self.prob_func = prob_func
        self.output_func = output_func
        self.reduction = reduction
        self.u_init = u_init
        self.safetycopy = safetycopy
        self.u0s, self.ps, self.tspans = u0s, ps, tspans


class EnsembleHIP:
    """Ensemble algorithm: all trajectories on one GPU through libsimplediffeq_hip."""

    def __init__(self, device=0, schedule="auto", precision="plain", memory="host"):
        self.device = device
        self.schedule = {"auto": _abi.SCHED_AUTO, "one_launch": _abi.SCHED_ONE_LAUNCH,
                         "persistent": _abi.SCHED_PERSISTENT, "staged": _abi.SCHED_STAGED}[schedule]
        self.precision = {"plain": _abi.PREC_PLAIN, "kahan": _abi.PREC_KAHAN}[precision]
        self.memory = memory


class EnsembleSolution:
    """u (the reduced outputs), converged, elapsedTime (seconds) as SciMLBase's;
    raw: the EnsembleBuffers of the solve (a list of them, one per batch, when
    batch_size split the ensemble), arrays in structure-of-arrays layout."""

    def __init__(self, u, converged, raw, elapsedTime=0.0):
        self.u = u
        self.converged = converged
        self.raw = raw
        self.elapsedTime = elapsedTime

    def __len__(self):
        return len(self.u)

    def __getitem__(self, i):
        return self.u[i]


# ---------------------------------------------------------------------------
# Right-hand sides
# ---------------------------------------------------------------------------
class BuiltinRHS:
    """A right-hand side the library compiles ahead of time."""

    def __init__(self, rid, n_state, n_param, shared=None, pyfunc=None):
        self.rid, self.n_state, self.n_param = rid, n_state, n_param
        self.shared = shared
        self.pyfunc = pyfunc

    def __call__(self, u, p, t):
        return self.pyfunc(u, p, t)


class rhs:
    """Built-in right-hand sides: rhs.lorenz, rhs.glv16(A), rhs.expdecay3,
    rhs.affine1, rhs.linear2 (definitions in include/simplediffeq_hip.h)."""

    lorenz = BuiltinRHS(_abi.RHS_LORENZ, 3, 3, pyfunc=lambda u, p, t: [
        p[0] * (u[1] - u[0]), u[0] * (p[1] - u[2]) - u[1], u[0] * u[1] - p[2] * u[2]])
    expdecay3 = BuiltinRHS(_abi.RHS_EXPDECAY3, 3, 1, pyfunc=lambda u, p, t: [p[0] * x for x in u])
    affine1 = BuiltinRHS(_abi.RHS_AFFINE1, 1, 2, pyfunc=lambda u, p, t: [p[0] * u[0] + p[1] * t])
    linear2 = BuiltinRHS(_abi.RHS_LINEAR2, 2, 4, pyfunc=lambda u, p, t: [
        p[0] * u[0] + p[1] * u[1], p[2] * u[0] + p[3] * u[1]])

    @staticmethod
    def glv16(A):
        A = np.asarray(A, dtype=np.float64).reshape(16, 16)

        def f(u, p, t):
            out = []
            for i in range(16):
                acc = p[i]
                for j in range(16):
                    acc = acc + A[i, j] * u[j]
                out.append(u[i] * acc)
            return out
        return BuiltinRHS(_abi.RHS_GLV16, 16, 16, shared=A.reshape(-1), pyfunc=f)


_CONTEXTS = {}
_TRACED = {}


def _context(device):
    if device not in _CONTEXTS:
        _CONTEXTS[device] = runtime.Context(device)
    return _CONTEXTS[device]


def _rhs_handle(ctx, f, n_state, n_param, scalar_u=False, scalar_p=False):
    if isinstance(f, BuiltinRHS):
        return f.rid, f.shared
    key = (id(ctx), f, n_state, n_param, scalar_u, scalar_p)
    if key not in _TRACED:
        src, fn = tracer.trace(f, n_state, n_param, "sdeh_traced_%d" % len(_TRACED),
                               scalar_u=scalar_u, scalar_p=scalar_p)
        _TRACED[key] = ctx.compile_rhs(src, fn, n_state, n_param)
    return _TRACED[key], None


def _f32(x):
    return float(np.float32(x))


# ---------------------------------------------------------------------------
# solve
# ---------------------------------------------------------------------------
def solve(prob, alg, ensemblealg=None, *, trajectories=None, batch_size=None, dt=None,
          saveat=None, save_everystep=True, abstol=_f32(1e-6), reltol=_f32(1e-3),
          max_saves=None, dtype=None, maxiters=0, quirks=True, dt0s=None, order=None,
          **kwargs):
    """solve(ODEProblem, alg; ...) -> ODESolution
    solve(EnsembleProblem, alg, EnsembleHIP(); trajectories, batch_size, ...)
        -> EnsembleSolution

    Unknown keywords are accepted and ignored, as the reference's `kwargs...`.
    dt defaults to 0.1f0, except for GPUSimpleRK4 and GPUSimpleEuler, which
    raise "dt is required for this algorithm" without it (gpurk4.jl:56).

    Ensembles follow SciMLBase's loop: trajectories run in batches of
    batch_size (default: all), one GPU solve per batch; each solution goes
    through output_func(sol, i) -> (out, rerun), and trajectories asking for a
    rerun are solved again from prob_func(prob, i, repeat + 1) (together, in one
    GPU solve per round) until none asks; then reduction(u, batch_outs, I) ->
    (u, converged) folds the batch in, and converged stops the loop early."""
    if not isinstance(alg, _Alg):
        raise TypeError("alg must be one of the GPUSimple* algorithms")
    # The six Tsit5/Verner solvers keep their step and end times in a vector of
    # dt's element type, ts = Vector{eltype(dt)} (gpuatsit5.jl:72,226; gpuvern7.jl:72,
    # 321; gpuvern9.jl:72,432), and push! rounds each binary64 t into it
    # (gpuatsit5.jl:114,303,325).  The default dt is the Float32 literal 0.1f0, so
    # their sol.t is binary32 unless dt is given in binary64 (a Python float); a
    # numpy float32 dt is the Float32 case.  saveat times are the caller's array, and
    # RK4/Euler return their Float64 range (gpurk4.jl:66).
    t_f32 = dt is None or isinstance(dt, np.float32)
    if dt is None:
        if alg.dt_required:
            raise ValueError("dt is required for this algorithm")
        dt = _f32(0.1)
    kw = dict(dt=float(dt), saveat=saveat, save_everystep=save_everystep, abstol=abstol,
              reltol=reltol, max_saves=max_saves, dtype=dtype, maxiters=maxiters,
              quirks=quirks, t_f32=t_f32)
    if isinstance(prob, ODEProblem):
        sols, _ = _solve_many(EnsembleProblem(prob), alg, ensemblealg or EnsembleHIP(), [0],
                              [1], dt0s, order, kw)
        return sols[0]
    eprob = prob
    ens = ensemblealg or EnsembleHIP()
    if trajectories is None:
        raise TypeError("trajectories is required for an EnsembleProblem")
    n = int(trajectories)
    bs = n if batch_size is None else int(batch_size)
    if bs < 1:
        raise ValueError("batch_size must be >= 1")
    t_start = time.perf_counter()
    u = [] if eprob.u_init is None else eprob.u_init
    converged = False
    raws = []
    for start in range(0, n, bs):
        idx = list(range(start, min(start + bs, n)))
        batch, raw = _solve_batch(eprob, alg, ens, idx, dt0s, order, kw)
        raws.append(raw)
        I = range(idx[0] + 1, idx[-1] + 2)  # 1-based, as Julia's
        if eprob.reduction is None:
            u = list(u) + batch  # SciMLBase's default: append!(u, data), false
            converged = False
        else:
            u, converged = eprob.reduction(u, batch, I)
        if converged:
            break
    return EnsembleSolution(u, converged, raws[0] if len(raws) == 1 else raws,
                            time.perf_counter() - t_start)


def _solve_batch(eprob, alg, ens, idx, dt0s, order, kw):
    """One batch: solve, output_func, and the reruns it asks for."""
    repeats = [1] * len(idx)
    sols, raw = _solve_many(eprob, alg, ens, idx, repeats, dt0s, order, kw)
    outs = [None] * len(idx)
    pending = list(range(len(idx)))
    while pending:
        again = []
        for j, s in zip(pending, sols):
            if eprob.output_func is None:
                outs[j] = s
                continue
            x = eprob.output_func(s, idx[j] + 1)
            if not isinstance(x, tuple) or len(x) != 2:
                raise TypeError("output_func must return (output, rerun)")
            outs[j] = x[0]
            if x[1]:
                again.append(j)
        pending = again
        if pending:
            for j in pending:
                repeats[j] += 1
            sols, _ = _solve_many(eprob, alg, ens, [idx[j] for j in pending],
                                  [repeats[j] for j in pending], dt0s, order, kw)
    return outs, raw


def _materialise(eprob, idx, repeats):
    """(u0s, ps, tspans) of trajectories idx (0-based) from prob_func(prob, i, repeat)
    (i 1-based; the prototype deep-copied first when safetycopy is set), or from
    the array form of the problem."""
    base = eprob.prob
    n = len(idx)
    if eprob.u0s is not None:
        cols = np.asarray(idx)
        u0s = np.asarray(eprob.u0s)[:, cols]
        ps = None if eprob.ps is None else np.asarray(eprob.ps)[:, cols]
        tspans = (None if eprob.tspans is None
                  else np.asarray(eprob.tspans, dtype=np.float64)[:, cols])
        return u0s, ps, tspans
    if eprob.prob_func is None:
        u0s = np.repeat(base.u0.reshape(-1, 1), n, axis=1)
        ps = None if base.p is None else np.repeat(base.p.reshape(-1, 1), n, axis=1)
        return u0s, ps, None
    probs = [eprob.prob_func(copy.deepcopy(base) if eprob.safetycopy else base, i + 1, r)
             for i, r in zip(idx, repeats)]
    u0s = np.stack([np.asarray(q.u0).reshape(-1) for q in probs], axis=1)
    ps = None if base.p is None else np.stack([np.asarray(q.p).reshape(-1) for q in probs],
                                              axis=1)
    spans = np.array([q.tspan for q in probs], dtype=np.float64).T
    tspans = None if (spans == np.array(base.tspan).reshape(2, 1)).all() else spans
    return u0s, ps, tspans


def _solve_many(eprob, alg, ens, idx, repeats, dt0s, order, kw):
    """One GPU solve of trajectories idx; returns their ODESolutions and buffers."""
    base = eprob.prob
    assert not base.isinplace(), "the GPUSimple solvers are out-of-place only"
    n = len(idx)
    u0s, ps, tspans = _materialise(eprob, idx, repeats)
    n_state = u0s.shape[0]
    n_param = 0 if ps is None else ps.shape[0]
    dtype = kw["dtype"]
    if dtype is None:
        dtype = _abi.F32 if u0s.dtype == np.float32 else _abi.F64
    elif dtype in ("f32", np.float32):
        dtype = _abi.F32
    elif dtype in ("f64", np.float64):
        dtype = _abi.F64
    # Precision policy: binary32 AVern9 below a tolerance of 1e-9 runs in binary64.
    # Binary32 rounds at about 6e-8 relative, so the embedded error estimate of a
    # step asked to stay below 1e-9 is mostly rounding (Verner 9's dense-output
    # polynomials also carry coefficients up to ~1e4); the states come back in the
    # caller's binary32.
    out_np = None
    if (dtype == _abi.F32 and alg.code == _abi.ALG_AVERN9
            and min(kw["abstol"], kw["reltol"]) < 1e-9):
        dtype, out_np = _abi.F64, np.float32
    cols = np.asarray(idx)
    if dt0s is not None:
        dt0s = np.asarray(dt0s)[cols]
    if order is not None:
        # a caller's trajectory order applies to a solve of the whole ensemble
        order = np.asarray(order) if n == len(order) else None

    ctx = _context(ens.device)
    rid, shared = _rhs_handle(ctx, base.f, n_state, n_param, base.scalar_u,
                              base.scalar_p and n_param == 1)
    if ps is None:
        ps = np.zeros((0, n))
    dt, saveat, max_saves = kw["dt"], kw["saveat"], kw["max_saves"]
    everystep = saveat is None and bool(kw["save_everystep"])
    # save_everystep rows use the packed layout: no per-trajectory cap, as the
    # reference's push! (gpuatsit5.jl:113-114, 302-303).  The first capacity is a
    # guess (max_saves rows per trajectory if given, else 64); when it is short,
    # the library reports the capacity that fits and the batch is solved again.
    rows_each = max_saves if max_saves else 64
    if alg.code in _abi.FIXED_GRID_ONLY_ALGS:
        # gpurk4.jl:66-88 / gpueuler.jl:66-82: the solution is (ts, us) over the
        # whole grid; save_everystep and saveat arrive in kwargs... and are ignored.
        saveat, everystep = None, True
        rows_each = max(int(runtime.fixed_grid(base.tspan[0], float(dt), base.tspan[1])[0]), 1)
    saveat_arr = None
    if saveat is not None:
        saveat_arr = np.atleast_1d(np.asarray(saveat, dtype=np.float64))
    d = runtime.make_desc(alg.code, dtype=dtype, rhs=rid, n_traj=n, tspan=base.tspan, dt=dt,
                          abstol=kw["abstol"], reltol=kw["reltol"], saveat=saveat_arr,
                          save_everystep=everystep, maxiters=kw["maxiters"], quirks=kw["quirks"],
                          schedule=ens.schedule, precision=ens.precision,
                          save_layout=_abi.SAVE_PACKED if everystep else _abi.SAVE_TRAJ_MAJOR)
    cap = n * int(rows_each) if everystep else 0
    allow = (_abi.E_DTMIN, _abi.E_SAVE_OVERFLOW)
    while True:
        bufs = runtime.EnsembleBuffers(d, n_state, u0s, ps if n_param else None, shared, tspans,
                                       dt0s, order, steps_capacity=cap)
        st = ctx.solve(d, bufs, allow=allow)
        if st != _abi.E_SAVE_OVERFLOW:
            break
        need = int(ctx.stats()["steps_needed"])
        if need <= cap:
            raise runtime.SolveError(st, ctx.last_error())
        cap = need
    if st == _abi.E_DTMIN:
        raise DtLessThanMin("dt<dtmin")

    t0s = tspans[0] if tspans is not None else np.full(n, base.tspan[0])
    # ts = Vector{eltype(dt)} for the six Tsit5/Verner solvers without saveat
    t_type = (np.float32 if kw["t_f32"] and saveat_arr is None
              and alg.code not in _abi.FIXED_GRID_ONLY_ALGS else np.float64)
    sols = []
    for i in range(n):
        rc = _abi.RETCODE_NAMES[int(bufs.retcode[i])]
        if saveat_arr is not None:
            t = saveat_arr
            u = list(bufs.u_saveat[i])
        elif everystep:
            t, us = bufs.steps_of(i)
            t = t.astype(t_type)
            u = list(us)
        else:
            t = np.array([t0s[i], bufs.t_end[i]]).astype(t_type)
            u = [u0s[:, i].astype(bufs.u_end.dtype), bufs.u_end[:, i]]
        if out_np is not None:
            u = [np.asarray(x).astype(out_np) for x in u]
        if base.scalar_u:
            u = [x[0] for x in u]
        sols.append(ODESolution(t, u, rc, bufs.naccept[i], bufs.nreject[i], alg))
    return sols, bufs
