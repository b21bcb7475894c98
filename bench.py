#!/usr/bin/env python3
"""Benchmark: adaptive ensemble throughput on one MI355X.

Default line (one "step" = one ensemble solve): the learned-binning evaluation
ensemble, Lorenz with 4,194,304 trajectories, sigma ~ U(8,12), rho ~ U(0,28),
beta ~ U(2,3), u0 ~ U(-1,1)^3, t_end ~ LogUniform(0.5,10) (Philox seed 4),
GPUSimpleATsit5 at abstol 1e-7, reltol 1e-5, dt0 = 0.1f0, endpoint output.
Inputs are resident in HBM before the timed region.

value = trajectories/s of the binary64 solve (the reference's precision) over the
K timed steps, in the launch shape the library's autotuner picks (schedule,
block, grid, queue chunk, register cap, trajectories per lane, refill batch;
config.launch).  Beside it:
  * learned: the cost model (6-64-64-2 MLP) trained on the GPU from 2^20 plain
    binary64 step logs (seed 40), then the binned binary32 Kahan-compensated
    solve of the same ensemble, planning (features, forward pass, hand-out order)
    included, against the plain binary64 one-launch kernel in index order, with an
    ablation (untrained weights, no order, the model-free orders) and the same
    binned solves with the model's predicted step size as each trajectory's dt0
    (dt0_seeding: time, attempts, rejections, error);
  * error of every reported dtype against binary64 AVern9 at 1e-14 on the first
    65,536 trajectories (median and 90th percentile of the normwise relative
    endpoint error);
  * lorenz_wide: the 1,048,576-trajectory Lorenz ensemble (rho ~ U(0,21), seed 1,
    tspan (0,10), abstol 1e-6, reltol 1e-3), binary64 and binary32.

roofline: the integrator kernel is bound by vector (VALU) binary64 throughput:
achieved = algorithmic FLOPs per solve / average device time per solve (HIP
events on the launch stream over the timed region), peak = 78.6 TFLOP/s.
Algorithmic FLOPs = sum over trajectories of (accepted + rejected attempts) x
(3 * 67 + 6 * 8 + 12) + 8 per trajectory (the initial FSAL evaluation): 67 FLOPs
of stage, error and scaling arithmetic per component per attempt, 6 Lorenz
evaluations of 8 FLOPs, 2 for the norm and 10 for the controller; FMA = 2,
add/mul/div/sqrt/pow = 1.

cpu_baseline: the CPU restatement of the same solver (oracle/, one thread per
host core of the GPU's share, contiguous chunks as EnsembleThreads does) on the
first 65,536 trajectories of the same ensemble.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "simplediffeq.jl_amd"))

N_HEADLINE = 1 << 22  # the learned-binning evaluation ensemble
N_WIDE = 1 << 20      # the wide Lorenz ensemble
N_TRAIN = 1 << 20     # step logs the cost model is trained on
N_ERR = 1 << 16       # trajectories checked against binary64 AVern9 at 1e-14
N_TRAJ = N_WIDE       # default size of the single-workload runs
# The host threads of the CPU baseline: the GPU's share of the host is 16 threads
# (the CPU count the process sees may be larger), capped at the CPUs it may use.
CPU_SHARE = 16
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}
HBM_PEAK_GBS = 8000.0

# HBM bytes per integrator launch, from two rocprofv3 --pmc passes (FETCH_SIZE x
# 2.000 + WRITE_SIZE x 1.000, factors from a 256 MiB device copy in the same run).
# Counters cannot be read from inside a timed run, so each figure applies only to the
# configuration it was measured on.
#   The wide ensemble, binary64, 2^20 trajectories, default shapes (refill batches of
#   4): persistent 49,660 KiB read + 104,942 KiB written; staged (first pass, key sort,
#   resumed pass) 250,358 KiB read + 390,722 KiB written, mostly the parked records and
#   the resumed pass's endpoint stores in hand-out order (partial lines).
#   The default line, binary64, 2^22 trajectories with their own t_end, persistent
#   (256 x 1,280, chunk 64, refill batches of 4), longest span first: 3,851,216 KiB
#   read + 930,024 KiB written; each trajectory's eight input loads land on eight
#   separate lines, since the span order hands neighbouring lanes trajectories about a
#   hundred indices apart.
MEASURED_TRAFFIC = {("f64", 1 << 20, 2): 158.30e6, ("f64", 1 << 20, 3): 656.42e6,
                    ("f64", 1 << 22, 2, "costmodel"): 4.896e9}


def measured_traffic(key, launch):
    """The counted bytes for `key`, or None: the figures were counted on the default
    kernel (no register cap, one trajectory per lane), so a launch the autotuner moved to
    another kernel variant reports none."""
    if launch.get("waves", 0) != 0 or launch.get("per_lane", 1) != 1:
        return None
    return MEASURED_TRAFFIC.get(key)
# Algorithmic bytes per trajectory of an endpoint solve with its own tspan (binary64
# Lorenz): reads u0, p (48), tspan (16), the hand-out slot (8); writes u_end (24),
# t_end (8), retcode, naccept, nreject (12).
ALG_BYTES_TSPAN = 48 + 16 + 8 + 24 + 8 + 12

# Algorithmic FLOPs (per component per attempt, RHS calls per attempt) by solver.
STAGE_FLOPS = {0: 48, 1: 67, 2: 83, 3: 117, 4: 167, 5: 211, 6: 13, 7: 2}
RHS_PER_ATTEMPT = {0: 6, 1: 6, 2: 9, 3: 9, 4: 15, 5: 15, 6: 4, 7: 1}
LORENZ_FLOPS = 8


def algorithmic_flops(alg, n_state, f_rhs, attempts, accepts, n_traj):
    adaptive = alg in (1, 3, 5)
    per_attempt = n_state * STAGE_FLOPS[alg] + RHS_PER_ATTEMPT[alg] * f_rhs + (12 if adaptive else 0)
    total = attempts * per_attempt
    if alg in (3, 5):
        total += accepts * f_rhs  # k1 once per accepted step
    if alg in (0, 1):
        total += n_traj * f_rhs  # initial FSAL evaluation
    return float(total)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def cpu_time(desc, n_state, u0, p, ps=None, tspan=None, threads=1, min_seconds=10.0,
             max_reps=400):
    """Oracle solves of the host ensemble (u0, p, ...) described by `desc`, repeated
    until at least `min_seconds` have elapsed; returns (trajectories/s, seconds, reps)."""
    from oracle import pyoracle
    from simplediffeq_hip import runtime
    n = int(desc.n_traj)
    bufs = runtime.EnsembleBuffers(desc, n_state, u0, p, ps, tspan)
    pyoracle.solve(desc, bufs, nthreads=threads)  # warm-up (page in, thread start)
    reps = 0
    t0 = time.perf_counter()
    while True:
        st = pyoracle.solve(desc, bufs, nthreads=threads)
        if st != 0:
            raise RuntimeError("oracle status %d" % st)
        reps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds or reps >= max_reps:
            break
    return n * reps / el, el, reps


def cpu_threads():
    """(threads used, CPUs the process may use, CPUs of the host)."""
    allowed = len(os.sched_getaffinity(0))
    return min(CPU_SHARE, allowed), allowed, os.cpu_count()


def cpu_leg(desc, n_state, u0, p, ps=None, tspan=None, min_seconds=3.0):
    """The `cpu_baseline` record of a workload (oracle on the host cores)."""
    threads, allowed, nproc = cpu_threads()
    rate, el, reps = cpu_time(desc, n_state, u0, p, ps, tspan, threads, min_seconds)
    return {"value": rate, "unit": "trajectories/s", "cores": threads, "kind": "port",
            "nproc": nproc, "affinity": allowed,
            "sample": "first %d trajectories, solved %d times in %.1f s on %s, %d threads "
                      "(the GPU's share of the host; %d CPUs allowed, %d in the host)"
                      % (int(desc.n_traj), reps, el, cpu_model(), threads, allowed, nproc)}


def _time_solves(ctx, d, bufs, stream, steps, warmup):
    """Average device time per solve (HIP events on the launch stream) and wall time."""
    import torch
    for _ in range(warmup):
        ctx.solve(d, bufs, stream=stream.cuda_stream, sync=False)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        ctx.solve(d, bufs, stream=stream.cuda_stream, sync=False)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps, (time.perf_counter() - t0) / steps


def workload_small(args):
    """The 1,024-trajectory Lorenz ensemble (rho ~ U(0,21), seed 0, tspan (0,1)):
    fixed Tsit5 dt = 0.01 and ATsit5 1e-6/1e-3, GPU vs the CPU restatement."""
    import numpy as np
    import torch
    from oracle import pyoracle
    from simplediffeq_hip import _abi, inputs, runtime
    u0, p = inputs.lorenz_small(1024, seed=0)
    ctx = runtime.Context(0)
    stream = torch.cuda.Stream()
    out = {"workload": "lorenz-small"}
    ref = _Reference(ctx, u0, p, t=(0.0, 1.0))
    for name, alg, dt in (("tsit5", _abi.ALG_TSIT5, 0.01), ("atsit5", _abi.ALG_ATSIT5, 0.1)):
        d = runtime.make_desc(alg, rhs=_abi.RHS_LORENZ, n_traj=1024, tspan=(0.0, 1.0), dt=dt,
                              abstol=1e-6, reltol=1e-3)
        g = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0")
        ms, wall = _time_solves(ctx, d, g, stream, args.steps, args.warmup)
        c = runtime.EnsembleBuffers(d, 3, u0, p)
        threads = cpu_threads()[0]
        t0 = time.perf_counter()
        for _ in range(5):
            pyoracle.solve(d, c, nthreads=threads)
        cpu = (time.perf_counter() - t0) / 5
        same = all(np.array_equal(g.numpy(f), c.numpy(f)) for f in ("u_end", "naccept", "nreject"))
        out[name] = {"gpu_kernel_ms": ms, "gpu_call_ms": wall * 1e3, "cpu_ms": cpu * 1e3,
                     "cpu_threads": threads, "bit_identical": bool(same),
                     "error_vs_avern9_1e-14": ref.error(g)}
    return out


def workload_uneven(args):
    """2,097,152 Lorenz trajectories, rho ~ U(0,60), t_end ~ LogUniform(0.5,20),
    u0 ~ N(0,1)^3 (seed 2), AVern7 binary64 at 1e-8/1e-8: one-launch vs persistent
    vs autotuned, and K repeats from a HIP graph vs K solve calls.  Trajectories are
    handed out longest span first (the library default with per-trajectory spans);
    the *_index_order entries time index order for comparison."""
    import torch
    from simplediffeq_hip import _abi, inputs, runtime
    n = args.n if args.n != N_TRAJ else 1 << 21
    u0, p, ts = inputs.lorenz_uneven(n, seed=2)
    ctx = runtime.Context(0)
    stream = torch.cuda.Stream()
    out = {"workload": "lorenz-uneven", "trajectories": n}
    base = dict(rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 1.0), dt=float(__import__("numpy").float32(0.1)),
                abstol=1e-8, reltol=1e-8)
    for sched_name, sched, so in (("one_launch", _abi.SCHED_ONE_LAUNCH, 0),
                                  ("persistent", _abi.SCHED_PERSISTENT, 0),
                                  ("one_launch_index_order", _abi.SCHED_ONE_LAUNCH, 1),
                                  ("persistent_index_order", _abi.SCHED_PERSISTENT, 1)):
        d = runtime.make_desc(_abi.ALG_AVERN7, schedule=sched, span_order=so, **base)
        g = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, backend="torch", device="cuda:0")
        ctx.solve(d, g, sync=True)
        st = ctx.stats()
        ms, wall = _time_solves(ctx, d, g, stream, args.steps, args.warmup)
        flops = algorithmic_flops(_abi.ALG_AVERN7, 3, LORENZ_FLOPS, st["attempts"], st["accepts"], n)
        out[sched_name] = {"kernel_ms": ms, "traj_per_s": n / (ms * 1e-3),
                           "tflops": flops / (ms * 1e-3) / 1e12, "attempts": st["attempts"],
                           "launch": {k: st[k] for k in ("block", "grid", "chunk")}}
    d = runtime.make_desc(_abi.ALG_AVERN7, **base)
    g = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, backend="torch", device="cuda:0")
    best = ctx.autotune(d, g)
    ms, _ = _time_solves(ctx, d, g, stream, args.steps, args.warmup)
    out["autotuned"] = {"kernel_ms": ms, "traj_per_s": n / (ms * 1e-3),
                        "launch": {k: best[k] for k in ("block", "grid", "chunk", "schedule")},
                        "error_vs_avern9_1e-14": _Reference(ctx, u0, p, tspan=ts).error(g)}
    if not args.no_cpu:
        k = 1 << 16
        dc = runtime.make_desc(_abi.ALG_AVERN7, **dict(base, n_traj=k))
        out["cpu_baseline"] = cpu_leg(dc, 3, u0[:, :k], p[:, :k], tspan=ts[:, :k])
    h = ctx.graph_capture(d, g)
    torch.cuda.synchronize()
    reps = max(args.steps, 5)
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.graph_launch(h)
    ctx.synchronize()
    t_graph = (time.perf_counter() - t0) / reps
    ctx.graph_destroy(h)
    out["graph_replay"] = {"ms_per_repeat": t_graph * 1e3, "traj_per_s": n / t_graph}
    return out


def workload_glv(args):
    """GLV-16, 131,072 trajectories (seed 3), tspan (0,50), saveat 256 evenly spaced
    points: AVern9 binary64 at 1e-10 and ATsit5 binary32 at 1e-6.  Saved states in
    the trajectory-major layout ([trajectory][point][component], each trajectory's
    solution contiguous); the point-major layout is timed beside it."""
    import numpy as np
    import torch
    from simplediffeq_hip import _abi, inputs, runtime
    n = args.n if args.n != N_TRAJ else 1 << 17
    u0, r, A = inputs.glv16(n, seed=3)
    saveat = np.linspace(0.0, 50.0, 256)
    ctx = runtime.Context(0)
    stream = torch.cuda.Stream()
    out = {"workload": "glv16-saveat", "trajectories": n}
    # endpoint error against binary64 AVern9 at 1e-14 on the first 16,384 trajectories
    ref = _Reference(ctx, u0, r, t=(0.0, 50.0), m=1 << 14, rhs=_abi.RHS_GLV16, ps=A)
    for name, alg, dtype, tol in (("avern9_f64", _abi.ALG_AVERN9, _abi.F64, 1e-10),
                                  ("atsit5_f32", _abi.ALG_ATSIT5, _abi.F32, 1e-6)):
        # default lane layout in both save layouts; binary64 also in each layout
        variants = [(_abi.SAVE_TRAJ_MAJOR, 0, ""), (_abi.SAVE_POINT_MAJOR, 0, "_point_major")]
        if dtype == _abi.F64:
            variants += [(_abi.SAVE_TRAJ_MAJOR, 16, "_lanes16"), (_abi.SAVE_TRAJ_MAJOR, 4, "_lanes4")]
        else:
            variants += [(_abi.SAVE_TRAJ_MAJOR, 1, "_lanes1"), (_abi.SAVE_TRAJ_MAJOR, 4, "_lanes4"),
                         (_abi.SAVE_POINT_MAJOR, 4, "_lanes4_point_major")]
        for layout, lanes, tag in variants:
            d = runtime.make_desc(alg, dtype=dtype, rhs=_abi.RHS_GLV16, n_traj=n,
                                  tspan=(0.0, 50.0), dt=0.1, abstol=tol, reltol=tol,
                                  saveat=saveat, save_layout=layout, lanes=lanes)
            g = runtime.EnsembleBuffers(d, 16, u0, r, A, backend="torch", device="cuda:0")
            ctx.solve(d, g, sync=True)
            st = ctx.stats()
            ms, _ = _time_solves(ctx, d, g, stream, max(1, args.steps // 4), 1)
            es = 8 if dtype == _abi.F64 else 4
            out_bytes = 256 * 16 * n * es
            per_attempt = 16 * STAGE_FLOPS[alg] + RHS_PER_ATTEMPT[alg] * 528 + 12
            flops = st["attempts"] * per_attempt + (st["accepts"] * 528 if alg == 5 else n * 528)
            out[name + tag] = {"kernel_ms": ms, "traj_per_s": n / (ms * 1e-3),
                               "attempts": st["attempts"], "lanes": st["lanes"],
                               "saveat_GBs": out_bytes / (ms * 1e-3) / 1e9,
                               "tflops_excl_saveat": flops / (ms * 1e-3) / 1e12}
            if not tag:
                out[name + tag]["error_vs_avern9_1e-14"] = ref.error(g)
            del g
        if not args.no_cpu:
            k = min(n, 1 << 14 if dtype == _abi.F64 else 1 << 16)
            dc = runtime.make_desc(alg, dtype=dtype, rhs=_abi.RHS_GLV16, n_traj=k,
                                   tspan=(0.0, 50.0), dt=0.1, abstol=tol, reltol=tol,
                                   saveat=saveat, save_layout=_abi.SAVE_TRAJ_MAJOR)
            out[name]["cpu_baseline"] = cpu_leg(dc, 16, u0[:, :k], r[:, :k], A)
    return out


def workload_everystep(args):
    """The headline ensemble (Lorenz, 2^20 trajectories, ATsit5 1e-6/1e-3, tspan
    (0,10)) with every accepted step saved: the packed layout (one pass into the
    chunk pool, then the offsets scan and compaction; capacity = the points
    offered) against fixed rows per trajectory (trajectory-major, rows = the
    largest step count) and endpoint-only output.  Times cover the whole solve on
    the stream (all kernels)."""
    import numpy as np
    import torch
    from simplediffeq_hip import _abi, inputs, runtime
    n = args.n
    u0, p = inputs.lorenz_wide(n, seed=1)
    ctx = runtime.Context(0)
    stream = torch.cuda.Stream()
    kw = dict(rhs=_abi.RHS_LORENZ, n_traj=n, tspan=(0.0, 10.0), dt=float(np.float32(0.1)),
              abstol=1e-6, reltol=1e-3)
    out = {"workload": "lorenz-wide-everystep", "trajectories": n}
    reps = max(1, args.steps // 4)
    d = runtime.make_desc(_abi.ALG_ATSIT5, **kw)
    g = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0")
    ms, _ = _time_solves(ctx, d, g, stream, reps, 1)
    out["endpoint"] = {"ms": ms}
    del g
    dp = runtime.make_desc(_abi.ALG_ATSIT5, save_everystep=True, save_layout=_abi.SAVE_PACKED,
                           **kw)
    probe = runtime.EnsembleBuffers(dp, 3, u0, p, backend="torch", device="cuda:0",
                                    steps_capacity=0)
    ctx.solve(dp, probe, sync=True, allow=(_abi.E_SAVE_OVERFLOW,))
    rows = int(ctx.stats()["steps_needed"])
    del probe
    g = runtime.EnsembleBuffers(dp, 3, u0, p, backend="torch", device="cuda:0",
                                steps_capacity=rows)
    ctx.solve(dp, g, sync=True)
    kernel_ms = ctx.stats()["kernel_ms"]
    counts = g.numpy("n_steps")
    assert int(counts.sum()) == rows
    ms, _ = _time_solves(ctx, dp, g, stream, reps, 1)
    out_bytes = rows * (3 * 8 + 8)
    out["packed"] = {"ms": ms, "integration_kernel_ms": kernel_ms, "rows": rows,
                     "rows_max": int(counts.max()), "rows_mean": rows / n,
                     "output_bytes": out_bytes, "output_GBs": out_bytes / (ms * 1e-3) / 1e9}
    del g
    ms_rows = int(counts.max())
    dt_ = runtime.make_desc(_abi.ALG_ATSIT5, save_everystep=True, max_saves=ms_rows,
                            save_layout=_abi.SAVE_TRAJ_MAJOR, **kw)
    g = runtime.EnsembleBuffers(dt_, 3, u0, p, backend="torch", device="cuda:0")
    ms, _ = _time_solves(ctx, dt_, g, stream, reps, 1)
    out["fixed_rows_traj_major"] = {"ms": ms, "rows_per_trajectory": ms_rows,
                                    "output_bytes": n * ms_rows * 32}
    del g
    return out


def _timed_steps(ctx, d, bufs, stream, steps, warmup):
    """The timed region: `warmup` untimed solves, then exactly `steps`
    solves bracketed by synchronisation.  Returns (device ms per solve from HIP
    events on the launch stream, wall seconds of the K solves)."""
    import torch
    for _ in range(warmup):
        ctx.solve(d, bufs, stream=stream.cuda_stream, sync=False)
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        ctx.solve(d, bufs, stream=stream.cuda_stream, sync=False)
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


def _events(stream, fn, reps=5):
    """Median device time (ms) of fn() on `stream` over `reps` runs after one warm-up."""
    import numpy as np
    import torch
    fn()
    torch.cuda.synchronize()
    out = []
    for _ in range(reps):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        out.append(e0.elapsed_time(e1))
    return float(np.median(out))


class _Reference:
    """Binary64 AVern9 at abstol = reltol = 1e-14 on the first `m` trajectories of
    an ensemble (the oracle runs of the error figures), solved on the GPU."""

    def __init__(self, ctx, u0, p, tspan=None, t=None, m=N_ERR, rhs=None, ps=None):
        import numpy as np
        from simplediffeq_hip import _abi, runtime
        m = min(m, u0.shape[1])
        kw = dict(tspan=t) if t is not None else {}
        d = runtime.make_desc(_abi.ALG_AVERN9, rhs=_abi.RHS_LORENZ if rhs is None else rhs,
                              n_traj=m, dt=float(np.float32(0.1)), abstol=1e-14, reltol=1e-14,
                              **kw)
        g = runtime.EnsembleBuffers(d, u0.shape[0], u0[:, :m], p[:, :m], ps,
                                    tspan=None if tspan is None else tspan[:, :m],
                                    backend="torch", device="cuda:0")
        ctx.solve(d, g, sync=True)
        self.m = m
        self.u = g.numpy("u_end").astype(np.float64)
        self.ok = g.numpy("retcode") == 0

    def error(self, bufs):
        """Median and 90th percentile of max_c |u - u_ref| / max_c |u_ref| over the
        trajectories the reference finished."""
        import numpy as np
        ue = bufs.numpy("u_end")[:, :self.m].astype(np.float64)
        e = np.max(np.abs(ue - self.u), axis=0) / np.max(np.abs(self.u), axis=0)
        e = e[self.ok]
        return {"median": float(np.median(e)), "p90": float(np.percentile(e, 90)),
                "trajectories": int(self.ok.sum())}


def _progress(msg):
    """One line on stderr as each phase of a run ends (the JSON line is stdout's only)."""
    print("bench: " + msg, file=sys.stderr, flush=True)


def _roofline(alg_flops, kernel_ms, peak, kernel, traffic=None, extra=None):
    ach = alg_flops / (kernel_ms * 1e-3) / 1e12
    r = {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
         "traffic": traffic, "kernel": kernel, "kernel_ms": kernel_ms,
         "algorithmic_flops_per_launch": alg_flops}
    if extra:
        r.update(extra)
    return r


_KERNEL = "sdeh_adaptive_kernel<%s, Lorenz, ATsit5, endpoint%s>"
_SCHED = {1: "one launch", 2: "persistent", 3: "staged: 32-attempt first pass, key sort, resumed "
          "persistent pass"}


def _kernel_name(dtype, sched, kahan=False):
    return (_KERNEL % ("double" if dtype == "f64" else "float", ", Kahan" if kahan else "")
            + " (" + _SCHED.get(sched, str(sched)) + ")")


def lorenz_wide(args, ctx, stream, steps, warmup):
    """The 1,048,576-trajectory Lorenz ensemble (seed 1, rho ~ U(0,21), tspan
    (0,10), ATsit5 1e-6/1e-3, dt0 0.1f0, endpoint): binary64 and binary32, each
    autotuned, K timed solves, error against AVern9 at 1e-14 on 65,536."""
    import numpy as np
    import torch
    from simplediffeq_hip import _abi, inputs, runtime
    n = N_WIDE if args.n in (N_TRAJ, N_HEADLINE) else args.n
    u0, p = inputs.lorenz_wide(n, seed=1)
    ref = _Reference(ctx, u0, p, t=(0.0, 10.0))
    out = {"workload": "Lorenz ensemble, GPUSimpleATsit5, abstol 1e-6, reltol 1e-3, tspan "
                       "(0,10), dt0 0.1f0, endpoint only",
           "trajectories": n, "sigma": "U(8,12)", "rho": "U(0,21)", "beta": "U(2,3)",
           "u0": "U(-1,1)^3", "seed": 1}
    for name, dtype in (("f64", _abi.F64), ("f32", _abi.F32)):
        d = runtime.make_desc(_abi.ALG_ATSIT5, dtype=dtype, rhs=_abi.RHS_LORENZ, n_traj=n,
                              tspan=(0.0, 10.0), dt=float(np.float32(0.1)), abstol=1e-6,
                              reltol=1e-3, schedule=args.schedule if name == "f64" else 0)
        g = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0")
        torch.cuda.synchronize()
        if args.tune:
            ctx.autotune(d, g)
        ctx.solve(d, g, stream=stream.cuda_stream, sync=True)
        st = ctx.stats()
        if not (g.numpy("retcode") == 0).all():
            raise RuntimeError("wide ensemble: trajectories did not finish")
        ms, wall = _timed_steps(ctx, d, g, stream, steps, warmup)
        flops = algorithmic_flops(_abi.ALG_ATSIT5, 3, LORENZ_FLOPS, st["attempts"], st["accepts"], n)
        sched = st["schedule"]
        out[name] = {"value": n * steps / wall, "unit": "trajectories/s", "kernel_ms": ms,
                     "ms_per_step": wall * 1e3 / steps,
                     "roofline": _roofline(flops, ms, PEAK_TFLOPS[name], _kernel_name(name, sched),
                                           measured_traffic((name, n, sched), st),
                                           {"attempts_per_launch": st["attempts"]}),
                     "error_vs_avern9_1e-14": ref.error(g),
                     "launch": {k: st[k] for k in ("block", "grid", "chunk", "schedule", "waves",
                                                   "per_lane", "refill")}}
        if name == "f32":
            out[name]["roofline"]["frac_of_78.6_unpacked"] = out[name]["roofline"]["achieved"] / 78.6
        del g
    return out


def headline(args):
    """The default line: the learned-binning evaluation ensemble (see the module
    docstring), binary64 `value`, the learned block, errors, the wide ensemble and
    the CPU baseline."""
    import numpy as np
    import torch
    from simplediffeq_hip import _abi, costmodel, inputs, runtime
    n = N_HEADLINE if args.n == N_TRAJ else args.n
    steps, warmup = args.steps, args.warmup
    ctx = runtime.Context(0)
    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)  # default priority
    f32_01 = float(np.float32(0.1))
    tol = dict(abstol=1e-7, reltol=1e-5)

    # 1. the cost model: step logs of the plain solver and training, on the GPU
    cm, hist, t_train, t_logs = costmodel.train_lorenz(ctx, n=min(N_TRAIN, n), seed=40,
                                                       steps=args.train_steps, **tol)
    _progress("cost model trained (%.2f s)" % t_train)
    # 2. the evaluation ensemble, resident in HBM
    u0, p, ts = inputs.lorenz_costmodel(n, seed=4)
    u0d, pd, tsd = (torch.as_tensor(a, device=dev) for a in (u0, p, ts))
    torch.cuda.synchronize()
    ref = _Reference(ctx, u0, p, tspan=ts)

    _progress("binary64 AVern9 1e-14 reference solved")

    def bufs(dtype=_abi.F64, sched=0, precision=_abi.PREC_PLAIN, order=None, span=True,
             dt0=None, **kw):
        d = runtime.make_desc(_abi.ALG_ATSIT5, dtype=dtype, rhs=_abi.RHS_LORENZ, n_traj=n,
                              dt=f32_01, schedule=sched, precision=precision,
                              span_order=0 if span else 1, **tol, **kw)
        g = runtime.EnsembleBuffers(d, 3, u0d, pd, tspan=tsd, dt0=dt0, order=order,
                                    backend="torch", device=dev)
        return d, g

    def solve_once(d, g):
        ctx.solve(d, g, stream=stream.cuda_stream, sync=True)
        st = ctx.stats()
        if not (g.numpy("retcode") == 0).all():
            raise RuntimeError("trajectories did not finish")
        return st

    def time_of(d, g, reps):
        return _events(stream, lambda: ctx.solve(d, g, stream=stream.cuda_stream, sync=False),
                       reps)

    # 3. value: binary64 in the autotuned shape, the timed region
    shape = {}
    if args.shape:
        sh = [int(v) for v in args.shape.split(",")]
        shape = dict(sched=sh[0], block=sh[1], grid=sh[2], chunk=sh[3],
                     refill=sh[4] if len(sh) > 4 else 0)
    d64, g64 = bufs(**(shape or dict(sched=args.schedule)))
    if args.tune and not shape:
        ctx.autotune(d64, g64)
    st64 = solve_once(d64, g64)
    ms64, wall64 = _timed_steps(ctx, d64, g64, stream, steps, warmup)
    flops64 = algorithmic_flops(_abi.ALG_ATSIT5, 3, LORENZ_FLOPS, st64["attempts"],
                                st64["accepts"], n)
    att = g64.numpy("naccept").astype(np.int64) + g64.numpy("nreject")
    err64 = ref.error(g64)
    sched64 = st64["schedule"]
    launch64 = {k: st64[k] for k in ("block", "grid", "chunk", "schedule", "waves", "per_lane",
                                     "refill")}
    del g64
    _progress("binary64 timed: %.3f ms per solve, launch %s" % (ms64, launch64))
    # the autotuner's kernel variants on the same solve (the tuner's choice is above)
    variants = {}
    for name, kw in (("waves_4", dict(waves=4)), ("waves_6", dict(waves=6)),
                     ("waves_8", dict(waves=8)),
                     ("two_per_lane_persistent", dict(per_lane=2, sched=2)),
                     ("two_per_lane_one_launch", dict(per_lane=2, sched=1))):
        kw.setdefault("sched", sched64 if sched64 != 3 or "waves" in kw else 2)
        d, g = bufs(**kw)
        stv = solve_once(d, g)
        variants[name] = {"kernel_ms": time_of(d, g, 5), "schedule": stv["schedule"],
                          "grid": stv["grid"]}
        del g
    _progress("kernel variants timed")

    # 4. binary32 (plain) in its autotuned shape, for its error and rate
    d32, g32 = bufs(dtype=_abi.F32)
    if args.tune:
        ctx.autotune(d32, g32)
    st32 = solve_once(d32, g32)
    ms32 = time_of(d32, g32, 5)
    f32 = {"kernel_ms": ms32, "value": n / (ms32 * 1e-3), "unit": "trajectories/s",
           "error_vs_avern9_1e-14": ref.error(g32),
           "roofline": _roofline(algorithmic_flops(_abi.ALG_ATSIT5, 3, LORENZ_FLOPS,
                                                   st32["attempts"], st32["accepts"], n),
                                 ms32, PEAK_TFLOPS["f32"], _kernel_name("f32", st32["schedule"])),
           "launch": {k: st32[k] for k in ("block", "grid", "chunk", "schedule", "waves",
                                           "per_lane", "refill")}}
    del g32
    _progress("binary32 timed")

    # 5. learned: planning on the device, binned binary32 Kahan solves, baselines
    def plan(model):
        return model.plan_lorenz(pd, u0d, tsd, stream, pred=False)

    plan_ms = _events(stream, lambda: plan(cm))
    _, order, _ = plan(cm)
    untrained = costmodel.CostModel(ctx, seed=0)
    untrained.y_mean, untrained.y_std = cm.y_mean, cm.y_std
    plan_ms_untrained = _events(stream, lambda: plan(untrained))
    _, order_untrained, _ = plan(untrained)
    torch.cuda.synchronize()
    reps = max(5, steps // 2)
    runs = {}
    # a binned run plans right before its solves, as a caller does: the context takes
    # the order sdeh_plan_order wrote last without the permutation check
    for name, dtype, sched, prec, model, span in (
            ("f64_one_launch_index_order", _abi.F64, 1, _abi.PREC_PLAIN, None, False),
            ("f32k_persistent_binned", _abi.F32, 2, _abi.PREC_KAHAN, cm, False),
            ("f32k_one_launch_binned", _abi.F32, 1, _abi.PREC_KAHAN, cm, False),
            ("f32k_persistent_binned_seeded", _abi.F32, 2, _abi.PREC_KAHAN, cm, False),
            ("f32k_one_launch_binned_seeded", _abi.F32, 1, _abi.PREC_KAHAN, cm, False),
            ("f32k_one_launch_binned_two_per_lane", _abi.F32, 1, _abi.PREC_KAHAN, cm, False),
            ("f32k_one_launch_binned_untrained", _abi.F32, 1, _abi.PREC_KAHAN, untrained,
             False),
            ("f32k_one_launch_index_order", _abi.F32, 1, _abi.PREC_KAHAN, None, False),
            ("f32k_one_launch_span_order", _abi.F32, 1, _abi.PREC_KAHAN, None, True),
            ("f32k_persistent_index_order", _abi.F32, 2, _abi.PREC_KAHAN, None, False),
            ("f32k_persistent_span_order", _abi.F32, 2, _abi.PREC_KAHAN, None, True),
            ("f32k_staged", _abi.F32, 3, _abi.PREC_KAHAN, None, True)):
        # every run, the baseline included, takes the fastest of a few launch shapes
        # of its own schedule and order (the library's default, blocks of 64 and 128,
        # for binary32 the 4- and 6-wave register caps, and for the persistent and
        # staged schedules refill batches of 8 and 64 beside the default 4)
        base_kw = dict(per_lane=2) if name.endswith("two_per_lane") else {}
        shapes = [{}, dict(block=64), dict(block=128)]
        if dtype == _abi.F32 and not base_kw and sched != 3:
            shapes += [dict(waves=4), dict(waves=6)]
        if sched in (2, 3) and not base_kw:
            shapes += [dict(refill=8), dict(refill=64)]
        seeded = name.endswith("_seeded")  # the model's dt0 as each trajectory's first dt
        best = None
        for shape in shapes:
            od, dt0 = (None, None) if model is None else plan(model)[1:]
            d, g = bufs(dtype, sched, prec, od, span, dt0 if seeded else None, **base_kw,
                        **shape)
            st = solve_once(d, g)
            ms = time_of(d, g, reps)
            if best is None or ms < best[0]:
                best = (ms, st, {k: st[k] for k in ("block", "grid", "waves", "per_lane",
                                                    "refill")})
            if shape != shapes[-1]:
                del g
        ms, st, launch = best
        runs[name] = {"kernel_ms": ms, "attempts": st["attempts"], "accepts": st["accepts"],
                      "schedule": st["schedule"], "launch": launch}
        if name == "f32k_persistent_binned_seeded":
            err_seeded = ref.error(g)
        if name == "f32k_persistent_binned":
            err32k = ref.error(g)
            ach = algorithmic_flops(_abi.ALG_ATSIT5, 3, LORENZ_FLOPS, st["attempts"],
                                    st["accepts"], n)
            roof32k = _roofline(ach, runs[name]["kernel_ms"], PEAK_TFLOPS["f32"],
                                _kernel_name("f32", 2, kahan=True))
        del g
    base_ms = runs["f64_one_launch_index_order"]["kernel_ms"]
    binned_ms = runs["f32k_persistent_binned"]["kernel_ms"] + plan_ms
    best_free = min(("f32k_one_launch_index_order", "f32k_one_launch_span_order",
                     "f32k_persistent_index_order", "f32k_persistent_span_order", "f32k_staged"),
                    key=lambda k: runs[k]["kernel_ms"])
    best_binned = min(("f32k_persistent_binned", "f32k_one_launch_binned",
                       "f32k_one_launch_binned_two_per_lane"),
                      key=lambda k: runs[k]["kernel_ms"])
    learned = {
        "value": n / (binned_ms * 1e-3), "unit": "trajectories/s",
        "dtype": "f32 state, Kahan-compensated update, f64 time",
        "workload": "persistent, binned by the trained model; planning included",
        "plan_ms": plan_ms, "solve_ms": runs["f32k_persistent_binned"]["kernel_ms"],
        "speedup_vs_f64_one_launch_index_order": base_ms / binned_ms,
        "target_3x_met": base_ms / binned_ms >= 3.0,
        "error_vs_avern9_1e-14": err32k, "target_median_1e-5_met": err32k["median"] <= 1e-5,
        "roofline": roof32k,
        "solves_ms": {k: v["kernel_ms"] for k, v in runs.items()},
        "solve_launches": {k: dict(v["launch"], schedule=v["schedule"]) for k, v in runs.items()},
        "gain": {
            "wave_efficiency_one_launch": {
                "trained": costmodel.wave_efficiency(att, order.cpu().numpy()),
                "untrained": costmodel.wave_efficiency(att, order_untrained.cpu().numpy()),
                "index_order": costmodel.wave_efficiency(att)},
            "plan_ms_untrained": plan_ms_untrained,
            "best_binned_with_plan_ms": runs[best_binned]["kernel_ms"] + plan_ms,
            "best_binned": best_binned,
            "best_model_free_ms": runs[best_free]["kernel_ms"], "best_model_free": best_free,
            "net_gain_over_best_model_free_ms":
                runs[best_free]["kernel_ms"] - (runs[best_binned]["kernel_ms"] + plan_ms)},
        "train": {"step_logs": min(N_TRAIN, n), "step_log_solve_s": t_logs, "adam_steps":
                  args.train_steps, "seconds": t_train, "loss": [hist[0][1], hist[-1][1]]},
        # the model's predicted step size as each trajectory's first dt instead of 0.1f0:
        # the controller still decides every step, so results change within tolerance
        # (the unseeded runs above are bit-identical to unbinned ones)
        "dt0_seeding": {
            "persistent_binned_seeded_ms": runs["f32k_persistent_binned_seeded"]["kernel_ms"],
            "one_launch_binned_seeded_ms": runs["f32k_one_launch_binned_seeded"]["kernel_ms"],
            "speedup_persistent_with_plan":
                base_ms / (runs["f32k_persistent_binned_seeded"]["kernel_ms"] + plan_ms),
            "attempts": {k: runs[k]["attempts"] for k in ("f32k_persistent_binned",
                                                          "f32k_persistent_binned_seeded")},
            "rejected": {k: runs[k]["attempts"] - runs[k]["accepts"]
                         for k in ("f32k_persistent_binned", "f32k_persistent_binned_seeded")},
            "error_vs_avern9_1e-14": err_seeded},
    }

    traffic64 = measured_traffic(("f64", n, sched64, "costmodel"), launch64)
    out = {
        "metric": "adaptive trajectories/s at stated tol vs FP64 oracle; % FP32/FP64 VALU peak",
        "value": n * steps / wall64,
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": wall64 * 1e3 / steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox seed 4; cost model trained on seed 40)",
        "config": {"workload": "Lorenz ensemble, GPUSimpleATsit5, abstol 1e-7, reltol 1e-5, "
                               "dt0 0.1f0, endpoint only, per-trajectory t_end (the learned-"
                               "binning evaluation ensemble)",
                   "trajectories": n, "sigma": "U(8,12)", "rho": "U(0,28)", "beta": "U(2,3)",
                   "u0": "U(-1,1)^3", "t_end": "LogUniform(0.5,10)", "launch": launch64},
        "roofline": _roofline(flops64, ms64, PEAK_TFLOPS["f64"], _kernel_name("f64", sched64),
                              traffic64,
                              {"attempts_per_launch": st64["attempts"],
                               # rejected attempts run the same branch-free attempt
                               # body as accepted ones; lane use needs the hardware
                               # counters (rocprofv3), which a timed run cannot read
                               "rejected_attempt_fraction":
                                   1.0 - st64["accepts"] / max(1, st64["attempts"]),
                               "kernel_ms_covers": "the whole solve: span order and integrator",
                               "algorithmic_bytes_gbs": n * ALG_BYTES_TSPAN / (ms64 * 1e-3) / 1e9,
                               "traffic_gbs": None if traffic64 is None else
                               traffic64 / (ms64 * 1e-3) / 1e9}),
        "error_vs_avern9_1e-14": err64,
        "kernel_variants_f64_ms": variants,
        "f32": f32,
        "learned": learned,
    }
    _progress("learned binning timed")
    if not args.no_wide:
        out["lorenz_wide"] = lorenz_wide(args, ctx, stream, steps, warmup)
        _progress("wide ensemble timed")
    if not args.no_cpu:
        k = min(n, 1 << 16)
        dc = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=k, dt=f32_01, **tol)
        out["cpu_baseline"] = cpu_leg(dc, 3, u0[:, :k], p[:, :k], tspan=ts[:, :k],
                                      min_seconds=10.0)
    ctx.close()
    return out


def wide_line(args):
    """--workload lorenz-wide: the wide ensemble alone, as its own JSON line."""
    import torch
    from simplediffeq_hip import runtime
    ctx = runtime.Context(0)
    stream = torch.cuda.Stream(device=torch.device("cuda:0"))
    out = lorenz_wide(args, ctx, stream, args.steps, args.warmup)
    if not args.no_cpu:
        import numpy as np
        from simplediffeq_hip import _abi, inputs
        k = 1 << 16
        u0, p = inputs.lorenz_wide(k, seed=1)
        dc = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=k, tspan=(0.0, 10.0),
                               dt=float(np.float32(0.1)), abstol=1e-6, reltol=1e-3)
        out["cpu_baseline"] = cpu_leg(dc, 3, u0, p, min_seconds=10.0)
    ctx.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_TRAJ,
                    help="ensemble size (default: each workload's own)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-wide", action="store_true",
                    help="default line: skip the wide-ensemble block")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 256 MiB device copy first (reference kernel for the HBM "
                         "byte counters)")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="skip the autotuner (by default sdeh_autotune picks the launch shape "
                         "once per variant, before the warmup and the timed steps)")
    ap.add_argument("--schedule", type=int, default=0, choices=[0, 1, 2, 3],
                    help="binary64 schedule: 0 library choice (autotuned), 1 one launch, "
                         "2 persistent, 3 staged")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "costmodel", "lorenz-wide", "lorenz-small",
                             "lorenz-uneven", "glv16", "everystep"],
                    help="headline (= costmodel) is the default line; the others report their "
                         "own JSON")
    ap.add_argument("--train-steps", type=int, default=1500)
    ap.add_argument("--shape", default="",
                    help="SCHEDULE,BLOCK,GRID,CHUNK[,REFILL]: the binary64 launch of the default "
                         "line, instead of the autotuner's (to repeat a measured shape, e.g. "
                         "under counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU; --gpus %d is not supported "
              "(trajectories are independent: an N-GPU run gives each GPU a contiguous slice)"
              % args.gpus, file=sys.stderr)
        return 2
    if args.copy_probe:
        import torch
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    fn = {"headline": headline, "costmodel": headline, "lorenz-wide": wide_line,
          "lorenz-small": workload_small, "lorenz-uneven": workload_uneven, "glv16": workload_glv,
          "everystep": workload_everystep}[args.workload]
    print(json.dumps(fn(args)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
