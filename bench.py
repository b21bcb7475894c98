#!/usr/bin/env python3
"""Benchmark: adaptive ensemble throughput on one MI355X.

Workload (one "step" = one ensemble solve): Lorenz, 1,048,576 trajectories,
sigma ~ U(8,12), rho ~ U(0,21), beta ~ U(2,3), u0 ~ U(-1,1)^3 (Philox seed 1),
tspan (0, 10), GPUSimpleATsit5 with abstol 1e-6, reltol 1e-3 and the reference
default dt0 = 0.1f0, endpoint output only.  Inputs are resident in HBM before
the timed region; each step runs the integrator kernel over the whole ensemble.

value = trajectories/s over the K timed steps (binary64 state, the headline);
the binary32 run of the same ensemble is reported beside it.  Before the warmup,
the library's autotuner picks each variant's launch shape (schedule, block, grid,
queue chunk) by timing candidate launches of the same solve; the chosen shape is
in config.launch.

roofline: the integrator kernel is bound by vector (VALU) binary64 throughput:
achieved = algorithmic FLOPs per launch / average launch time (HIP events on
the launch stream over the timed region), peak = 78.6 TFLOP/s.  Algorithmic
FLOPs per launch = sum over trajectories of (accepted + rejected attempts) x
(3 * 67 + 6 * 8 + 12) + 8 (the initial FSAL evaluation): 67 FLOPs of stage,
error and scaling arithmetic per component per attempt, 6 Lorenz evaluations of
8 FLOPs, 2 for the norm and 10 for the controller; FMA = 2, add/mul/div/sqrt/pow = 1.

cpu_baseline: the CPU restatement of the same solver (oracle/, one thread per
host core, contiguous chunks as EnsembleThreads does) on a bounded sample of the
same ensemble.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "simplediffeq.jl_amd"))

N_TRAJ = 1 << 20
PEAK_TFLOPS = {"f64": 78.6, "f32": 157.3}
HBM_PEAK_GBS = 8000.0

# HBM bytes per solve of the binary64 headline at 2^20 trajectories, by schedule,
# from two rocprofv3 --pmc passes (FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000, factors
# from a 256 MiB device copy in the same run).  Persistent: 49,405 KiB read + 82,225
# KiB written.  Staged (first pass, key sort, resumed pass): 234,824 KiB read +
# 389,626 KiB written, mostly the parked records and the resumed pass's endpoint
# stores in hand-out order (partial lines).  Counters cannot be read from inside a
# timed run, so each figure applies only to the configuration it was measured on.
MEASURED_TRAFFIC = {("f64", 1 << 20, 2): 134.79e6, ("f64", 1 << 20, 3): 639.44e6}

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


def cpu_leg(desc, n_state, u0, p, ps=None, tspan=None, min_seconds=3.0):
    """The `cpu_baseline` record of a workload (oracle on the host cores)."""
    threads = min(16, len(os.sched_getaffinity(0)))
    rate, el, reps = cpu_time(desc, n_state, u0, p, ps, tspan, threads, min_seconds)
    return {"value": rate, "unit": "trajectories/s", "cores": threads, "kind": "port",
            "sample": "first %d trajectories, solved %d times in %.1f s on %s"
                      % (int(desc.n_traj), reps, el, cpu_model())}


def cpu_baseline(u0, p, desc_kwargs, sample, threads, min_seconds=10.0, max_reps=400):
    """Oracle solves of the first `sample` trajectories, repeated until at least
    `min_seconds` of CPU time have elapsed; returns (trajectories/s, seconds, reps)."""
    from simplediffeq_hip import runtime
    d = runtime.make_desc(n_traj=sample, **desc_kwargs)
    return cpu_time(d, 3, u0[:, :sample], p[:, :sample], threads=threads,
                    min_seconds=min_seconds, max_reps=max_reps)


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
    for name, alg, dt in (("tsit5", _abi.ALG_TSIT5, 0.01), ("atsit5", _abi.ALG_ATSIT5, 0.1)):
        d = runtime.make_desc(alg, rhs=_abi.RHS_LORENZ, n_traj=1024, tspan=(0.0, 1.0), dt=dt,
                              abstol=1e-6, reltol=1e-3)
        g = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device="cuda:0")
        ms, wall = _time_solves(ctx, d, g, stream, args.steps, args.warmup)
        c = runtime.EnsembleBuffers(d, 3, u0, p)
        threads = min(16, len(os.sched_getaffinity(0)))
        t0 = time.perf_counter()
        for _ in range(5):
            pyoracle.solve(d, c, nthreads=threads)
        cpu = (time.perf_counter() - t0) / 5
        same = all(np.array_equal(g.numpy(f), c.numpy(f)) for f in ("u_end", "naccept", "nreject"))
        out[name] = {"gpu_kernel_ms": ms, "gpu_call_ms": wall * 1e3, "cpu_ms": cpu * 1e3,
                     "cpu_threads": threads, "bit_identical": bool(same)}
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
                        "launch": {k: best[k] for k in ("block", "grid", "chunk", "schedule")}}
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
    for name, alg, dtype, tol in (("avern9_f64", _abi.ALG_AVERN9, _abi.F64, 1e-10),
                                  ("atsit5_f32", _abi.ALG_ATSIT5, _abi.F32, 1e-6)):
        # default lane layout in both save layouts; binary64 also in each layout
        variants = [(_abi.SAVE_TRAJ_MAJOR, 0, ""), (_abi.SAVE_POINT_MAJOR, 0, "_point_major")]
        if dtype == _abi.F64:
            variants += [(_abi.SAVE_TRAJ_MAJOR, 16, "_lanes16"), (_abi.SAVE_TRAJ_MAJOR, 4, "_lanes4")]
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


def workload_costmodel(args):
    """Learned cost binning + dt0.  Train the 6-64-64-2 MLP on the GPU from the
    accepted-step counts of 2^20 plain binary64 ATsit5 solves (Lorenz, sigma~U(8,12),
    rho~U(0,28), beta~U(2,3), u0~U(-1,1)^3, t_end~LogUniform(0.5,10); seed 40) at
    abstol 1e-7, reltol 1e-5; evaluate on 2^22 fresh trajectories (seed 4):
    binary32 + compensated update, persistent schedule, binned by predicted cost and
    dt0-seeded, against the plain unsorted binary64 one-launch kernel.  Endpoint error
    of both against binary64 AVern9 at 1e-14 on a 65,536-trajectory subset."""
    import numpy as np
    import torch
    from simplediffeq_hip import _abi, costmodel, inputs, runtime
    n_train = min(args.n, 1 << 20) if args.n != N_TRAJ else 1 << 20
    n_eval = 1 << 22 if args.n == N_TRAJ else args.n
    ctx = runtime.Context(0)
    stream = torch.cuda.Stream()
    dev = "cuda:0"
    out = {"workload": "costmodel", "train": n_train, "eval": n_eval}
    tol = dict(abstol=1e-7, reltol=1e-5)
    f32_01 = float(np.float32(0.1))

    # 1. step logs of the plain solver
    u0, p, ts = inputs.lorenz_costmodel(n_train, seed=40)
    d = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=n_train, dt=f32_01, **tol)
    g = runtime.EnsembleBuffers(d, 3, u0, p, tspan=ts, backend="torch", device=dev)
    ctx.solve(d, g, sync=True)
    x = costmodel.features(p, u0, ts)
    y = costmodel.targets(g.numpy("naccept"), ts)
    del g
    # 2. train on the GPU (forward/backward/Adam kernels)
    cm = costmodel.CostModel(ctx, seed=0)
    t0 = time.perf_counter()
    hist = cm.fit(x, y, steps=args.train_steps, batch=65536, lr=3e-3, log_every=200)
    out["train_s"] = time.perf_counter() - t0
    out["loss"] = [hist[0][1], hist[-1][1]]

    # 3. evaluation ensemble, resident in HBM.  The binning step (features, MLP
    # predictions, stable sort by decreasing predicted cost, seeded dt0) runs on the
    # device on the bench stream; its time counts against the binned solves below.
    # One library launch (sdeh_mlp_plan_lorenz) does all but the sort; the same step
    # from torch-built features and the plain forward pass is timed beside it.
    u0, p, ts = inputs.lorenz_costmodel(n_eval, seed=4)
    u0d, pd, tsd = (torch.as_tensor(a, device=dev) for a in (u0, p, ts))
    torch.cuda.synchronize()

    def plan_unfused():
        with torch.cuda.stream(stream):
            xd = costmodel.features_device(pd, u0d, tsd)
        return cm.plan_device(xd, stream)

    def plan():
        return cm.plan_lorenz(pd, u0d, tsd, stream)

    def time_plan(fn):
        fn()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(5):
            res = fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 5, res

    plan_ms_unfused, (pred_u, order_u, dt0_u) = time_plan(plan_unfused)
    plan_ms, (pred_f, order, dt0) = time_plan(plan)
    out["plan_on_device_ms"] = plan_ms
    out["plan_on_device_unfused_ms"] = plan_ms_unfused
    out["fused_plan_equals_unfused"] = bool(
        torch.equal(pred_f, pred_u) and torch.equal(order, order_u) and torch.equal(dt0, dt0_u))
    # the same step with the predictions sorted on the host, for comparison
    t0 = time.perf_counter()
    pred_h = cm.predict(costmodel.features(p, u0, ts))
    order_h = cm.order_by_cost(pred_h)
    out["plan_on_host_s"] = time.perf_counter() - t0
    out["device_and_host_orders_equal"] = bool(np.array_equal(order.cpu().numpy(), order_h))

    def run(dtype, sched, precision, order_=None, dt0_=None, span=False):
        # "unsorted" runs hand trajectories out in index order (span_order 1); with
        # span=True the library's longest-span-first order applies (no model)
        dd = runtime.make_desc(_abi.ALG_ATSIT5, dtype=dtype, rhs=_abi.RHS_LORENZ, n_traj=n_eval,
                               dt=f32_01, schedule=sched, precision=precision,
                               span_order=0 if span else 1, **tol)
        gg = runtime.EnsembleBuffers(dd, 3, u0, p, tspan=ts, dt0=dt0_, order=order_,
                                     backend="torch", device=dev)
        ctx.solve(dd, gg, sync=True)
        st = ctx.stats()
        ms, _ = _time_solves(ctx, dd, gg, stream, args.steps, 1)
        return gg, ms, st

    base, ms_base, st_base = run(_abi.F64, _abi.SCHED_ONE_LAUNCH, _abi.PREC_PLAIN)
    out["f64_one_launch_unsorted"] = {"kernel_ms": ms_base, "traj_per_s": n_eval / (ms_base * 1e-3),
                                      "attempts": st_base["attempts"]}
    if not args.no_cpu:
        k = min(n_eval, 1 << 16)
        dc = runtime.make_desc(_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, n_traj=k, dt=f32_01, **tol)
        out["cpu_baseline"] = cpu_leg(dc, 3, u0[:, :k], p[:, :k], tspan=ts[:, :k])
    rec = {}
    for name, dtype, sched, prec, od, dz, sp in (
            ("f64_persistent_unsorted", _abi.F64, _abi.SCHED_PERSISTENT, _abi.PREC_PLAIN, None, None, False),
            ("f64_one_launch_span_order", _abi.F64, _abi.SCHED_ONE_LAUNCH, _abi.PREC_PLAIN, None, None, True),
            ("f32k_one_launch_span_order", _abi.F32, _abi.SCHED_ONE_LAUNCH, _abi.PREC_KAHAN, None, None, True),
            ("f64_persistent_binned", _abi.F64, _abi.SCHED_PERSISTENT, _abi.PREC_PLAIN, order, None, False),
            ("f32k_persistent_unsorted", _abi.F32, _abi.SCHED_PERSISTENT, _abi.PREC_KAHAN, None, None, False),
            ("f32k_one_launch_binned", _abi.F32, _abi.SCHED_ONE_LAUNCH, _abi.PREC_KAHAN, order, None, False),
            ("f32k_persistent_binned", _abi.F32, _abi.SCHED_PERSISTENT, _abi.PREC_KAHAN, order, None, False),
            ("f32k_persistent_binned_dt0", _abi.F32, _abi.SCHED_PERSISTENT, _abi.PREC_KAHAN, order, dt0, False)):
        gg, ms, st = run(dtype, sched, prec, od, dz, sp)
        rec[name] = gg
        out[name] = {"kernel_ms": ms, "traj_per_s": n_eval / (ms * 1e-3), "attempts": st["attempts"],
                     "speedup_vs_f64_one_launch": ms_base / ms}
        if od is not None:  # the binning step is part of the binned solve's cost
            out[name]["with_plan_ms"] = ms + plan_ms
            out[name]["speedup_with_plan"] = ms_base / (ms + plan_ms)
    # binning leaves every trajectory's result unchanged
    out["binned_bit_identical"] = bool(
        np.array_equal(rec["f64_persistent_binned"].numpy("u_end"), base.numpy("u_end")) and
        np.array_equal(rec["f32k_persistent_binned"].numpy("u_end"),
                       rec["f32k_persistent_unsorted"].numpy("u_end")))
    # 4. error against binary64 AVern9 at 1e-14 on a subset
    ns = min(65536, n_eval)
    dref = runtime.make_desc(_abi.ALG_AVERN9, rhs=_abi.RHS_LORENZ, n_traj=ns, dt=f32_01,
                             abstol=1e-14, reltol=1e-14)
    gref = runtime.EnsembleBuffers(dref, 3, u0[:, :ns], p[:, :ns], tspan=ts[:, :ns],
                                   backend="torch", device=dev)
    ctx.solve(dref, gref, sync=True)
    uref = gref.numpy("u_end").astype(np.float64)
    ok = gref.numpy("retcode") == 0

    def err(gg):
        ue = gg.numpy("u_end")[:, :ns].astype(np.float64)
        e = np.max(np.abs(ue - uref), axis=0) / np.max(np.abs(uref), axis=0)
        e = e[ok]
        return {"median": float(np.median(e)), "p90": float(np.percentile(e, 90))}
    out["error_vs_avern9_1e-14"] = {"f64_one_launch_unsorted": err(base),
                                    "f32k_persistent_binned": err(rec["f32k_persistent_binned"]),
                                    "f32k_persistent_binned_dt0": err(rec["f32k_persistent_binned_dt0"]),
                                    "reference_ok": int(ok.sum())}
    a_base = np.asarray(base.numpy("naccept"), dtype=np.int64) + base.numpy("nreject")
    out["wave_efficiency"] = {"unsorted": costmodel.wave_efficiency(a_base),
                              "binned": costmodel.wave_efficiency(a_base, order.cpu().numpy())}
    out["target_speedup_3x_met"] = out["f32k_persistent_binned_dt0"]["speedup_vs_f64_one_launch"] >= 3.0
    out["target_speedup_3x_met_with_plan"] = (
        max(v["speedup_with_plan"] for v in out.values()
            if isinstance(v, dict) and "speedup_with_plan" in v) >= 3.0)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_TRAJ)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-f32", action="store_true", help="skip the binary32 run")
    ap.add_argument("--cpu-sample", type=int, default=1 << 18)
    ap.add_argument("--copy-probe", action="store_true",
                    help="run one 256 MiB device copy first (reference kernel for the HBM "
                         "byte counters)")
    ap.add_argument("--no-tune", dest="tune", action="store_false",
                    help="skip the autotuner (by default sdeh_autotune picks the launch shape "
                         "once per variant, before the warmup and the timed steps)")
    ap.add_argument("--schedule", type=int, default=0, choices=[0, 1, 2, 3],
                    help="headline schedule: 0 library choice (autotuned), 1 one launch, "
                         "2 persistent, 3 staged")
    ap.add_argument("--grid", type=int, default=0,
                    help="persistent schedule: workgroups of 256 (0 = library default, "
                         "CUs x occupancy)")
    ap.add_argument("--workload", default="lorenz-wide",
                    choices=["lorenz-wide", "lorenz-small", "lorenz-uneven", "glv16", "costmodel",
                             "everystep"],
                    help="lorenz-wide is the headline line; the others report their own JSON")
    ap.add_argument("--train-steps", type=int, default=1500)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(trajectories are independent: an N-GPU run gives each GPU a contiguous slice)"
              % args.gpus, file=sys.stderr)
        return 2

    extra = {"lorenz-small": workload_small, "lorenz-uneven": workload_uneven,
             "glv16": workload_glv, "costmodel": workload_costmodel,
             "everystep": workload_everystep}
    if args.workload in extra:
        print(json.dumps(extra[args.workload](args)))
        return 0

    import numpy as np
    import torch
    from simplediffeq_hip import _abi, inputs, runtime

    if args.copy_probe:
        src = torch.ones(32 << 20, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    n = args.n
    u0, p = inputs.lorenz_wide(n, seed=1)
    dt0 = float(np.float32(0.1))
    common = dict(alg=_abi.ALG_ATSIT5, rhs=_abi.RHS_LORENZ, tspan=(0.0, 10.0), dt=dt0,
                  abstol=1e-6, reltol=1e-3, save_everystep=False)
    dev = torch.device("cuda:0")
    ctx = runtime.Context(0)
    stream = torch.cuda.Stream(device=dev)  # default priority
    sh = stream.cuda_stream

    results = {}
    for dtype_name, dtype in (("f64", _abi.F64), ("f32", _abi.F32)):
        if dtype_name == "f32" and args.no_f32:
            continue
        d = runtime.make_desc(dtype=dtype, n_traj=n, grid=args.grid, schedule=args.schedule,
                              **common)
        bufs = runtime.EnsembleBuffers(d, 3, u0, p, backend="torch", device=dev)
        torch.cuda.synchronize()
        if args.tune:
            ctx.autotune(d, bufs)
        ctx.solve(d, bufs, stream=sh, sync=True)  # counters of this (deterministic) workload
        st = ctx.stats()
        rc = bufs.numpy("retcode")
        if not (rc == 0).all():
            raise RuntimeError("%d trajectories did not finish" % int((rc != 0).sum()))
        for _ in range(args.warmup):
            ctx.solve(d, bufs, stream=sh, sync=False)
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            ctx.solve(d, bufs, stream=sh, sync=False)
        e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        dev_ms = e0.elapsed_time(e1) / args.steps
        flops = algorithmic_flops(_abi.ALG_ATSIT5, 3, LORENZ_FLOPS, st["attempts"], st["accepts"], n)
        es = 8 if dtype == _abi.F64 else 4
        alg_bytes = n * (3 * es + 3 * es + 3 * es + 8 + 12)  # u0, p, u_end, t_end, 3 int32
        host_ms = None
        if dtype == _abi.F64:
            # the same solve from host (pageable numpy) buffers: the library stages
            # inputs to HBM and outputs back, so this rate includes both PCIe copies
            hb = runtime.EnsembleBuffers(d, 3, u0, p)
            ctx.solve(d, hb)
            t1 = time.perf_counter()
            reps = 3
            for _ in range(reps):
                ctx.solve(d, hb)
            host_ms = (time.perf_counter() - t1) * 1e3 / reps
            del hb
        results[dtype_name] = dict(
            wall_s=wall, ms_per_step=wall * 1e3 / args.steps, kernel_ms=dev_ms,
            traj_per_s=n * args.steps / wall, attempts=st["attempts"], accepts=st["accepts"],
            flops=flops, tflops=flops / (dev_ms * 1e-3) / 1e12, bytes=alg_bytes,
            gbs=alg_bytes / (dev_ms * 1e-3) / 1e9, host_ms=host_ms,
            shape=dict(block=st["block"], grid=st["grid"],
                                                             chunk=st["chunk"],
                                                             schedule=st["schedule"]))
        del bufs

    r64 = results["f64"]
    out = {
        "metric": "adaptive trajectories/s at stated tol vs FP64 oracle; % FP32/FP64 VALU peak",
        "value": r64["traj_per_s"],
        "unit": "trajectories/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (Philox seed 1)",
        "config": {"workload": "Lorenz ensemble, GPUSimpleATsit5, abstol 1e-6, reltol 1e-3, "
                               "tspan (0,10), dt0 0.1f0, endpoint only",
                   "trajectories": n, "sigma": "U(8,12)", "rho": "U(0,21)", "beta": "U(2,3)",
                   "u0": "U(-1,1)^3", "launch": r64["shape"]},
        "roofline": {"bound": "valu", "achieved": r64["tflops"], "peak": PEAK_TFLOPS["f64"],
                     "unit": "TFLOP/s", "frac": r64["tflops"] / PEAK_TFLOPS["f64"],
                     "traffic": MEASURED_TRAFFIC.get(("f64", n, r64["shape"]["schedule"])),
                     "kernel": "sdeh_adaptive_kernel<double, Lorenz, ATsit5, endpoint>" + (
                         " (staged: first pass, key sort, resumed pass)"
                         if r64["shape"]["schedule"] == _abi.SCHED_STAGED else ""),
                     "kernel_ms": r64["kernel_ms"], "algorithmic_flops_per_launch": r64["flops"],
                     "attempts_per_launch": r64["attempts"],
                     "hbm_gbs": r64["gbs"], "hbm_frac": r64["gbs"] / HBM_PEAK_GBS},
        # not `value`: the rate when the caller hands over host buffers (PCIe both ways)
        "host_buffers": {"ms_per_solve": r64["host_ms"],
                         "traj_per_s": n / (r64["host_ms"] * 1e-3)},
    }
    if "f32" in results:
        r32 = results["f32"]
        out["f32"] = {"value": r32["traj_per_s"], "unit": "trajectories/s",
                      "kernel_ms": r32["kernel_ms"], "achieved_tflops": r32["tflops"],
                      "frac_of_157.3": r32["tflops"] / 157.3,
                      "frac_of_78.6_unpacked": r32["tflops"] / 78.6,
                      "attempts_per_launch": r32["attempts"], "launch": r32["shape"]}
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        sample = min(args.cpu_sample, n)
        kw = dict(common)
        kw["dtype"] = _abi.F64
        rate, el, reps = cpu_baseline(u0, p, kw, sample, threads)
        out["cpu_baseline"] = {"value": rate, "unit": "trajectories/s", "cores": threads,
                               "kind": "port",
                               "sample": "first %d trajectories of the same ensemble, binary64, "
                                         "solved %d times in %.1f s on %s"
                                         % (sample, reps, el, cpu_model())}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
