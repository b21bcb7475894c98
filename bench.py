#!/usr/bin/env python3
"""Benchmark: adaptive ensemble throughput on one MI355X.

Workload (one "step" = one ensemble solve): Lorenz, 1,048,576 trajectories,
sigma ~ U(8,12), rho ~ U(0,21), beta ~ U(2,3), u0 ~ U(-1,1)^3 (Philox seed 1),
tspan (0, 10), GPUSimpleATsit5 with abstol 1e-6, reltol 1e-3 and the reference
default dt0 = 0.1f0, endpoint output only.  Inputs are resident in HBM before
the timed region; each step runs the integrator kernel over the whole ensemble.

value = trajectories/s over the K timed steps (binary64 state, the headline);
the binary32 run of the same ensemble is reported beside it.

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

# Algorithmic FLOPs (per component per attempt, RHS calls per attempt) by solver.
STAGE_FLOPS = {0: 48, 1: 67, 2: 83, 3: 117, 4: 167, 5: 211}
RHS_PER_ATTEMPT = {0: 6, 1: 6, 2: 9, 3: 9, 4: 15, 5: 15}
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


def cpu_baseline(u0, p, desc_kwargs, sample, threads):
    from oracle import pyoracle
    from simplediffeq_hip import _abi, runtime
    d = runtime.make_desc(n_traj=sample, **desc_kwargs)
    bufs = runtime.EnsembleBuffers(d, 3, u0[:, :sample], p[:, :sample])
    pyoracle.solve(d, bufs, nthreads=threads)  # warm-up (page in, thread start)
    t0 = time.perf_counter()
    st = pyoracle.solve(d, bufs, nthreads=threads)
    el = time.perf_counter() - t0
    if st != 0:
        raise RuntimeError("oracle status %d" % st)
    return sample / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=N_TRAJ)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-f32", action="store_true", help="skip the binary32 run")
    ap.add_argument("--cpu-sample", type=int, default=1 << 17)
    ap.add_argument("--tune", action="store_true", help="autotune the launch shape first")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; --gpus %d is not supported "
              "(trajectories are independent, see DESIGN.md for how the ensemble shards)"
              % args.gpus, file=sys.stderr)
        return 2

    import numpy as np
    import torch
    from simplediffeq_hip import _abi, inputs, runtime

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
        d = runtime.make_desc(dtype=dtype, n_traj=n, **common)
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
        results[dtype_name] = dict(
            wall_s=wall, ms_per_step=wall * 1e3 / args.steps, kernel_ms=dev_ms,
            traj_per_s=n * args.steps / wall, attempts=st["attempts"], accepts=st["accepts"],
            flops=flops, tflops=flops / (dev_ms * 1e-3) / 1e12, bytes=alg_bytes,
            gbs=alg_bytes / (dev_ms * 1e-3) / 1e9, shape=dict(block=st["block"], grid=st["grid"],
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
                     "traffic": None,
                     "kernel": "sdeh_adaptive_kernel<double, Lorenz, ATsit5, endpoint>",
                     "kernel_ms": r64["kernel_ms"], "algorithmic_flops_per_launch": r64["flops"],
                     "attempts_per_launch": r64["attempts"],
                     "hbm_gbs": r64["gbs"], "hbm_frac": r64["gbs"] / HBM_PEAK_GBS},
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
        rate, el = cpu_baseline(u0, p, kw, sample, threads)
        out["cpu_baseline"] = {"value": rate, "unit": "trajectories/s", "cores": threads,
                               "kind": "port",
                               "sample": "first %d trajectories of the same ensemble, binary64, "
                                         "%.1f s" % (sample, el)}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
