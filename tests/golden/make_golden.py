#!/usr/bin/env python3
"""Write tests/golden/*.json: high-accuracy endpoint and saveat values for the
test problems, from scipy's DOP853 at rtol = atol = 1e-13 (dense output for
saveat).  They stand in for the OrdinaryDiffEq Vern9 solutions the reference's
tests compare against (test/gpu_ode_regression.jl:29-33,
test/gpusimpleatsit5_tests.jl:41,75), which are not available here.

Inputs come from simplediffeq_hip.inputs (Philox), so the fixtures are
regenerable; run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np
from scipy.integrate import solve_ivp

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "simplediffeq.jl_amd"))
from simplediffeq_hip import inputs  # noqa: E402

TOL = dict(method="DOP853", rtol=1e-13, atol=1e-13)


def lorenz(t, u, s, r, b):
    return [s * (u[1] - u[0]), u[0] * (r - u[2]) - u[1], u[0] * u[1] - b * u[2]]


def main():
    out = {}
    # Lorenz small ensemble (sigma 10, beta 8/3, rho ~ U(0,21), u0 (1,0,0), tspan (0,1)): first 32
    u0, p = inputs.lorenz_small(32, seed=0)
    ends = []
    for i in range(32):
        sol = solve_ivp(lorenz, (0.0, 1.0), u0[:, i], args=tuple(p[:, i]), **TOL)
        ends.append(sol.y[:, -1].tolist())
    out["lorenz_small"] = {"n": 32, "seed": 0, "tspan": [0.0, 1.0], "u_end": ends}
    # Lorenz wide ensemble (tspan (0,10)): first 16
    u0, p = inputs.lorenz_wide(16, seed=1)
    ends = []
    for i in range(16):
        sol = solve_ivp(lorenz, (0.0, 10.0), u0[:, i], args=tuple(p[:, i]), **TOL)
        ends.append(sol.y[:, -1].tolist())
    out["lorenz_wide"] = {"n": 16, "seed": 1, "tspan": [0.0, 10.0], "u_end": ends}
    # The reference's Lorenz problem (gpusimpleatsit5_tests.jl:24-27,73): u0 = 10 ones,
    # p = (10, 28, 8/3), tspan (0, 10); saveat 0:0.1:10 (point values)
    ts = [i / 10 for i in range(101)]
    sol = solve_ivp(lorenz, (0.0, 10.0), [10.0, 10.0, 10.0], args=(10.0, 28.0, 8.0 / 3.0),
                    t_eval=ts, **TOL)
    out["lorenz_reference"] = {"u0": [10.0, 10.0, 10.0], "p": [10.0, 28.0, 8.0 / 3.0],
                               "tspan": [0.0, 10.0], "saveat": ts, "u": sol.y.T.tolist()}
    # GLV-16 (shared A), first 4 trajectories, 9 save points on (0, 50)
    u0, r, A = inputs.glv16(4, seed=3)
    Am = A.reshape(16, 16)
    ts = np.linspace(0.0, 50.0, 9).tolist()
    us = []
    for i in range(4):
        f = lambda t, u, rr=r[:, i]: u * (rr + Am @ u)
        sol = solve_ivp(f, (0.0, 50.0), u0[:, i], t_eval=ts, **TOL)
        us.append(sol.y.T.tolist())
    out["glv16"] = {"n": 4, "seed": 3, "tspan": [0.0, 50.0], "saveat": ts, "u": us}
    with open(os.path.join(HERE, "dop853_fixtures.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("wrote dop853_fixtures.json")


if __name__ == "__main__":
    main()
