"""The oracle's error norm and step-size controller against a model of them in
exact arithmetic (mpmath, 40 digits) that shares no code with the oracle.

The problem is u' = diag(l1, l2) u (the LINEAR2 right-hand side with zero
off-diagonal entries), for which one Runge-Kutta attempt is closed-form: with
z = l dt, the stage vector is Y(z) = (I - z A)^-1 1, the step multiplies u by
R(z) = 1 + z b.Y(z) and the embedded error estimate dt * sum(btilde_j k_j) is
u E(z), E(z) = z btilde.Y(z).  The model then applies the reference's controller
to those numbers (gpuatsit5.jl:272-299, gpuvern7.jl:397-426, gpuvern9.jl:556-595,
constants src/SimpleDiffEq.jl:67-77):

    EEst = sqrt(sum_i (err_i / (abstol + max(|uprev_i|, |u_i|) reltol))^2 / N)
    q11 = EEst^beta1;  q = EEst == 0 ? 1/qmax : q11 / qold^beta2
    reject (EEst > 1): dt = dt / min(1/qmin, q11/gamma)
    accept: q = max(1/qmax, min(1/qmin, q/gamma)); qold = max(EEst, qoldinit);
            dt_new = min(|dt/q|, |tf - t - dt|); t = tf - t - dt < snap ? tf : t + dt

and predicts every accepted time, state and the reject count, which the oracle
must reproduce (save_everystep output) to within the rounding of its
binary64 error estimate (modelled per attempt; below 1e-5 over the first steps
of every scenario, against mutant differences at the percent level).  The scenarios are built so that
the first attempt's error sits where a plausible mistake changes the outcome:
RMS <= 1 < max-norm, mean-abs < 1 < RMS, huge error (the 1/qmin clamp),
tiny error (the 1/qmax clamp and the qoldinit floor), growth and decay (which of
|uprev|, |u| scales the error).  test_every_mutant_is_caught runs the same model
with each such mistake and checks that the oracle's output rejects it."""
import mpmath as mp
import numpy as np
import pytest

from oracle import pyoracle
from simplediffeq_hip import _abi, runtime
from tableaus import tsit5_method, vern7_method, vern9_method

mp.mp.dps = 40

# src/SimpleDiffEq.jl:67-77 (binary64 values of the literals)
BETA1, BETA2, QMAX, QMIN, GAMMA, QOLDINIT = 7 / 50, 2 / 25, 10.0, 1 / 5, 9 / 10, 1.0e-4

METHODS = {
    _abi.ALG_ATSIT5: (tsit5_method, 1e-14, 1e-14),    # dtmin, snap (gpuatsit5.jl:256,294)
    _abi.ALG_AVERN7: (vern7_method, 1e-14, 1e-14),    # gpuvern7.jl:383,421
    _abi.ALG_AVERN9: (vern9_method, float(np.float32(1e-7)), float(np.float32(1e-7))),  # 1f-7
}
NAMES = {_abi.ALG_ATSIT5: "ATsit5", _abi.ALG_AVERN7: "AVern7", _abi.ALG_AVERN9: "AVern9"}


def rk_polys(method, z):
    """(R(z), E(z)) of one explicit step of `method` on u' = l u, z = l dt."""
    c, a, b, bt, s = method
    Y = {}
    for i in range(1, s + 1):
        Y[i] = 1 + z * mp.fsum(mp.mpf(v) * Y[j] for (ii, j), v in a.items() if ii == i)
    R = 1 + z * mp.fsum(mp.mpf(b[j]) * Y[j] for j in b)
    E = z * mp.fsum(mp.mpf(bt[j]) * Y[j] for j in bt)
    return R, E


def norm(e, kind):
    if kind == "rms":
        return mp.sqrt(mp.fsum(x * x for x in e) / len(e))
    if kind == "max":
        return max(abs(x) for x in e)
    if kind == "mean_abs":
        return mp.fsum(abs(x) for x in e) / len(e)
    if kind == "no_len":
        return mp.sqrt(mp.fsum(x * x for x in e))
    raise ValueError(kind)


def model(alg, lam, u0, tf, dt0, abstol, reltol, mut=None, max_attempts=400):
    """Exact-arithmetic solve of u' = diag(lam) u from t = 0.  `mut` names one
    deliberate mistake (see MUTANTS).  Returns (ts, us, nreject, EEst of every
    attempt, |EEst - 1| margin)."""
    mk, dtmin, snap = METHODS[alg]
    method = mk()
    beta1, beta2 = mp.mpf(BETA1), mp.mpf(BETA2)
    if mut == "betas_swapped":
        beta1, beta2 = beta2, beta1
    gamma = mp.mpf(1) if mut == "no_gamma" else mp.mpf(GAMMA)
    inv_qmax, inv_qmin, qoldinit = 1 / mp.mpf(QMAX), 1 / mp.mpf(QMIN), mp.mpf(QOLDINIT)
    kind = {"max_norm": "max", "mean_abs_norm": "mean_abs", "norm_without_len": "no_len"}.get(
        mut, "rms")
    lam = [mp.mpf(x) for x in lam]
    u = [mp.mpf(x) for x in u0]
    t, tf, dt, qold = mp.mpf(0), mp.mpf(tf), mp.mpf(dt0), qoldinit
    abstol, reltol = mp.mpf(abstol), mp.mpf(reltol)
    ts, us, nrej, hist = [t], [list(u)], 0, []
    # The binary64 error estimate is a sum whose weights add to zero; its relative
    # rounding is about 2 eps (1 + sum|btilde|) |z| / |E(z)| (stage values of size
    # |u|).  tols[k] bounds how far that moves the k-th accepted time (relative):
    # EEst^(7/50) carries 0.14 of it into every later step size.
    bsum = 1 + sum(abs(v) for v in method[3].values())
    # (Clamped factors and the qoldinit floor carry none of it.)
    drift, tols, r_qold = mp.mpf(0), [mp.mpf(1e-12)], mp.mpf(0)
    attempts = 0
    while t < tf:
        uprev = u
        while True:
            assert dt >= dtmin, "model: dt < dtmin"
            attempts += 1
            assert attempts <= max_attempts
            RE = [rk_polys(method, l * dt) for l in lam]
            unew = [up * R for up, (R, E) in zip(uprev, RE)]
            err = [up * E for up, (R, E) in zip(uprev, RE)]
            if mut == "scale_by_uprev":
                m = [abs(x) for x in uprev]
            elif mut == "scale_by_u":
                m = [abs(x) for x in unew]
            else:
                m = [max(abs(x), abs(y)) for x, y in zip(uprev, unew)]
            e = [er / (abstol + mm * reltol) for er, mm in zip(err, m)]
            EEst = norm(e, kind)
            hist.append(EEst)
            z1 = lam[0] * dt
            r = 2 * mp.mpf(2) ** -53 * bsum * abs(z1) / max(abs(RE[0][1]), mp.mpf(1e-300))
            q11 = EEst ** beta1
            q = inv_qmax if EEst == 0 else q11 / qold ** beta2
            if EEst > 1:
                if q11 / gamma < inv_qmin:  # unclamped: EEst's rounding reaches dt
                    drift += mp.mpf(0.14) * r
                if mut == "reject_uses_q":
                    dt = dt / min(inv_qmin, q / gamma)
                else:
                    dt = dt / min(inv_qmin, q11 / gamma)
                nrej += 1
                continue
            if EEst != 0 and inv_qmax < q / gamma < inv_qmin:
                drift += mp.mpf(0.14) * r + mp.mpf(0.08) * r_qold
            r_qold = r if EEst > qoldinit else 0
            q = max(inv_qmax, min(inv_qmin, q / gamma))
            if mut == "qold_not_floored":
                qold = EEst
            elif mut == "qold_never_updated":
                pass
            else:
                qold = max(EEst, qoldinit)
            dtold = dt
            dt = min(abs(dt / q), abs(tf - t - dtold))
            t = tf if tf - t - dtold < snap else t + dtold
            u = unew
            ts.append(t)
            us.append(list(u))
            tols.append(mp.mpf(1e-12) * len(ts) + 10 * drift)
            break
    margin = min(abs(h - 1) for h in hist)
    return ts, us, nrej, hist, margin, tols


def pick_dt(alg, lam, level):
    """dt at which the embedded error of one step is `level` of the step's
    change, |E(z)| = level |z| (bisection in z): large enough that the
    binary64 error estimate, a sum whose weights add to zero, is far above its
    rounding (relative rounding about 1e-15 / level)."""
    method = METHODS[alg][0]()
    lo, hi = mp.mpf("1e-3"), mp.mpf(6)
    for _ in range(80):
        mid = (lo + hi) / 2
        z = mid if lam > 0 else -mid
        if abs(rk_polys(method, z)[1]) < level * mid:
            lo = mid
        else:
            hi = mid
    return float(lo / abs(lam))


def design(alg, lam, u0, dt0, target, rel_share=0.05):
    """abstol, reltol (binary64) that put the first attempt's scaled error of
    component 1 at `target`, reltol carrying `rel_share` of the scale."""
    mk = METHODS[alg][0]
    R, E = rk_polys(mk(), mp.mpf(lam[0]) * mp.mpf(dt0))
    err = abs(mp.mpf(u0[0]) * E)
    m = max(abs(mp.mpf(u0[0])), abs(mp.mpf(u0[0]) * R))
    reltol = float(rel_share * err / (target * m))
    abstol = float(err / target - m * mp.mpf(reltol))
    assert abstol > 0 and reltol > 0
    return abstol, reltol


# (name, lam, u0, first-attempt target of component 1's scaled error, dt0 divisor,
# span in units of dt0).  dt0 = pick_dt(lam1, 1e-2) / divisor.
SCENARIOS = [
    ("rms_accepts_max_rejects", (1.5, 0.15), (1.0, -0.7), 1.2, 1, 10),
    ("mean_abs_accepts_rms_rejects", (-2.0, -0.2), (0.8, 1.1), 1.7, 1, 10),
    ("huge_error_qmin_clamp", (-3.0, 0.3), (1.0, 0.5), 1e5, 1, 6),
    ("tiny_error_qmax_clamp_qoldinit", (-0.5, 0.05), (1.0, 1.0), 1e-12, 20, 60),
    ("moderate_growth", (0.9, -0.09), (-0.6, 0.4), 0.3, 1, 12),
]


def scenario(alg, sc):
    name, lam, u0, target, div, span = sc
    dt0 = pick_dt(alg, lam[0], 1e-2) / div
    abstol, reltol = design(alg, lam, u0, dt0, target)
    return dict(lam=lam, u0=u0, tf=span * dt0, dt0=dt0, abstol=abstol, reltol=reltol)


def oracle_run(alg, s, max_saves=512):
    p = np.array([[s["lam"][0]], [0.0], [0.0], [s["lam"][1]]])
    u0 = np.array(s["u0"], dtype=np.float64).reshape(2, 1)
    d = runtime.make_desc(alg, rhs=_abi.RHS_LINEAR2, n_traj=1, tspan=(0.0, s["tf"]),
                          dt=s["dt0"], abstol=s["abstol"], reltol=s["reltol"],
                          save_everystep=True, max_saves=max_saves)
    b = runtime.EnsembleBuffers(d, 2, u0, p)
    st = pyoracle.solve(d, b, nthreads=1)
    assert st == 0
    k = int(b.n_steps[0])
    return dict(lam=s["lam"], ts=b.t_steps[:k, 0].copy(), us=b.u_steps[:k, :, 0].copy(),
                nacc=int(b.naccept[0]), nrej=int(b.nreject[0]), t_end=float(b.t_end[0]))


def agrees(o, m, tols, factor=1.0):
    """The oracle's accepted times and states equal the model's within
    factor * tols[k] (relative), with the same accepted and rejected counts."""
    ts, us, nrej = m[0], m[1], m[2]
    if o["nrej"] != nrej or len(o["ts"]) != len(ts):
        return False
    lam = max(abs(o["lam"][0]), abs(o["lam"][1]))
    for k in range(len(ts)):
        tol = factor * float(tols[k])
        tk = float(ts[k])
        if abs(o["ts"][k] - tk) > tol * max(1.0, abs(tk)):
            return False
        for c in range(2):  # a time error dt moves u = e^(l t) u0 by l dt relative
            ref = float(us[k][c])
            if abs(o["us"][k, c] - ref) > tol * (1 + lam * abs(tk)) * max(abs(ref), 1e-300):
                return False
    return True


CASES = [(alg, sc) for alg in METHODS for sc in SCENARIOS]
IDS = ["%s-%s" % (NAMES[alg], sc[0]) for alg, sc in CASES]


@pytest.mark.parametrize("alg,sc", CASES, ids=IDS)
def test_oracle_follows_exact_controller(alg, sc):
    s = scenario(alg, sc)
    m = model(alg, **s)
    ts, us, nrej, hist, margin, tols = m
    # the rounding bound of the first accepted steps, where the scenario's designed
    # error acts, stays far below the percent-level differences the mutants make
    assert max(tols[:4]) < 1e-5
    # the scenario is what it says: first attempt at the designed error, and no
    # decision within rounding of the accept/reject boundary
    assert abs(hist[0] / norm([mp.mpf(sc[3]), 0], "rms") - 1) < 0.05 or sc[3] < 1e-6
    assert margin > 1e-6
    o = oracle_run(alg, s)
    assert o["nrej"] == nrej and o["nacc"] == len(ts) - 1
    assert o["t_end"] == s["tf"]
    assert agrees(o, m, tols), (o["ts"], [float(x) for x in ts])


def test_first_attempts_discriminate_the_norm():
    # RMS <= 1 < max-norm (and < the norm without /N); mean-abs < 1 < RMS
    e1 = [mp.mpf(1.2), mp.mpf(1e-4)]
    assert norm(e1, "rms") <= 1 < min(norm(e1, "max"), norm(e1, "no_len"))
    e2 = [mp.mpf(1.7), mp.mpf(1e-4)]
    assert norm(e2, "mean_abs") < 1 < norm(e2, "rms")
    # the tiny-error scenario's first step is clamped at 1/qmax (dt grows 10x, not 9x)
    q = (mp.mpf(1e-12) ** mp.mpf(BETA1)) / mp.mpf(QOLDINIT) ** mp.mpf(BETA2) / mp.mpf(GAMMA)
    assert q < 1 / mp.mpf(QMAX)
    # the huge-error scenario's rejection is clamped at 1/qmin (dt shrinks 5x)
    assert mp.mpf(1e7) ** mp.mpf(BETA1) / mp.mpf(GAMMA) > 1 / mp.mpf(QMIN)


MUTANTS = ["max_norm", "mean_abs_norm", "norm_without_len", "scale_by_uprev", "scale_by_u",
           "betas_swapped", "no_gamma", "reject_uses_q", "qold_not_floored",
           "qold_never_updated"]


@pytest.mark.parametrize("alg", list(METHODS), ids=[NAMES[a] for a in METHODS])
def test_every_mutant_is_caught(alg):
    # each deliberate mistake in the model changes the predicted accepted times or
    # reject count of at least one scenario by far more (100x) than rounding explains
    runs = []
    for sc in SCENARIOS:
        s = scenario(alg, sc)
        runs.append((s, oracle_run(alg, s), model(alg, **s)[5]))
    for mut in MUTANTS:
        caught = False
        for s, o, tols in runs:
            try:
                m = model(alg, mut=mut, **s)
            except AssertionError:  # the mistake drives dt below dtmin or runs away
                caught = True
                break
            n = min(len(m[0]), len(tols))
            if m[2] != o["nrej"] or len(m[0]) != len(o["ts"]) or \
                    not agrees(o, (m[0][:n], m[1][:n], m[2]), tols[:n], factor=100.0):
                caught = True
                break
        assert caught, mut


@pytest.mark.skipif(not __import__("os").path.isdir("/root/reference"),
                    reason="reference tree not present")
def test_controller_constants_match_reference_text():
    # build_adaptive_controller_cache, src/SimpleDiffEq.jl:67-77: the six literals
    import os
    import re
    text = open("/root/reference/src/SimpleDiffEq.jl").read()
    want = {"beta1": BETA1, "beta2": BETA2, "qmax": QMAX, "qmin": QMIN, "gamma": GAMMA,
            "qoldinit": QOLDINIT}
    for name, val in want.items():
        m = re.search(r"\b%s = T\(([^)]*)\)" % name, text)
        assert m, name
        assert eval(m.group(1), {"__builtins__": {}}) == val, name  # e.g. "7 / 50"
    # and the same literals in the kernels and the oracle
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    for path in ("simplediffeq.jl_amd/csrc/sdeh_integrators.h", "oracle/oracle.hpp"):
        src = open(os.path.join(root, path)).read()
        for name, val in want.items():
            m = re.search(r"\b%s = T\(([^)]*)\)" % name, src)
            assert m, (path, name)
            assert eval(m.group(1), {"__builtins__": {}}) == val, (path, name)
