"""Synthetic ensemble inputs from a counter-based generator.

Every random number is Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel
random numbers: as easy as 1, 2, 3", SC'11) of the counter
(traj mod 2^32, traj >> 32, var, 0) under the key (seed mod 2^32, seed >> 32),
so any subset of an ensemble regenerates identically.  From the four output
words: U53 = ((w0 >> 5) * 2^26 + (w1 >> 6)) / 2^53 in [0, 1), and a second
uniform from (w2, w3).  Normals are Box-Muller of the two uniforms
(sqrt(-2 log(1 - u1)) cos(2 pi u2)); LogUniform(a, b) = exp(U(log a, log b)).

The workloads below are the ensembles the benchmark and tests use.
"""
import numpy as np

_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint64(0x9E3779B9)
_W1 = np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Vectorised Philox4x32 on uint64 arrays holding 32-bit words."""
    c0, c1, c2, c3 = (np.asarray(x, dtype=np.uint64) & _MASK for x in (c0, c1, c2, c3))
    k0 = np.uint64(int(k0) & 0xFFFFFFFF)
    k1 = np.uint64(int(k1) & 0xFFFFFFFF)
    for _ in range(rounds):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = (hi1 ^ c1 ^ k0) & _MASK, lo1, (hi0 ^ c3 ^ k1) & _MASK, lo0
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return c0, c1, c2, c3


def uniforms(seed, traj, var):
    """Two U[0,1) doubles per (traj, var)."""
    traj = np.asarray(traj, dtype=np.uint64)
    w0, w1, w2, w3 = philox4x32(traj & _MASK, traj >> np.uint64(32), np.full(traj.shape, var),
                                np.zeros(traj.shape), seed & 0xFFFFFFFF, seed >> 32)
    s = 1.0 / 9007199254740992.0
    u1 = ((w0 >> np.uint64(5)).astype(np.float64) * 67108864.0 + (w1 >> np.uint64(6)).astype(np.float64)) * s
    u2 = ((w2 >> np.uint64(5)).astype(np.float64) * 67108864.0 + (w3 >> np.uint64(6)).astype(np.float64)) * s
    return u1, u2


def uniform(seed, traj, var, lo, hi):
    u, _ = uniforms(seed, traj, var)
    return lo + (hi - lo) * u


def normal(seed, traj, var, mu=0.0, sigma=1.0):
    u1, u2 = uniforms(seed, traj, var)
    return mu + sigma * np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def loguniform(seed, traj, var, a, b):
    return np.exp(uniform(seed, traj, var, np.log(a), np.log(b)))


def lorenz_small(n=1024, seed=0, start=0):
    """Lorenz, sigma = 10, beta = 8/3, rho ~ U(0, 21); u0 = (1, 0, 0); tspan (0, 1).
    Returns u0 (3, n), p (3, n)."""
    i = np.arange(start, start + n, dtype=np.uint64)
    p = np.empty((3, n))
    p[0] = 10.0
    p[1] = uniform(seed, i, 0, 0.0, 21.0)
    p[2] = 8.0 / 3.0
    u0 = np.zeros((3, n))
    u0[0] = 1.0
    return u0, p


def lorenz_wide(n=1 << 20, seed=1, start=0, rho_hi=21.0):
    """Lorenz, sigma ~ U(8,12), rho ~ U(0,rho_hi), beta ~ U(2,3), u0 ~ U(-1,1)^3."""
    i = np.arange(start, start + n, dtype=np.uint64)
    p = np.empty((3, n))
    p[0] = uniform(seed, i, 0, 8.0, 12.0)
    p[1] = uniform(seed, i, 1, 0.0, rho_hi)
    p[2] = uniform(seed, i, 2, 2.0, 3.0)
    u0 = np.empty((3, n))
    for c in range(3):
        u0[c] = uniform(seed, i, 3 + c, -1.0, 1.0)
    return u0, p


def lorenz_uneven(n=1 << 21, seed=2, start=0):
    """sigma = 10, beta = 8/3, rho ~ U(0, 60), t_end ~ LogUniform(0.5, 20), u0 ~ N(0,1)^3.
    Returns u0 (3, n), p (3, n), tspan (2, n)."""
    i = np.arange(start, start + n, dtype=np.uint64)
    p = np.empty((3, n))
    p[0] = 10.0
    p[1] = uniform(seed, i, 0, 0.0, 60.0)
    p[2] = 8.0 / 3.0
    tspan = np.zeros((2, n))
    tspan[1] = loguniform(seed, i, 1, 0.5, 20.0)
    u0 = np.empty((3, n))
    for c in range(3):
        u0[c] = normal(seed, i, 2 + c)
    return u0, p, tspan


def glv16(n=1 << 17, seed=3, start=0):
    """GLV, 16 species: r_i ~ U(0.5, 1.5), u0_i ~ U(0.1, 1) per trajectory; one shared
    interaction matrix A with A_ij ~ N(0, 0.1^2) off the diagonal and A_ii = -1.
    Returns u0 (16, n), r (16, n), A (256,) row-major."""
    i = np.arange(start, start + n, dtype=np.uint64)
    r = np.empty((16, n))
    u0 = np.empty((16, n))
    for k in range(16):
        r[k] = uniform(seed, i, k, 0.5, 1.5)
        u0[k] = uniform(seed, i, 16 + k, 0.1, 1.0)
    idx = np.arange(256, dtype=np.uint64)
    A = normal(seed, idx, 100, 0.0, 0.1)
    A[np.arange(16) * 17] = -1.0
    return u0, r, A


def lorenz_costmodel(n, seed, start=0):
    """The cost-model distribution: the wide Lorenz ensemble with rho ~ U(0, 28) and
    t_end ~ LogUniform(0.5, 10).  Returns u0, p, tspan."""
    u0, p = lorenz_wide(n, seed=seed, start=start, rho_hi=28.0)
    i = np.arange(start, start + n, dtype=np.uint64)
    tspan = np.zeros((2, n))
    tspan[1] = loguniform(seed, i, 6, 0.5, 10.0)
    return u0, p, tspan
