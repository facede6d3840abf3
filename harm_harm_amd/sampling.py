"""Seeded synthetic inputs for the batched-inversion workload.

2^24 zones by default: Lorentz factor 1-50, magnetisation b^2/rho 1e-4-1e3 and
u/rho 1e-6-1e2, all log-uniform, rho = 1, at random Kerr-Schild points
r in [1.3, 1000], theta in (0, pi), spin a in [0, 0.99].  Velocity and field
directions are isotropic in the local orthonormal frame of the spatial metric.
The primitive guess handed to the inversion is the true state perturbed by up
to +-20% per component (a stand-in for "the previous step's primitives").
"""
import ctypes
import numpy as np

from . import lib, G_NFIELD

_SYM = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (0, 3): 3, (1, 1): 4, (1, 2): 5, (1, 3): 6, (2, 2): 7, (2, 3): 8, (3, 3): 9}


def _s(a, b):
    return _SYM[(min(a, b), max(a, b))]


def ks_geometry(a, r, th):
    """SoA geometry records (27, n) from the library's host geometry."""
    n = a.size
    a = np.ascontiguousarray(a, dtype=np.float64)
    r = np.ascontiguousarray(r, dtype=np.float64)
    th = np.ascontiguousarray(th, dtype=np.float64)
    rec = np.empty((G_NFIELD, n), dtype=np.float64)
    lib().harm_geom_ks_batch(n, a.ctypes.data, r.ctypes.data, th.ctypes.data, rec.ctypes.data)
    return rec


def _unit_vectors(rng, geom):
    """Random directions with unit norm in the spatial metric g_ij."""
    n = geom.shape[1]
    G = np.empty((n, 3, 3))
    for i in range(3):
        for j in range(3):
            G[:, i, j] = geom[_s(i + 1, j + 1)]
    L = np.linalg.cholesky(G)  # G = L L^T
    e = rng.standard_normal((n, 3))
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    # v = L^{-T} e  =>  v^T G v = 1
    v = np.linalg.solve(np.transpose(L, (0, 2, 1)), e[:, :, None])[:, :, 0]
    return v


def _loguniform(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n))


def sample_inversion_inputs(n, seed=7, chunk=1 << 20):
    """Return (p_true, p_guess, geom) as float64 SoA arrays (8,n), (8,n), (27,n)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    p = np.empty((8, n))
    geom = np.empty((G_NFIELD, n))
    for c0 in range(0, n, chunk):
        m = min(chunk, n - c0)
        a = rng.uniform(0.0, 0.99, m)
        r = rng.uniform(1.3, 1000.0, m)
        th = rng.uniform(0.0, np.pi, m)
        th = np.clip(th, 1e-6, np.pi - 1e-6)
        g = ks_geometry(a, r, th)
        gamma = _loguniform(rng, 1.0, 50.0, m)
        sigma = _loguniform(rng, 1e-4, 1e3, m)
        uor = _loguniform(rng, 1e-6, 1e2, m)
        rho = np.ones(m)
        vu = _unit_vectors(rng, g)
        ut = np.sqrt(np.maximum(gamma * gamma - 1.0, 0.0))[:, None] * vu  # tilde u^i
        # four-velocity for the field normalisation
        alpha = g[22]
        beta = np.stack([g[24], g[25], g[26]], axis=1)
        uton = gamma / alpha
        ucon = np.empty((m, 4))
        ucon[:, 0] = uton
        ucon[:, 1:] = ut - uton[:, None] * beta
        ucov_sp = np.zeros((m, 3))
        for i in range(3):
            for mu in range(4):
                ucov_sp[:, i] += g[_s(i + 1, mu)] * ucon[:, mu]
        vb = _unit_vectors(rng, g)  # B^i with g_ij B^i B^j = 1
        udotB = np.sum(vb * ucov_sp, axis=1)
        bsq1 = (1.0 + udotB * udotB) / (uton * uton)
        scale = np.sqrt(sigma * rho / bsq1)
        B = vb * scale[:, None]
        p[0, c0:c0 + m] = rho
        p[1, c0:c0 + m] = uor * rho
        p[2:5, c0:c0 + m] = ut.T
        p[5:8, c0:c0 + m] = B.T
        geom[:, c0:c0 + m] = g
    guess = p.copy()
    pert = rng.uniform(-0.2, 0.2, (5, n))
    guess[0:5] *= 1.0 + pert
    return p, guess, geom
