"""Pins the host's connection table against the closed form (CPU, no GPU).

The library builds Gamma^a_{bc} of the modified Kerr-Schild metric from
numerical derivatives of g_{mu nu} (Ridders' extrapolation of central
differences, the method of the reference's conn_func_numerical1,
metric.tools.c:1771).  Here sympy derives the same connection exactly: the
Kerr-Schild metric (M = 1, spin a) pulled back to MKS, r = R0 + exp(X1),
theta = pi X2 + (1 - h)/2 sin(2 pi X2), phi = 2 pi X3, differentiated
symbolically and evaluated with 30-digit arithmetic at zone centres all over
the grid (near the poles too).  All 40 independent components must agree to
1e-11 of the largest component at that zone (measured: 3.5e-13), so a dropped term, a wrong sign
or index, or a transposed pair of lower indices fails; the trace identity
Gamma^a_{a i} = d_i ln gdet is checked on the closed form as well.
"""
import functools

import numpy as np
import pytest

import _oracle
import harm_harm_amd as H

sp = pytest.importorskip("sympy")
mpmath = pytest.importorskip("mpmath")

SYM = {(0, 0): 0, (0, 1): 1, (0, 2): 2, (0, 3): 3, (1, 1): 4, (1, 2): 5, (1, 3): 6, (2, 2): 7, (2, 3): 8,
       (3, 3): 9}


@functools.lru_cache(maxsize=None)
def metric_and_derivatives(a_val, h_val, R0_val):
    """mpmath functions of (X1, X2): g_{mu nu} and its X1, X2 derivatives,
    from the symbolic MKS Kerr-Schild metric"""
    X1, X2 = sp.symbols("X1 X2", real=True)
    a, h, R0 = sp.Float(a_val, 30), sp.Float(h_val, 30), sp.Float(R0_val, 30)
    r = R0 + sp.exp(X1)
    th = sp.pi * X2 + (1 - h) / 2 * sp.sin(2 * sp.pi * X2)
    rho2 = r ** 2 + a ** 2 * sp.cos(th) ** 2
    m = 2 * r / rho2
    s2 = sp.sin(th) ** 2
    gks = sp.zeros(4, 4)
    gks[0, 0] = -1 + m
    gks[0, 1] = gks[1, 0] = m
    gks[0, 3] = gks[3, 0] = -m * a * s2
    gks[1, 1] = 1 + m
    gks[1, 3] = gks[3, 1] = -a * s2 * (1 + m)
    gks[2, 2] = rho2
    gks[3, 3] = s2 * (rho2 + a ** 2 * s2 * (1 + m))
    J = sp.diag(1, sp.diff(r, X1), sp.diff(th, X2), 2 * sp.pi)
    g = J.T * gks * J
    flat = [g[i, j] for i in range(4) for j in range(4)]
    fg = sp.lambdify((X1, X2), flat, modules="mpmath")
    fd1 = sp.lambdify((X1, X2), [sp.diff(e, X1) for e in flat], modules="mpmath")
    fd2 = sp.lambdify((X1, X2), [sp.diff(e, X2) for e in flat], modules="mpmath")
    return fg, fd1, fd2


def closed_form_at(fns, X1, X2):
    """Gamma^a_{bc} (dict over b <= c) and d_i ln sqrt(-g) at (X1, X2), 30 digits"""
    fg, fd1, fd2 = fns
    X1, X2 = mpmath.mpf(X1), mpmath.mpf(X2)
    g = mpmath.matrix(4, 4)
    for k, v in enumerate(fg(X1, X2)):
        g[k // 4, k % 4] = v
    gi = g ** -1
    dg = [[[mpmath.mpf(0)] * 4 for _ in range(4)] for _ in range(4)]
    for d, fd in ((1, fd1), (2, fd2)):
        for k, v in enumerate(fd(X1, X2)):
            dg[d][k // 4][k % 4] = v
    conn = {}
    for A in range(4):
        for B in range(4):
            for C in range(B, 4):
                conn[(A, B, C)] = sum(gi[A, D] * (dg[C][D][B] + dg[B][D][C] - dg[D][B][C]) for D in range(4)) / 2
    # d_i ln sqrt(-det g) = 1/2 g^{mn} d_i g_mn
    dln = [sum(gi[m_, n] * dg[i][m_][n] for m_ in range(4) for n in range(4)) / 2 for i in (1, 2)]
    return conn, dln


@pytest.mark.parametrize("n", [(48, 32, 1), (64, 64, 1)])
def test_connection_table_matches_closed_form(n):
    cfg = H.default_config(*n, recon="para", flux="laxf", timeorder=2)
    og = _oracle.OGrid(cfg, nthreads=2)
    fns = metric_and_derivatives(float(cfg.a), float(cfg.hslope), float(cfg.R0))
    mpmath.mp.dps = 30
    sx = [og.g.startx[a] for a in range(4)]
    dx = og.dx
    worst = 0.0
    rng = np.random.default_rng(0)
    for _ in range(40):
        i = int(rng.integers(-og.ng1, cfg.n1 + og.ng1))
        j = int(rng.integers(0, cfg.n2))  # active theta rows, down to the zones next to the poles
        X1 = sx[1] + (i + 0.5) * dx[1]
        X2 = sx[2] + (j + 0.5) * dx[2]
        conn, dln = closed_form_at(fns, X1, X2)
        table = og.conn[j + og.ng2, i + og.ng1]
        scale = max(abs(float(v)) for v in conn.values())
        for (A, B, C), ex in conn.items():
            worst = max(worst, abs(table[A * 10 + SYM[(B, C)]] - float(ex)) / scale)
        # the trace identity Gamma^a_{a i} = d_i ln sqrt(-g) holds for the closed form
        for i_, t in zip((1, 2), dln):
            trace = sum(conn[(A_, min(A_, i_), max(A_, i_))] for A_ in range(4))
            assert abs(trace - t) <= mpmath.mpf(10) ** -24 * max(1, abs(t))
    print("worst |table - closed form| / max|Gamma| at the zone:", worst)
    assert worst < 1e-11
