"""Pins the oracle's WENO5BND line reconstruction (CPU, no GPU) against what
the mathematics fixes, so that a plausible mistake in the restatement (a
wrong interpolation or optimal weight, a swapped face, a wrong SMONO
derivative stencil, the wrong reduction threshold, neighbour or ramp) fails
one of these checks:

* the SMONO indicator's derivative stencils against the derivatives of the
  five-point interpolating polynomial at the neighbouring zone centres,
  computed by numpy;
* exactness: monotone quartics (SMONO path) and quadratics with an extremum
  (WENO path, weights limited to equal, so the optimal weights must combine
  the three sub-stencils into the five-point interpolant);
* measured fifth-order convergence (error ratio ~32 per halving) on smooth
  monotone and oscillating data, and no spurious reduction there;
* a single-zone spike, where the AND stencil reduction must switch the zone
  to WENO3 with equal weights (a value derived by hand below);
* the pre/post-shock ramp: lof = max(lof0, s * lof0[c-1], s * lof0[c+1]) with
  s = min((|dP| - 1/4) / (1/4), 1), and the output affine in lof;
* mirror symmetry, and no new extrema next to a jump.
"""
import numpy as np
import pytest

import _oracle

EPS = 500.0 * np.finfo(float).eps


def line(f, N, ng=4):
    h = 1.0 / N
    xc = (np.arange(-ng, N + ng) + 0.5) * h
    return xc, f(xc), h


# ---------------------------------------------------------------------------
# SMONO indicator
# ---------------------------------------------------------------------------
def _face_derivs(y5):
    """derivatives 1..3 of the quartic through (-2..2, y5) at the two
    neighbouring zone centres x = -1, +1"""
    c = np.polyfit(np.arange(-2, 3, dtype=float), y5, 4)
    out = []
    for x in (-1.0, 1.0):
        out.append([np.polyval(np.polyder(c, n), x) for n in (1, 2, 3)])
    return np.array(out)


def _minmod(a, b):
    return 0.0 if a * b <= 0 else (a if abs(a) < abs(b) else b)


def test_smono_indicator_matches_polynomial_derivatives():
    rng = np.random.default_rng(3)
    outcomes = {0.0: 0, 1.0: 0}
    for _ in range(4000):
        kind = rng.integers(3)
        if kind == 0:  # growth with growing, accelerating increments
            y = np.cumsum(np.cumsum(np.cumsum(rng.uniform(0.1, 1.0, 5))))
        elif kind == 1:  # random monotone
            y = np.cumsum(rng.uniform(0.0, 1.0, 5))
        else:  # arbitrary
            y = rng.normal(size=5)
        y = y * rng.choice([-1, 1]) + rng.normal()
        if rng.random() < 0.5:
            y = y[::-1].copy()
        d = _face_derivs(y)
        norm = np.abs(y).sum()
        tol = 1e-9 * norm
        L1 = _minmod(d[0][0], y[1] - y[0])
        R1 = _minmod(d[1][0], y[4] - y[3])
        pairs = [(L1, R1), (d[0][1], d[1][1]), (d[0][2], d[1][2])]
        # a sign decision that polyfit's rounding could flip is ambiguous
        if any(min(abs(a), abs(b)) < tol for a, b in pairs) or min(abs(d[0][0]), abs(y[1] - y[0]),
                                                                   abs(d[1][0]), abs(y[4] - y[3])) < tol:
            continue
        mins = min(abs(_minmod(a, b)) for a, b in pairs)
        cut = norm * EPS + np.finfo(float).tiny
        expect = max(0.0, min(1.0, mins / cut))
        got = _oracle.smono_indicator(y)
        assert got == expect, (y, got, expect)
        outcomes[expect] += 1
    # both outcomes are exercised often (the ramp between them is ~1e-13 wide)
    assert outcomes[0.0] > 500 and outcomes[1.0] > 500, outcomes


def test_smono_indicator_flags_extremum_and_monotone():
    assert _oracle.smono_indicator(np.exp(0.3 * np.arange(-2, 3.0))) == 1.0  # all derivatives positive
    assert _oracle.smono_indicator(np.array([1.0, 2.0, 3.0, 2.0, 1.0])) == 0.0  # extremum
    assert _oracle.smono_indicator(np.full(5, 3.0)) == 0.0  # flat: no monotone derivative


# ---------------------------------------------------------------------------
# exactness
# ---------------------------------------------------------------------------
def test_monotone_quartic_is_reproduced_exactly():
    f = lambda x: 1 + x + 0.3 * x ** 2 + 0.1 * x ** 3 + 0.02 * x ** 4
    for N in (8, 16, 40):
        xc, y, h = line(f, N)
        yl, yr = _oracle.weno5bnd_line(y)
        c = np.arange(3, N + 5)
        assert np.allclose(yl[c], f(xc[c] - h / 2), rtol=2e-15, atol=0)
        assert np.allclose(yr[c], f(xc[c] + h / 2), rtol=2e-15, atol=0)


def test_quadratic_with_extremum_uses_optimal_weights():
    """Near the vertex of a parabola SMONO is off (the first derivative
    changes sign) and the three smoothness indicators are within the 20 %
    band, so the weights are limited to exactly equal: the optimal weights
    must then combine the sub-stencils into the five-point interpolant,
    exact on a quadratic."""
    N = 64
    f = lambda x: 2.0 + (x - 0.3) ** 2
    xc, y, h = line(f, N)
    yl, yr = _oracle.weno5bnd_line(y)
    mono = np.array([_oracle.smono_indicator(y[c - 2:c + 3]) for c in range(3, N + 5)])
    c = np.arange(3, N + 5)
    wenocells = c[mono == 0]
    assert wenocells.size >= 2  # the vertex cells go through the WENO weights
    assert np.allclose(yl[wenocells], f(xc[wenocells] - h / 2), rtol=1e-14, atol=0)
    assert np.allclose(yr[wenocells], f(xc[wenocells] + h / 2), rtol=1e-14, atol=0)
    # and every cell is exact on the quadratic
    assert np.allclose(yl[c], f(xc[c] - h / 2), rtol=1e-14, atol=0)


@pytest.mark.parametrize("f", [np.exp, lambda x: np.sin(2 * np.pi * x + 0.3)])
def test_fifth_order_convergence(f):
    errs = []
    for N in (32, 64, 128):
        xc, y, h = line(f, N)
        dP = np.zeros_like(y)
        yl, yr, lof, lof0 = _oracle.weno5bnd_line_lof(y, dP)
        c = np.arange(3, N + 5)
        if N >= 64:  # resolved smooth data: no stencil reduction anywhere
            assert np.all(lof[c] == 0.0)
        errs.append(max(np.abs(yl[c] - f(xc[c] - h / 2)).max(), np.abs(yr[c] - f(xc[c] + h / 2)).max()))
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 26.0 < r1 < 40.0 and 28.0 < r2 < 36.0, errs


# ---------------------------------------------------------------------------
# stencil reduction and the dP ramp
# ---------------------------------------------------------------------------
def test_single_zone_spike_reduces_to_weno3():
    """y = 1 with y[12] = 2.  At the spike every WENO5 sub-stencil is rough
    and the zone's outer sub-stencils weigh more than the neighbours two
    zones away give the same three zones, so the AND test reduces the zone
    fully to WENO3; its two 2-point sub-stencils (2,1) and (1,2) are equally
    smooth, so the left face value is 1/4 (3/2*2 - 1/2*1) + 3/4 (1/2*1 +
    1/2*2) = 1.75, and by symmetry the right face too.  Unreduced WENO5 with
    these data gives about 1.70."""
    y = np.ones(24)
    y[12] = 2.0
    yl, yr, lof, lof0 = _oracle.weno5bnd_line_lof(y, np.zeros(24))
    assert lof[12] == 1.0
    assert yl[12] == 1.75 and yr[12] == 1.75
    # constant data elsewhere are reproduced whatever the weights
    assert np.all(yl[3:12] == 1.0) and np.all(yr[13:21] == 1.0)


def test_dp_reduction_ramp_and_affine_blend():
    rng = np.random.default_rng(5)
    n = 40
    y = np.cumsum(rng.normal(size=n)) + 5.0 * (np.arange(n) > 20)
    zero = np.zeros(n)
    yl0, yr0, lof_a, base = _oracle.weno5bnd_line_lof(y, zero)
    c = np.arange(4, n - 4)
    assert np.any(base[c] > 0)  # reductions present on these data
    for dPv in (0.2, 0.25, 0.3, 0.4, 0.5, 0.8, 3.0):
        dP = np.full(n, dPv)
        yl, yr, lof, lof0 = _oracle.weno5bnd_line_lof(y, dP)
        assert np.array_equal(lof0, base)  # the ratio fractions do not depend on dP
        s = min((abs(dPv) - 0.25) / 0.25, 1.0)
        expect = np.maximum(np.maximum(base[c], s * base[c - 1]), s * base[c + 1])
        assert np.array_equal(lof[c], expect), dPv
    # output affine in lof: two runs with different lof give the same slope
    d1, d2 = np.full(n, 0.35), np.full(n, 3.0)
    yl1, yr1, lof1, _ = _oracle.weno5bnd_line_lof(y, d1)
    yl2, yr2, lof2, _ = _oracle.weno5bnd_line_lof(y, d2)
    a = np.arange(5, n - 5)
    mono = np.array([_oracle.smono_indicator(y[k - 2:k + 3]) for k in a])
    sel = a[(mono == 0) & (lof2[a] > lof1[a] + 0.1) & (lof1[a] > lof_a[a] + 0.1)]
    assert sel.size > 0
    for yy0, yy1, yy2 in ((yl0, yl1, yl2), (yr0, yr1, yr2)):
        s1 = (yy1[sel] - yy0[sel]) / (lof1[sel] - lof_a[sel])
        s2 = (yy2[sel] - yy0[sel]) / (lof2[sel] - lof_a[sel])
        assert np.allclose(s1, s2, rtol=1e-9, atol=1e-12)


def test_mirror_symmetric_and_no_new_extrema_at_jump():
    rng = np.random.default_rng(11)
    n = 32
    y = np.where(np.arange(n) < 16, 1.0, 3.0) + 0.01 * rng.normal(size=n)
    dP = np.abs(rng.normal(size=n))
    yl, yr = _oracle.weno5bnd_line(y, dP)
    ylm, yrm = _oracle.weno5bnd_line(y[::-1].copy(), dP[::-1].copy())
    c = np.arange(3, n - 3)
    assert np.allclose(yl[c], yrm[::-1][c], rtol=1e-14, atol=1e-14)
    assert np.allclose(yr[c], ylm[::-1][c], rtol=1e-14, atol=1e-14)
    for k in (14, 15, 16, 17):
        lo, hi = y[k - 1:k + 2].min(), y[k - 1:k + 2].max()
        assert lo - 0.05 <= yl[k] <= hi + 0.05 and lo - 0.05 <= yr[k] <= hi + 0.05


def test_dp_line_closed_form_fluid_at_rest():
    """A fluid at rest in flat space: u^dir = 0, so par = veff = 0 and
    upar = |-(u + p + b^2) * 0 ... | = p + b^2/2 (u_t = -1); dP is the
    relative jump of total pressure between the two neighbours."""
    # flat metric as a 27-field record: gcov and gcon packed symmetric
    # (00,01,02,03,11,12,13,22,23,33), 1 + g_tt, gdet, alpha, beta^2/alpha^2, beta^i
    rec = np.zeros(27)
    for off in (0, 10):
        rec[off + 0] = -1.0
        rec[off + 4] = rec[off + 7] = rec[off + 9] = 1.0
    rec[21] = rec[22] = 1.0
    cells = np.zeros((5, 8))
    cells[:, 0] = 1.0
    cells[:, 1] = [1.0, 2.0, 3.0, 5.0, 8.0]
    dP, st = _oracle.dp_line(cells, rec, 1)
    p = (4.0 / 3.0 - 1.0) * cells[:, 1]
    assert np.allclose(st[:, 3], p, rtol=1e-15)
    for c in (1, 2, 3):
        assert dP[c] == pytest.approx(abs(p[c + 1] - p[c - 1]) / p[c], rel=1e-14)
