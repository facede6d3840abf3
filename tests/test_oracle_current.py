"""Pins the oracle's four-current (o_current_precalc, o_current_calc; the
reference's current_precalc phys.c:622 and current_calc_1 :846 with
CURRENTCALC1) on the CPU.

* The Faraday samples against F^{mu nu} = -(1/gdet) [mu nu kappa lambda]
  u_kappa b_lambda built here from the Levi-Civita symbol and the oracle's own
  u_mu, b_mu (the dump row): slot 0 = (F^{rt}, F^{ht}, F^{pt}), slot 1 =
  (F^{tr}, F^{hr}, F^{pr}), slot 2 = (F^{th}, F^{rh}, F^{ph}), slot 3 =
  (F^{tr}, F^{rp}, F^{hp}) -- the reference's CURTYPEZ sample 0 is
  F^{tr} where its comment and the x3 derivative of J^t want F^{tp}
  (phys.c:679); the oracle keeps the reference's value -- and the shift of the
  two previous F^{it} samples into slots 4 and 5.
* The spatial differences: with gdet times every sample linear in the zone
  index along its derivative's direction, J^mu is the exact slope over gdet.
* The time derivative: samples of a quadratic in t at t0, t0 + dtl,
  t0 + dtl + dtr give its exact derivative at the newest time, and none is
  taken before the third call.
"""
import itertools

import numpy as np
import pytest

import _oracle
import harm_harm_amd as H


def _levi(a, b, c, d):
    perm = [a, b, c, d]
    if len(set(perm)) < 4:
        return 0
    sign = 1
    for x in range(4):
        for y in range(x + 1, 4):
            if perm[x] > perm[y]:
                sign = -sign
    return sign


def _grid(n=(12, 10, 6)):
    cfg = H.default_config(*n, recon="para", flux="laxf", timeorder=2)
    return cfg, _oracle.OGrid(cfg, nthreads=2)


def test_faraday_samples_match_levi_civita_form():
    cfg, og = _grid()
    p = og.init_torus()
    rng = np.random.default_rng(4)
    a = p[og.ng3:og.ng3 + cfg.n3, og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1]
    a[..., 2:8] += 0.05 * rng.standard_normal(a[..., 2:8].shape)  # velocities and fields in every component
    cf = og.current_slots()
    cf[..., 0, :] = rng.standard_normal(cf[..., 0, :].shape)
    cf[..., 5, :] = rng.standard_normal(cf[..., 5, :].shape)
    old0, old5 = cf[..., 0, :].copy(), cf[..., 5, :].copy()
    og.current_precalc(p, cf)
    assert np.array_equal(cf[..., 4, :], old5) and np.array_equal(cf[..., 5, :], old0)
    rows = og.dump_rows(p, og.zeros())
    ucov, bcov, gdet = rows[:, 24:28], rows[:, 32:36], rows[:, 42]
    F = np.zeros((len(rows), 4, 4))
    for m, nn, k, l in itertools.product(range(4), repeat=4):
        e = _levi(m, nn, k, l)
        if e:
            F[:, m, nn] += e * ucov[:, k] * bcov[:, l]
    F *= -1.0 / gdet[:, None, None]
    act = cf[og.ng3:og.ng3 + cfg.n3, og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1].reshape(-1, 6, 3)
    want = {0: [(1, 0), (2, 0), (3, 0)], 1: [(0, 1), (2, 1), (3, 1)], 2: [(0, 2), (1, 2), (3, 2)],
            3: [(0, 1), (1, 3), (2, 3)]}
    scale = np.abs(F).max()
    for slot, comps in want.items():
        for c, (m, nn) in enumerate(comps):
            err = np.abs(act[:, slot, c] - F[:, m, nn]).max() / scale
            assert err < 1e-13, (slot, c, err)
    # the reference's slot 3 sample 0 is F^{tr}, not F^{tp}: they differ here
    assert np.abs(act[:, 3, 0] - F[:, 0, 3]).max() > 1e-6 * scale


def test_spatial_differences_are_exact_on_linear_data():
    cfg, og = _grid()
    cf = og.current_slots()
    ng1, ng2, ng3 = og.ng1, og.ng2, og.ng3
    gd = og.geom[0][..., 21]  # (ns2, ns1)
    i = (np.arange(og.ns1) - ng1)[None, None, :]
    j = (np.arange(og.ns2) - ng2)[None, :, None]
    k = (np.arange(og.ns3) - ng3)[:, None, None]
    rng = np.random.default_rng(7)
    a = rng.uniform(0.5, 2.0, 9)
    # gdet * sample = a * (index along the derivative's direction)
    lin = {(1, 0): a[0] * i, (1, 1): a[1] * i, (1, 2): a[2] * i, (2, 0): a[3] * j, (2, 1): a[4] * j,
           (2, 2): a[5] * j, (3, 0): a[6] * k, (3, 1): a[7] * k, (3, 2): a[8] * k}
    for (slot, c), v in lin.items():
        cf[..., slot, c] = np.broadcast_to(v, cf.shape[:3]) / gd[None, :, :]
    J = og.current_calc(cf, 0, 1.0, 1.0)
    gt = gd[ng2:ng2 + cfg.n2, ng1:ng1 + cfg.n1][None, :, :]
    dx = og.dx
    want = [a[0] / dx[1] + a[3] / dx[2] + a[6] / dx[3], a[4] / dx[2] + a[7] / dx[3],
            a[1] / dx[1] + a[8] / dx[3], a[2] / dx[1] + a[5] / dx[2]]
    for mu in range(4):
        exp = want[mu] / gt
        err = np.abs(J[..., mu] - exp).max() / np.abs(exp).max()
        assert err < 1e-12, (mu, err)


def test_time_derivative_of_a_quadratic():
    cfg, og = _grid((8, 6, 1))
    cf = og.current_slots()
    rng = np.random.default_rng(9)
    c0, c1, c2 = rng.standard_normal((3, 3))
    t0, dtl, dtr = 0.3, 0.07, 0.05
    f = lambda t: c0 + c1 * t + c2 * t * t
    cf[..., 4, :] = f(t0)
    cf[..., 5, :] = f(t0 + dtl)
    cf[..., 0, :] = f(t0 + dtl + dtr)
    J = og.current_calc(cf, 1, dtl, dtr)
    want = c1 + 2 * c2 * (t0 + dtl + dtr)
    for c in range(3):
        assert np.allclose(J[..., 1 + c], want[c], rtol=1e-10, atol=1e-12), c
    assert np.all(J[..., 0] == 0.0)
    assert np.all(og.current_calc(cf, 0, dtl, dtr) == 0.0)  # no time term before the third call
