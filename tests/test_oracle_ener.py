"""Pins of the oracle's energy-file diagnostics (o_grid.c: ener_pdot,
ener_dt4, ener_source, diag_fixup_acc, o_ener_totals), CPU only.

* Conservation identity (what the diagnostics exist to close, dump_ener.c:
  316-329): for a variable without a geometric source, the change of its
  total equals the boundary inflow minus outflow plus the fix-up additions,
  U_final - U_init = sum_dir (pcum[DN] - pcum[UP]) + fladd.  The conservative
  update telescopes, so this holds to round-off when the inversion succeeds;
  it pins pdot/pcum (faces, surfaces, substep weights, signs), the conserved
  totals and the floor part of fladd.  Pole death (x2 boundary) rewrites the
  polar rows of active zones without accounting, as the reference does, so
  the torus density closes to ~1e-12 instead of round-off.
* div B: closed forms for gdet B linear in one index.
* Conserved totals: closed form for a fluid at rest in the normal frame.
* Source sampling: the weights of the dt4diag schedule (advance.c:251-262)
  times the source total of the sampled state, obtained independently
  through the update with zero fluxes.
"""
import ctypes

import numpy as np
import pytest

import _oracle
import harm_harm_amd as H


def _perturbed(og, amp, seed=3):
    p = og.init_torus()
    rng = np.random.default_rng(seed)
    a = p[og.ng3:og.ng3 + og.cfg.n3, og.ng2:og.ng2 + og.cfg.n2, og.ng1:og.ng1 + og.cfg.n1]
    a[..., :5] *= 1.0 + amp * rng.uniform(-1, 1, a[..., :5].shape)
    og.bound(p)
    return p


def _identity(og, p, nsteps):
    """run nsteps with the accumulators on; residual and scale per variable"""
    st = og.new_state(p)
    F = [og.zeros() for _ in range(3)]
    dt = 0.5 * min(og.fluxcalc(p, *F))
    og.track_ener(0)
    e0 = og.ener(p)
    nfail = 0
    for _ in range(nsteps):
        dt = og.step(st, dt)
        nfail += int((st["pflag"] > 0).sum())
    e1 = og.ener(st["p"])
    pc = e1["pcum"]
    inflow = pc[1] - pc[0] + pc[3] - pc[2] + pc[5] - pc[4]  # X1DN-X1UP + X2DN-X2UP + X3DN-X3UP
    resid = e1["U"] - e0["U"] - inflow - e1["fladd"]
    cfg = og.cfg
    gd = og.geom[0][og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1, 21][None]
    act = p[og.ng3:og.ng3 + cfg.n3, og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1]
    dV = og.dx[1] * og.dx[2] * og.dx[3]
    scale = np.array([np.abs(act[..., pl] * gd).sum() * dV for pl in range(8)])
    scale = scale + np.abs(pc).sum(0) + np.abs(e1["fladd"])
    return resid, scale, e1, nfail


def test_conservation_identity_torus_2d():
    cfg = H.default_config(32, 32, 1, recon="para", flux="laxf", timeorder=2)
    og = _oracle.OGrid(cfg)
    resid, scale, e, nfail = _identity(og, _perturbed(og, 0.02), 10)
    rel = np.abs(resid) / scale
    print("failures", nfail, "fladd", e["fladd"][:2], "relative residual", rel[[0, 5, 6, 7]])
    assert nfail == 0
    # the floors added mass: a dropped or mis-signed fladd would leave this much
    assert abs(e["fladd"][0]) / scale[0] > 1e-8
    assert rel[0] < 1e-10          # density: pole death's unaccounted rows
    assert rel[5:].max() < 1e-13   # field: round-off


def test_conservation_identity_bondi():
    """Bondi flow (no floors, no pole death): the density closes to round-off
    through the radial boundaries and the polar faces."""
    cfg = H.bondi_config(32, 32, recon="para", flux="laxf", timeorder=2)
    og = _oracle.OGrid(cfg)
    resid, scale, e, nfail = _identity(og, og.init_bondi(), 10)
    print("relative residual", resid[0] / scale[0], "inflow", e["pcum"][1][0], "outflow", e["pcum"][0][0])
    assert nfail == 0
    assert abs(e["pcum"][1][0]) > 1e-3 * scale[0]  # a real inflow through the inner boundary
    assert abs(resid[0]) / scale[0] < 1e-14


def test_conservation_identity_field_3d():
    """3-D torus, WENO5/HLL/RK3: the field closes to round-off through all six
    boundaries (periodic phi faces included)."""
    cfg = H.default_config(16, 12, 8, recon="weno5", flux="hll", timeorder=3)
    og = _oracle.OGrid(cfg)
    resid, scale, e, nfail = _identity(og, _perturbed(og, 0.02), 3)
    rel = np.abs(resid) / scale
    print("failures", nfail, "relative residual B", rel[5:], "X3 pdot B3", e["pdot"][4][7], e["pdot"][5][7])
    assert rel[5:].max() < 1e-13
    assert np.any(e["pcum"][4:6, 5:] != 0.0)  # the phi faces carry field flux


def _set_linear_field(og, c0, c1, comp, axis):
    """gdet B^comp = c0 + c1 * (zone index along axis) on every storage zone"""
    p = og.zeros()
    p[..., 0] = 1.0
    p[..., 1] = 1.0
    gd = og.geom[0][..., 21]
    ng = [og.ng1, og.ng2, og.ng3][axis - 1]
    idx = np.arange(p.shape[3 - axis]) - ng
    shape = [1, 1, 1]
    shape[3 - axis] = -1
    F = c0 + c1 * idx.reshape(shape)
    p[..., 5 + comp - 1] += F / gd[None, :, :]
    return p


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_divb_closed_form(axis):
    """gdet B^a = c0 + c1 m (m the zone index along a, other components 0):
    the corner divergence is c1 / (2 c0 + c1 (2m - 1)) at every zone, so the
    maximum is at m = 0 and the sum is n_other * sum_m of it."""
    cfg = H.default_config(12, 10, 6, recon="para", flux="laxf", timeorder=2)
    og = _oracle.OGrid(cfg)
    c0, c1 = 3.0, 0.25
    p = _set_linear_field(og, c0, c1, axis, axis)
    tot = np.zeros(10)
    _oracle.lib().o_ener_totals(ctypes.byref(og.g), p.ctypes.data, tot.ctypes.data)
    n = [cfg.n1, cfg.n2, cfg.n3][axis - 1]
    m = np.arange(n)
    per = c1 / (2 * c0 + c1 * (2 * m - 1))
    nother = cfg.n1 * cfg.n2 * cfg.n3 // n
    print("axis", axis, "max", tot[8], per.max(), "sum", tot[9], nother * per.sum())
    assert np.isclose(tot[8], per.max(), rtol=1e-13)
    assert np.isclose(tot[9], nother * per.sum(), rtol=1e-12)


def test_divb_closed_form_mixed():
    """gdet B^1 = a0 + a1 i and gdet B^2 = b0 + b1 j together: divb =
    (a1/dx1 + b1/dx2) / ((2 a0 + a1 (2i-1))/dx1 + (2 b0 + b1 (2j-1))/dx2),
    which weighs the directions by their dx."""
    cfg = H.default_config(8, 6, 1, recon="para", flux="laxf", timeorder=2)
    og = _oracle.OGrid(cfg)
    p1 = _set_linear_field(og, 2.0, 0.5, 1, 1)
    p2 = _set_linear_field(og, 5.0, -0.75, 2, 2)
    p = p1.copy()
    p[..., 6] = p2[..., 6]
    tot = np.zeros(10)
    _oracle.lib().o_ener_totals(ctypes.byref(og.g), p.ctypes.data, tot.ctypes.data)
    dx1, dx2 = og.dx[1], og.dx[2]
    i = np.arange(cfg.n1)[None, :]
    j = np.arange(cfg.n2)[:, None]
    num = np.abs(0.5 / dx1 + (-0.75) / dx2)
    den = np.abs((2 * 2.0 + 0.5 * (2 * i - 1)) / dx1 + (2 * 5.0 - 0.75 * (2 * j - 1)) / dx2)
    ref = num / den
    assert np.isclose(tot[8], ref.max(), rtol=1e-13)
    assert np.isclose(tot[9], ref.sum(), rtol=1e-12)


def test_conserved_totals_fluid_at_rest():
    """u-tilde = 0, B = 0: u^t = 1/alpha, u_t = -alpha, so U[RHO] = gdet rho / alpha
    and U[UU] = T^t_t + rho u^t = gdet (rho/alpha - (rho + u)); momenta and
    field vanish."""
    cfg = H.default_config(10, 8, 4, recon="para", flux="laxf", timeorder=2)
    og = _oracle.OGrid(cfg)
    rng = np.random.default_rng(5)
    p = og.zeros()
    p[..., 0] = rng.uniform(0.5, 2.0, p.shape[:3])
    p[..., 1] = rng.uniform(0.1, 1.0, p.shape[:3])
    tot = np.zeros(10)
    _oracle.lib().o_ener_totals(ctypes.byref(og.g), p.ctypes.data, tot.ctypes.data)
    a = p[og.ng3:og.ng3 + cfg.n3, og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1]
    G = og.geom[0][og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1][None]
    gd, al = G[..., 21], G[..., 22]
    dV = og.dx[1] * og.dx[2] * og.dx[3]
    rho, u = a[..., 0], a[..., 1]
    U0 = (gd * rho / al).sum() * dV
    U1 = (gd * (rho / al - (rho + u))).sum() * dV
    assert np.isclose(tot[0], U0, rtol=1e-13)
    assert np.isclose(tot[1], U1, rtol=1e-12)
    assert np.all(np.abs(tot[5:8]) == 0.0)


def _source_total(og, pb, dt):
    """sum dV dUgeom(pb) through the update with zero fluxes (uf = dt dU)"""
    F = [og.zeros() for _ in range(3)]
    ui, uf, ucum = og.zeros(), og.zeros(), og.zeros()
    og.update(pb, pb, *F, ui, uf, ucum, [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 0.0], dt, 0)
    cfg = og.cfg
    a = uf[og.ng3:og.ng3 + cfg.n3, og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1]
    return a.reshape(-1, 8).sum(0) * (og.dx[1] * og.dx[2] * og.dx[3]) / dt


@pytest.mark.parametrize("timeorder", [2, 3])
def test_source_sampling_schedule(timeorder):
    """Sampled on the last substep of steps with nstep % 4 == 0, from that
    substep's starting state, weighted by the dt of every substep since the
    previous sample (the current one excluded)."""
    cfg = H.default_config(24, 16, 1, recon="para", flux="laxf", timeorder=timeorder)
    og = _oracle.OGrid(cfg)
    p = _perturbed(og, 0.02)
    st = og.new_state(p)
    F = [og.zeros() for _ in range(3)]
    dt = 0.5 * min(og.fluxcalc(p, *F))
    og.track_ener(0)
    nsub = timeorder
    last_state = "pk0" if timeorder == 2 else "pk1"
    dts, expect = [], np.zeros(8)
    gathered = 0.0
    for n in range(6):
        dts.append(dt)
        dt = og.step(st, dt)
        if n % 4 == 0:
            w = gathered + (nsub - 1) * dts[-1]
            expect = expect + w * _source_total(og, st[last_state], 1.0)
            gathered = 0.0
        else:
            gathered += nsub * dts[-1]
        got = og.ener(st["p"])["sourceadd"]
        assert np.allclose(got[1:5], expect[1:5], rtol=1e-12, atol=1e-14 * np.abs(expect).max()), (n, got, expect)
    assert np.all(got[[0, 5, 6, 7]] == 0.0)
