"""GPU parity of the Bondi problem (TESTNUMBER 153 boundary set) with the
oracle, and the flow staying steady.  The boundary pass starts from scrambled
ghosts; whole steps compare primitives and dt; a mixed batch of tori and
Bondi flows advances bit for bit as its grids do alone.  Tolerance: 1e-12
relative per variable against its largest magnitude (the Bondi boundaries use
device pow/exp/sqrt, which may round the last bit differently from glibc)."""
import numpy as np
import pytest

import _oracle
import harm_harm_amd as H
from harm_harm_amd.batch import Batch
from harm_harm_amd.grid import Grid

pytestmark = pytest.mark.gpu


def relmax(a, b):
    a = a.reshape(-1, a.shape[-1])
    b = b.reshape(-1, b.shape[-1])
    out = 0.0
    for v in range(a.shape[-1]):
        s = max(np.abs(a[:, v]).max(), np.abs(b[:, v]).max())
        if s > 0:
            out = max(out, np.abs(a[:, v] - b[:, v]).max() / s)
    return out


def relmax_flow(a, b):
    """relmax with one scale for the three velocity components (the largest
    of them): in the radial Bondi flow the polar and azimuthal velocities are
    rounding noise, so their own maximum is no scale."""
    a = a.reshape(-1, a.shape[-1])
    b = b.reshape(-1, b.shape[-1])
    out = 0.0
    vs = max(np.abs(a[:, 2:5]).max(), np.abs(b[:, 2:5]).max())
    for v in range(a.shape[-1]):
        s = vs if 2 <= v <= 4 else max(np.abs(a[:, v]).max(), np.abs(b[:, v]).max())
        if s > 0:
            out = max(out, np.abs(a[:, v] - b[:, v]).max() / s)
    return out


def act(og, a):
    return a[og.ng3:og.ng3 + og.cfg.n3, og.ng2:og.ng2 + og.cfg.n2, og.ng1:og.ng1 + og.cfg.n1]


@pytest.mark.parametrize("n", [(64, 16), (96, 24)])
def test_bondi_bound_matches_oracle(n):
    cfg = H.bondi_config(*n)
    og, g = _oracle.OGrid(cfg), Grid(cfg)
    p = og.init_bondi()
    q = np.random.default_rng(2).uniform(-3.0, 3.0, p.shape)
    act(og, q)[...] = act(og, p)
    g.set("p", q)
    g.bound("p")
    og.bound(q)
    err = relmax_flow(g.get("p"), q)
    print(n, "bound relmax", err)
    assert err <= 1e-12


@pytest.mark.parametrize("scheme", [("para", "laxf", 2), ("weno5", "hll", 3)])
def test_bondi_steps_match_oracle(scheme):
    recon, flux, to = scheme
    cfg = H.bondi_config(64, 16, recon=recon, flux=flux, timeorder=to)
    og, g = _oracle.OGrid(cfg), Grid(cfg)
    p = og.init_bondi()
    dt = 1.0 / sum(1.0 / x for x in og.fluxcalc(p, og.zeros(), og.zeros(), og.zeros()))
    g.set("p", p)
    g.set_time(0.0, dt)
    st = og.new_state(p.copy())
    for n in range(4):  # per step, as the torus comparison
        dt = og.step(st, dt)
        g.step(1)
        a, b = act(og, g.get("p")), act(og, st["p"])
        err = relmax_flow(a, b)
        _, dt_g = g.time()
        print(scheme, "step", n, "p relmax", err, "max |u~2|", float(np.abs(b[..., 3]).max()), "dt", dt_g, dt)
        assert err <= 1e-12
        assert abs(dt_g - dt) <= 1e-12 * dt


def test_bondi_flow_stays_steady():
    """500 steps of the 64x16 Bondi flow (PARA, LAX-F, RK2): the density and
    the radial velocity stay close to the exact steady flow (the departure is
    the scheme's truncation error), and nothing fails."""
    cfg = H.bondi_config(64, 16)
    g = Grid(cfg)
    g.init_bondi()
    p0 = g.active(g.get("p")).copy()
    ndt = g.fluxcalc("p")
    g.set_time(0.0, 1.0 / sum(1.0 / x for x in ndt))
    g.step(500)
    p1 = g.active(g.get("p"))
    t, _ = g.time()
    _, (_, nfail) = g.newton_stats()
    drho = np.abs(p1[..., 0] / p0[..., 0] - 1).max()
    du1 = np.abs(p1[..., 2] - p0[..., 2]).max() / np.abs(p0[..., 2]).max()
    print("t", t, "max |rho/rho_exact - 1|", drho, "max |du~1|/max|u~1|", du1, "failures", nfail)
    assert nfail == 0
    assert drho < 0.02 and du1 < 0.02


def test_mixed_batch_equals_single_grids():
    """Tori and Bondi flows of uneven sizes in one batch: the per-grid
    boundary selection and everything else bit for bit as when stepped alone."""
    cs = []
    for n1, n2, kind in ((64, 16, "bondi"), (32, 32, "torus"), (96, 24, "bondi"), (40, 32, "torus")):
        c = H.bondi_config(n1, n2) if kind == "bondi" else H.default_config(n1, n2, 1)
        cs.append(c)
    B = Batch(cs)
    singles = [Grid(c) for c in cs]
    for g, c in zip(B.grids, cs):
        if c.problem == 1:
            g.init_bondi()
        else:
            g.init_torus(seed=3)
        ndt = g.fluxcalc("p")
        g.set_time(0.0, 0.5 / sum(1.0 / x for x in ndt))
    for g, s in zip(B.grids, singles):
        s.set("p", g.get("p"))
        s.set_time(*g.time())
    for n in range(4):
        B.step(1, graph=(n % 2 == 1))
        for s in singles:
            s.step(1)
        for i, (g, s) in enumerate(zip(B.grids, singles)):
            assert np.array_equal(g.get("p"), s.get("p")), (n, i)
            assert g.time() == s.time()
    for s in singles:
        s.close()
    B.close()
