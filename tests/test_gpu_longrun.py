"""Long-run agreement with the oracle: 1000 steps of the 2-D torus (PARA,
LAX-F, RK2, 64x64), comparing the accumulated mass and energy accretion
through the inner and outer radial boundaries (diag_flux_pureflux) and the
final primitives.  Differences start at the last-ulp level (device exp/log in
the radial boundary extrapolation) and grow with the flow.  Measured on an
MI355X: accretion sums 0.8-3.5e-12 relative, dt 2e-11, primitives 1.0e-8
(largest in the low-density atmosphere); the bounds below are those values
times ~100."""
import numpy as np
import pytest

import _oracle
import harm_harm_amd as H
from harm_harm_amd.grid import Grid

pytestmark = pytest.mark.gpu


def test_accretion_rates_match_over_1000_steps():
    cfg = H.default_config(64, 64, 1, recon="para", flux="laxf", timeorder=2)
    og = _oracle.OGrid(cfg, nthreads=16)
    g = Grid(cfg)
    p = og.init_torus()
    dt = 0.5 / sum(1.0 / x for x in og.fluxcalc(p, og.zeros(), og.zeros(), og.zeros()))
    g.set("p", p)
    g.set_time(0.0, dt)
    st = og.new_state(p.copy())
    fa = np.zeros(4)
    for _ in range(1000):
        dt = og.step(st, dt, fa)
    g.step(1000)
    fg = np.array(g.boundary_fluxes())
    t_g, dt_g = g.time()
    rel_f = np.abs(fg - fa) / np.abs(fa)
    a = g.get("p")[0, og.ng2:og.ng2 + 64, og.ng1:og.ng1 + 64]
    b = st["p"][0, og.ng2:og.ng2 + 64, og.ng1:og.ng1 + 64]
    rel_p = max(float(np.abs(a[..., v] - b[..., v]).max() / np.abs(b[..., v]).max()) for v in range(8))
    print("accretion (mdot_in, edot_in, mdot_out, edot_out) gpu", fg, "oracle", fa, "rel", rel_f)
    print("primitives relmax", rel_p, "dt", dt_g, dt)
    assert np.all(fa[:2] != 0.0)
    assert np.all(rel_f <= 1e-9)
    assert rel_p <= 1e-6
    assert abs(dt_g - dt) <= 1e-9 * dt


def test_config_2d_128x128_para_laxf_rk2_200_steps():
    """The reference's CPU configuration: the 2-D 128x128 torus with PARA,
    LAX-F, Toth CT and RK2 at Courant 0.8, 200 steps, GPU against the oracle:
    final primitives, dt, the accumulated accretion sums and Toth's corner
    divergence of B (kept at round-off by both).  Measured on an MI355X:
    primitives 5.6e-12, accretion sums 1.7e-15, dt 4e-14, div B 8e-15
    relative; the bounds are about 100 times those."""
    cfg = H.default_config(128, 128, 1, recon="para", flux="laxf", timeorder=2)
    assert abs(cfg.cour - 0.8) < 1e-15
    og = _oracle.OGrid(cfg, nthreads=16)
    g = Grid(cfg)
    p = og.init_torus(seed=1)
    dt = 1.0 / sum(1.0 / x for x in og.fluxcalc(p, og.zeros(), og.zeros(), og.zeros()))
    g.set("p", p)
    g.set_time(0.0, dt)
    st = og.new_state(p.copy())
    fa = np.zeros(4)
    for _ in range(200):
        dt = og.step(st, dt, fa)
    g.step(200)
    fg = np.array(g.boundary_fluxes())
    _, dt_g = g.time()
    rel_f = np.abs(fg - fa) / np.abs(fa)
    J, I = slice(og.ng2, og.ng2 + 128), slice(og.ng1, og.ng1 + 128)
    a, b = g.get("p")[0, J, I], st["p"][0, J, I]
    rel_p = max(float(np.abs(a[..., v] - b[..., v]).max() / np.abs(b[..., v]).max()) for v in range(8))
    # corner divergence of gdet B (Toth), relative to |gdet B| / dx
    gd = og.geom[0][..., 21]
    Bg = [g.get("p")[0, ..., 5 + m] * gd for m in range(2)]
    c = lambda A, di, dj: A[og.ng2 + dj:og.ng2 + dj + 127, og.ng1 + di:og.ng1 + di + 127]
    divb = ((c(Bg[0], 1, 1) + c(Bg[0], 1, 0) - c(Bg[0], 0, 1) - c(Bg[0], 0, 0)) / (2 * og.dx[1])
            + (c(Bg[1], 1, 1) + c(Bg[1], 0, 1) - c(Bg[1], 1, 0) - c(Bg[1], 0, 0)) / (2 * og.dx[2]))
    scale = max(np.abs(c(Bg[0], 0, 0)).max(), np.abs(c(Bg[1], 0, 0)).max()) / min(og.dx[1], og.dx[2])
    print("accretion gpu", fg, "oracle", fa, "rel", rel_f)
    print("primitives relmax", rel_p, "dt", dt_g, dt, "divB/scale", np.abs(divb).max() / scale)
    assert np.all(fa[:2] != 0.0)
    assert np.all(rel_f <= 1e-12)
    assert rel_p <= 1e-9
    assert abs(dt_g - dt) <= 1e-11 * dt
    assert np.abs(divb).max() / scale < 1e-12
