"""Pins of the oracle's main-dump row (o_dump_row, dump.c:439-600), CPU only:
physical identities of the columns rather than their formulas.

* u^mu u_mu = -1 and b^mu u_mu = 0 (the state's four-vectors);
* ideal MHD: F^{mu nu} u_nu = 0 and F^{mu nu} b_nu = 0 (F is built from u
  and b; a swapped index pair or a sign error breaks both), and lowering F
  with the metric then raising it again returns F (lowerf's index order);
* the ideal-gas columns: p = (gam-1) u, cs2 = gam p / (rho + u + p), the
  entropy density rho ln(p^n / rho^(n+1)) with n = 1/(gam-1), and its
  special cases (rho < SMALL gives 0, p < SMALL gives -BIG);
* vmin <= vmax in every direction, U copied from ucum, gdet from the table.
"""
import numpy as np

import _oracle
import harm_harm_amd as H

C = {"P": 0, "PRESS": 8, "CS2": 9, "SDEN": 10, "U": 11, "DIVB": 19, "UCON": 20, "UCOV": 24, "BCON": 28, "BCOV": 32,
     "V": 36, "GDET": 42, "FCON": 43, "FCOV": 49}


def _rows():
    cfg = H.default_config(12, 10, 4, recon="para", flux="laxf", timeorder=2)
    og = _oracle.OGrid(cfg)
    p = og.init_torus()
    rng = np.random.default_rng(7)
    a = p[og.ng3:og.ng3 + cfg.n3, og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1]
    a[..., :5] *= 1.0 + 0.1 * rng.uniform(-1, 1, a[..., :5].shape)
    a[..., 2:5] += 0.3 * rng.uniform(-1, 1, a[..., 2:5].shape)  # some velocity everywhere
    og.bound(p)
    ucum = rng.normal(size=p.shape)
    return cfg, og, p, ucum, og.dump_rows(p, ucum)


def _full(f):
    F = np.zeros((len(f), 4, 4))
    for c, (a, b) in enumerate(((0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3))):
        F[:, a, b] = f[:, c]
        F[:, b, a] = -f[:, c]
    return F


def test_dump_row_identities():
    cfg, og, p, ucum, r = _rows()
    ucon, ucov = r[:, C["UCON"]:C["UCON"] + 4], r[:, C["UCOV"]:C["UCOV"] + 4]
    bcon, bcov = r[:, C["BCON"]:C["BCON"] + 4], r[:, C["BCOV"]:C["BCOV"] + 4]
    uu = ucon * ucov  # relative to the terms, which reach 1e3 where u_phi is large
    assert np.all(np.abs(uu.sum(1) + 1.0) / np.abs(uu).sum(1) < 1e-14)
    bu = np.abs((bcon * ucov).sum(1))
    scale = np.sqrt(np.abs((bcon * bcov).sum(1))) * np.abs(ucon).max(1) * np.abs(ucov).max(1) + 1e-300
    assert np.all(bu / scale < 1e-12)
    Fu = _full(r[:, C["FCON"]:C["FCON"] + 6])
    Fl = _full(r[:, C["FCOV"]:C["FCOV"] + 6])
    fscale = np.abs(Fu).max((1, 2))[:, None] * np.abs(ucov).max(1)[:, None] + 1e-300
    assert np.all(np.abs(np.einsum("zmn,zn->zm", Fu, ucov)) / fscale < 1e-12)
    bscale = np.abs(Fu).max((1, 2))[:, None] * np.abs(bcov).max(1)[:, None] + 1e-300
    assert np.all(np.abs(np.einsum("zmn,zn->zm", Fu, bcov)) / bscale < 1e-12)
    # raise the lowered F again with g^{ab}: F^{ab} = g^{am} g^{bn} F_{mn}
    G = og.geom[0][og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1]
    gcon = np.zeros((cfg.n2, cfg.n1, 4, 4))
    idx = [(0, 0), (0, 1), (0, 2), (0, 3), (1, 1), (1, 2), (1, 3), (2, 2), (2, 3), (3, 3)]
    for m, (a, b) in enumerate(idx):
        gcon[..., a, b] = gcon[..., b, a] = G[..., 10 + m]
    gz = np.broadcast_to(gcon[None], (cfg.n3, cfg.n2, cfg.n1, 4, 4)).reshape(-1, 4, 4)
    Fr = np.einsum("zam,zbn,zmn->zab", gz, gz, Fl)
    assert np.allclose(Fr, Fu, rtol=0, atol=1e-10 * np.abs(Fu).max())
    # EOS columns
    gam = cfg.gam
    rho, u = r[:, 0], r[:, 1]
    pres = (gam - 1) * u
    n = 1.0 / (gam - 1)
    assert np.array_equal(r[:, C["PRESS"]], pres)
    assert np.allclose(r[:, C["CS2"]], gam * pres / (rho + u + pres), rtol=1e-14)
    assert np.allclose(r[:, C["SDEN"]], rho * np.log(pres ** n / rho ** (n + 1)), rtol=1e-12)
    # characteristic speeds, copies
    v = r[:, C["V"]:C["V"] + 6].reshape(-1, 3, 2)
    assert np.all(v[:, :, 0] <= v[:, :, 1])
    act = lambda x: x[og.ng3:og.ng3 + cfg.n3, og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1].reshape(-1, 8)
    assert np.array_equal(r[:, C["U"]:C["U"] + 8], act(ucum))
    assert np.array_equal(r[:, :8], act(p))
    assert np.array_equal(r[:, C["GDET"]], np.broadcast_to(G[None, ..., 21], (cfg.n3, cfg.n2, cfg.n1)).ravel())


def test_entropy_special_cases():
    cfg = H.default_config(8, 6, 1, recon="para", flux="laxf", timeorder=2)
    og = _oracle.OGrid(cfg)
    p = og.zeros()
    p[..., 0], p[..., 1] = 1.0, 1.0
    # zone (0,0): rho below SMALL; zone (1,0): u = 0 (p below SMALL); zone (2,0): both
    a = p[0, og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1]
    a[0, 0, 0] = 0.0
    a[0, 1, 1] = 0.0
    a[0, 2, :2] = 0.0
    r = og.dump_rows(p, og.zeros())
    assert r[0, C["SDEN"]] == 0.0
    assert r[1, C["SDEN"]] == -1e100
    assert r[2, C["SDEN"]] == 0.0
