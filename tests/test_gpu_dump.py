"""HARM-format dumps from the device (harm_dump_columns, harm_write_dump,
harm_gdump_columns, harm_write_gdump) through the C ABI.

* the main dump's physics columns against the oracle's o_dump_row (pinned in
  test_oracle_dump.py) after three steps, per column to 1e-12 of the
  column's scale (the inversion, state and EOS columns are bit-exact; the
  fast speeds and |div B| of the device's own arithmetic), the coordinates
  against harm_get_coords, the current columns 0 before any harm_currents;
* the four-current columns after harm_currents at the start and after two
  steps (the third call takes the time derivative) against the oracle's
  o_current_precalc / o_current_calc on the same states, and J_mu the
  lowered J^mu;
* the written file parses back to the reference's header (30 fields, with
  the column count last) and one row of 72 values per zone equal to the
  columns at the printed precision;
* the grid dump: g^{ab} g_{bc} = delta, Gamma^a_{bc} symmetric in b, c and
  equal to the library's packed table, dV/dX the closed-form MKS Jacobian.
"""
import ctypes
import math

import numpy as np
import pytest

import _oracle
import harm_harm_amd as H
from harm_harm_amd.grid import Grid

pytestmark = pytest.mark.gpu


def _state(n=(24, 16, 8), recon="weno5", flux="hll", timeorder=3):
    cfg = H.default_config(*n, recon=recon, flux=flux, timeorder=timeorder)
    og = _oracle.OGrid(cfg)
    p = og.init_torus()
    rng = np.random.default_rng(3)
    a = p[og.ng3:og.ng3 + cfg.n3, og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1]
    a[..., :5] *= 1.0 + 0.02 * rng.uniform(-1, 1, a[..., :5].shape)
    og.bound(p)
    g = Grid(cfg)
    g.set("p", p)
    g.set_time(0.0, 0.5 * min(og.fluxcalc(p, *[og.zeros() for _ in range(3)])))
    g.step(3)
    return cfg, og, g


@pytest.mark.parametrize("case", [((24, 16, 8), "weno5", "hll", 3), ((32, 24, 1), "para", "laxf", 2)])
def test_dump_columns_match_oracle(case):
    cfg, og, g = _state(*case)
    rows = g.dump_columns()
    ref = og.dump_rows(g.get("p"), g.get("ucum"))
    dev = np.concatenate([rows[:, 9:52], rows[:, 60:72]], axis=1)
    assert dev.shape == ref.shape
    worst = 0.0
    for c in range(ref.shape[1]):
        s = np.abs(ref[:, c]).max()
        err = np.abs(dev[:, c] - ref[:, c]).max() / (s if s > 0 else 1.0)
        worst = max(worst, err)
        assert err <= 1e-12, (c, err)
    print("worst column relative difference", worst)
    assert np.all(rows[:, 52:60] == 0.0)
    n1, n2 = cfg.n1, cfg.n2
    for t in sorted({0, 1, n1 + 2, min(n1 * n2 + 5, len(rows) - 1), len(rows) - 1}):
        i, j, k = t % n1, (t // n1) % n2, t // (n1 * n2)
        assert rows[t, 0] == i and rows[t, 1] == j and rows[t, 2] == k
        Xc, Vc = (ctypes.c_double * 4)(), (ctypes.c_double * 4)()
        H.check(H.lib().harm_get_coords(g.h, i, j, k, Xc, Vc), "coords")
        assert list(rows[t, 3:6]) == list(Xc[1:4]) and list(rows[t, 6:9]) == list(Vc[1:4])


@pytest.mark.parametrize("case", [((24, 16, 8), "weno5", "hll", 3), ((32, 24, 1), "para", "laxf", 2)])
def test_current_columns_match_oracle(case):
    cfg, og, g = _state(*case)
    cf = og.current_slots()
    calls, lastdt = 0, 0.0
    for n in range(3):
        if n == 0:
            g.currents(at_start=True)
            dt = g.time()[1]
        else:
            dt = g.time()[1]  # the dt the next step takes
            g.step(1)
            g.currents()
        og.current_precalc(g.get("p"), cf)
        J = og.current_calc(cf, calls >= 2, lastdt, dt)
        calls, lastdt = calls + 1, dt
        rows = g.dump_columns()
        jc = rows[:, 52:56]
        want = J.reshape(-1, 4)
        for m in range(4):
            s_ = np.abs(want[:, m]).max()
            err = np.abs(jc[:, m] - want[:, m]).max() / (s_ if s_ > 0 else 1.0)
            print(case, "call", n + 1, "J component", m, "scale", s_, "relative difference", err)
            assert err <= 1e-12, (n, m, err)
        # J_mu = g_{mu nu} J^nu at the zone centre
        gc = og.geom[0][og.ng2:og.ng2 + cfg.n2, og.ng1:og.ng1 + cfg.n1, :10]
        sym = [[0, 1, 2, 3], [1, 4, 5, 6], [2, 5, 7, 8], [3, 6, 8, 9]]
        G = np.array([[gc[..., sym[a][b]] for b in range(4)] for a in range(4)])  # (4, 4, n2, n1)
        Jz = J  # (n3, n2, n1, 4)
        low = np.einsum("abji,kjib->kjia", G, Jz).reshape(-1, 4)
        assert np.allclose(rows[:, 56:60], low, rtol=1e-12, atol=1e-12 * max(np.abs(low).max(), 1e-300))
    assert np.abs(rows[:, 53:56]).max() > 0.0  # the time term and the spatial terms are there


def test_write_dump_format(tmp_path):
    cfg, og, g = _state()
    path = tmp_path / "dump0000"
    g.write_dump(path)
    rows = g.dump_columns()
    with open(path) as f:
        head = f.readline().split()
        body = np.loadtxt(f)
    assert len(head) == 30
    t, dt = g.time()
    assert float(head[0]) == pytest.approx(t, rel=1e-14)
    assert [int(x) for x in head[1:4]] == [cfg.n1, cfg.n2, cfg.n3]
    assert int(head[10]) == g.nstep() and float(head[17]) == pytest.approx(dt, rel=1e-14)
    assert [int(x) for x in head[21:27]] == [0, cfg.n1 - 1, 0, cfg.n2 - 1, 0, cfg.n3 - 1]
    assert int(head[27]) == 3 and int(head[29]) == 72
    assert body.shape == rows.shape
    assert np.allclose(body, rows, rtol=1e-14, atol=0)


def test_gdump_identities(tmp_path):
    cfg = H.default_config(16, 12, 8, recon="para", flux="laxf", timeorder=2)
    g = Grid(cfg)
    r = g.gdump_columns()
    conn = r[:, 9:73].reshape(-1, 4, 4, 4)
    gcon = r[:, 73:89].reshape(-1, 4, 4)
    gcov = r[:, 89:105].reshape(-1, 4, 4)
    J = r[:, 110:126].reshape(-1, 4, 4)
    eye = np.einsum("zab,zbc->zac", gcon, gcov)
    assert np.allclose(eye, np.eye(4)[None], atol=1e-11)
    assert np.array_equal(conn, conn.transpose(0, 1, 3, 2))
    assert np.all(r[:, 106:110] == 0.0)
    packed = g.connection()  # (ns2, ns1, 40), Gamma^a_{bc} at a*10 + sym(b, c)
    sym = [[0, 1, 2, 3], [1, 4, 5, 6], [2, 5, 7, 8], [3, 6, 8, 9]]
    for t in (0, 7, 40, len(r) - 1):
        i, j = t % cfg.n1, (t // cfg.n1) % cfg.n2
        row = packed[j + g.ng2, i + g.ng1]
        for a_ in range(4):
            for b in range(4):
                for c in range(4):
                    assert conn[t, a_, b, c] == row[a_ * 10 + sym[b][c]]
    X1, X2 = r[:, 3], r[:, 4]
    assert np.allclose(J[:, 1, 1], np.exp(X1), rtol=1e-14)  # r = R0 + exp(X1)
    assert np.allclose(J[:, 2, 2], math.pi * (1.0 + (1.0 - cfg.hslope) * np.cos(2 * math.pi * X2)), rtol=1e-13)
    assert np.allclose(J[:, 3, 3], 2 * math.pi, rtol=1e-15)
    off = J.copy()
    for m in range(4):
        off[:, m, m] = 0.0
    assert np.all(off == 0.0)
    assert np.array_equal(r[:, 105], g.geometry(0)[g.ng2:g.ng2 + cfg.n2, g.ng1:g.ng1 + cfg.n1, 21][None]
                          .repeat(cfg.n3, 0).ravel())
    path = tmp_path / "gdump"
    g.write_gdump(path)
    with open(path) as f:
        head = f.readline().split()
        body = np.loadtxt(f)
    assert int(head[27]) == 4 and int(head[29]) == 126 and body.shape == r.shape
