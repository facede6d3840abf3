"""Batched stepping of independent grids: every grid of a batch advances
exactly as it would alone (same kernels, bit for bit), graph replays equal
eager launches, tuned tile shapes change nothing, and batch members agree with
the oracle."""
import numpy as np
import pytest

import _oracle
import harm_harm_amd as H
from harm_harm_amd.batch import Batch
from harm_harm_amd.grid import Grid

pytestmark = pytest.mark.gpu

# (n1, n2, spin, beta, seed): uneven 2-D tori
MEMBERS = [(32, 32, 0.9375, 100.0, 1), (48, 40, 0.5, 30.0, 2), (24, 28, 0.0, 300.0, 3), (40, 32, 0.99, 60.0, 4)]


def cfgs(recon="para", flux="laxf", timeorder=2):
    out = []
    for n1, n2, a, beta, seed in MEMBERS:
        c = H.default_config(n1, n2, 1, recon=recon, flux=flux, timeorder=timeorder)
        c.a = a
        out.append(c)
    return out


def init(grids):
    for g, (n1, n2, a, beta, seed) in zip(grids, MEMBERS):
        g.init_torus(beta=beta, seed=seed)
        ndt = g.fluxcalc("p")
        g.set_time(0.0, 0.5 / sum(1.0 / x for x in ndt))


@pytest.mark.parametrize("scheme", [("para", "laxf", 2), ("weno5", "hll", 3)])
def test_batch_equals_single_grids(scheme):
    cs = cfgs(*scheme)
    B = Batch(cs)
    init(B.grids)
    singles = [Grid(c) for c in cs]
    for g, s in zip(B.grids, singles):
        s.set("p", g.get("p"))
        s.set_time(*g.time())
    for n in range(5):
        B.step(1, graph=False)
        for s in singles:
            s.step(1)
        for i, (g, s) in enumerate(zip(B.grids, singles)):
            assert np.array_equal(g.get("p"), s.get("p")), (scheme, n, i)
            assert g.time() == s.time()
    for s in singles:
        s.close()
    B.close()


def test_graph_replays_equal_eager():
    cs = cfgs("weno5", "hll", 3)
    A, E = Batch(cs), Batch(cs)
    init(A.grids)
    for a, e in zip(A.grids, E.grids):
        e.set("p", a.get("p"))
        e.set_time(*a.time())
    for n in range(4):
        A.step(1, graph=True)
        E.step(1, graph=False)
        A.sync()
        E.sync()
        for i, (a, e) in enumerate(zip(A.grids, E.grids)):
            # eager host work between replays
            assert np.array_equal(a.get("p"), e.get("p")), (n, i)
            assert a.time() == e.time()
    A.close()
    E.close()


def test_tuned_tiles_change_nothing():
    cs = cfgs("weno5", "hll", 3)
    A, T = Batch(cs), Batch(cs)
    init(A.grids)
    for a, t in zip(A.grids, T.grids):
        t.set("p", a.get("p"))
        t.set_time(*a.time())
    T.tune(reps=2)
    A.step(3, graph=True)
    T.step(3, graph=True)
    A.sync()
    T.sync()
    for a, t in zip(A.grids, T.grids):
        assert np.array_equal(a.get("p"), t.get("p"))


def relmax(a, b):
    a = a.reshape(-1, a.shape[-1])
    b = b.reshape(-1, b.shape[-1])
    out = 0.0
    for v in range(a.shape[-1]):
        sc = max(np.abs(a[:, v]).max(), np.abs(b[:, v]).max())
        if sc > 0:
            out = max(out, np.abs(a[:, v] - b[:, v]).max() / sc)
    return out


@pytest.mark.parametrize("scheme", [("para", "laxf", 2), ("weno5", "hll", 3)])
def test_batch_members_match_oracle(scheme):
    cs = cfgs(*scheme)
    B = Batch(cs)
    init(B.grids)
    oracle = []
    for g, c in zip(B.grids, cs):
        og = _oracle.OGrid(c)
        st = og.new_state(g.get("p"))
        dt = g.time()[1]
        for n in range(4):
            dt = og.step(st, dt)
        oracle.append((og, st["p"], dt))
    B.step(4, graph=True)
    B.sync()
    for i, (g, (og, po, dto)) in enumerate(zip(B.grids, oracle)):
        a = g.get("p")[:, og.ng2:og.ng2 + og.cfg.n2, og.ng1:og.ng1 + og.cfg.n1]
        b = po[:, og.ng2:og.ng2 + og.cfg.n2, og.ng1:og.ng1 + og.cfg.n1]
        err = relmax(a, b)
        print(scheme, "member", i, "relmax", err, "dt", g.time()[1], dto)
        assert err <= 1e-12
        assert abs(g.time()[1] - dto) <= 1e-12 * dto
    B.close()


def test_single_grid_graph_replays_equal_eager():
    """harm_step replays a captured graph unless kernel timing is on; both
    paths must give the same bits, replay after replay, with host work
    between replays."""
    c = H.default_config(40, 32, 8, recon="weno5", flux="hll", timeorder=3)
    A, E = Grid(c), Grid(c)
    A.init_torus(seed=5)
    ndt = A.fluxcalc("p")
    A.set_time(0.0, 0.5 / sum(1.0 / x for x in ndt))
    E.set("p", A.get("p"))
    E.set_time(*A.time())
    E.profile(True)
    for n in range(4):
        A.step(1)
        E.step(1)
        assert np.array_equal(A.get("p"), E.get("p")), n
        assert A.time() == E.time()
        assert np.array_equal(np.array(A.boundary_fluxes()), np.array(E.boundary_fluxes()))
    A.close()
    E.close()
