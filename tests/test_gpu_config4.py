"""The uneven batch at its own workload: 64 independent 2-D problems, grids
drawn from {64^2, 96^2, 128^2, 192^2, 256^2}, Fishbone-Moncrief tori (even
seeds; spins 0-0.99, beta 30-300) and Bondi flows (odd seeds), PARA / LAX-F /
RK2, stepped together (one launch per stage over all grids, the RK step
replayed from a hipGraph, tile shapes autotuned per grid) for 500 steps.

  * every member equals the same grid stepped alone, bit for bit, after 500
    steps (and its clock too);
  * every member's step agrees with its own CPU solve (the oracle) of that
    step from the same state, per zone to 1e-12, at steps 1, 101, 251 and
    500 of the run (the per-step comparison: over many steps the two
    trajectories part at the rate the flow amplifies the last-bit
    differences of the boundaries' exp and log, device against glibc).

Tolerance: tests/_parity's per-zone bound.  In the Bondi flows the polar
velocity is rounding noise around zero, so the three velocities share one
absolute floor, 1e-14 of the largest velocity component (a per-component
floor would be a floor on noise).
"""
import os

import numpy as np
import pytest

import _oracle
import harm_harm_amd as H
from _parity import FLOOR, PRIM_POSITIVE, zone_ratio
from harm_harm_amd.batch import Batch
from harm_harm_amd.grid import Grid

pytestmark = pytest.mark.gpu

NGRIDS, STEPS = 64, 500
CHECK_AT = (0, 100, 250, 499)  # steps before each compared step
SIZES = [64, 96, 128, 192, 256]


def members():
    rng = np.random.default_rng(2026)
    out = []
    for seed in range(NGRIDS):
        n = int(rng.choice(SIZES))
        kind = "torus" if seed % 2 == 0 else "bondi"
        out.append((kind, n, float(rng.uniform(0.0, 0.99)), float(np.exp(rng.uniform(np.log(30), np.log(300)))),
                    seed))
    return out


def config(kind, n, a):
    if kind == "torus":
        c = H.default_config(n, n, 1, recon="para", flux="laxf", timeorder=2)
        c.a = a
        return c
    return H.bondi_config(n, n, recon="para", flux="laxf", timeorder=2)


def prim_ratio(kind, a, b):
    """per-zone ratio (<= 1 passes) of active primitives a against b"""
    r = zone_ratio(a, b, positive=PRIM_POSITIVE)
    if kind == "bondi":
        # velocities against one floor from the largest velocity component
        vs = float(np.abs(b[..., 2:5]).max())
        others = zone_ratio(np.delete(a, [2, 3, 4], axis=-1), np.delete(b, [2, 3, 4], axis=-1), positive=(0, 1))
        dv = np.abs(a[..., 2:5] - b[..., 2:5])
        den = 1e-12 * np.abs(b[..., 2:5]) + FLOOR * vs
        r = max(others, float(np.where(dv == 0, 0.0, dv / den).max()))
    return r


def test_uneven_batch_of_64_grids_500_steps():
    ms = members()
    cfgs = [config(kind, n, a) for kind, n, a, _, _ in ms]
    B = Batch(cfgs)
    init = []
    for g, (kind, n, a, beta, seed) in zip(B.grids, ms):
        if kind == "torus":
            g.init_torus(beta=beta, seed=seed)
        else:
            g.init_bondi()
        ndt = g.fluxcalc("p")
        dt = 0.5 / sum(1.0 / x for x in ndt)
        g.set_time(0.0, dt)
        init.append((g.get("p"), dt))
    B.tune(reps=2)
    print("members:", [(k, n) for k, n, _, _, _ in ms], flush=True)

    threads = min(16, os.cpu_count() or 1)
    done = 0
    for at in CHECK_AT:
        if at > done:
            B.step(at - done, graph=True)
            done = at
        B.sync()
        starts = [(g.get("p"), g.time()[1]) for g in B.grids]
        B.step(1, graph=True)
        done += 1
        B.sync()
        worst = 0.0
        for i, (g, c, (kind, n, _, _, _), (p0, dt)) in enumerate(zip(B.grids, cfgs, ms, starts)):
            og = _oracle.OGrid(c, nthreads=threads)
            st = og.new_state(p0)
            dt = og.step(st, dt)
            r = prim_ratio(kind, g.active(g.get("p")), og_active(og, st["p"]))
            worst = max(worst, r)
            assert r <= 1.0, (done, i, kind, n, r)
            assert abs(g.time()[1] - dt) <= 1e-12 * dt, (done, i, g.time()[1], dt)
        print(f"step {done}: all {NGRIDS} members against their CPU solve of the step, worst per-zone ratio {worst}",
              flush=True)
    assert done == STEPS

    # every member against the grid stepped alone
    for i, (g, c, (p0, dt)) in enumerate(zip(B.grids, cfgs, init)):
        s = Grid(c)
        s.set("p", p0)
        s.set_time(0.0, dt)
        s.step(STEPS)
        assert np.array_equal(g.get("p"), s.get("p")), i
        assert g.time() == s.time(), i
        s.close()
    print(f"all {NGRIDS} members equal their single-grid runs after {STEPS} steps", flush=True)
    B.close()


def og_active(og, p):
    return p[:, og.ng2:og.ng2 + og.cfg.n2, og.ng1:og.ng1 + og.cfg.n1]
