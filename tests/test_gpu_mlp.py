"""GPU tests of the learned inversion guess: the forward and backward kernels
against the CPU reference (oracle/mlp_ref.py, BF16 rounding points
emulated), deterministic training, and the seeded Newton solve.

Tolerances: the GPU and CPU sides round the same operands to BF16 but may
accumulate products in a different order (FP32 on the matrix cores), so a
hidden activation can land one BF16 ulp apart; 1e-3 of each quantity's
largest magnitude covers that.
"""
import os
import sys

import numpy as np
import pytest

import harm_harm_amd as H
from harm_harm_amd.mlp import MLP, NPARAM
from harm_harm_amd.sampling import sample_inversion_inputs

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402

pytestmark = pytest.mark.gpu
D = H.DeviceArray


def random_norm(rng):
    return np.concatenate([rng.normal(0, 0.5, 8), rng.uniform(0.5, 2, 8), rng.normal(0, 1, 2),
                           rng.uniform(0.5, 2, 2)]).astype(np.float32)


def test_forward_matches_cpu():
    rng = np.random.default_rng(0)
    m = MLP(seed=3)
    w, _ = m.params()
    nm = random_norm(rng)
    m.set_params(None, nm)
    n = 1000  # not a multiple of the block size
    X = rng.normal(0, 1.5, (n, 8)).astype(np.float32)
    Y = m.forward(D.from_numpy(X), n).numpy()
    Xn, _ = mlp_ref.standardise(X, None, nm)
    Yr, _ = mlp_ref.forward(w, Xn)
    Yr = Yr * nm[18:20] + nm[16:18]
    err = np.abs(Y - Yr).max() / np.abs(Yr).max()
    print("forward relmax", err)
    assert err <= 1e-3


def test_gradients_match_cpu():
    rng = np.random.default_rng(1)
    m = MLP(seed=5)
    w = rng.normal(0, 0.3, NPARAM).astype(np.float32)
    nm = random_norm(rng)
    m.set_params(w, nm)
    X = rng.normal(0, 1.5, (300, 8)).astype(np.float32)
    T = rng.normal(0, 1, (300, 2)).astype(np.float32)
    order = rng.choice(300, 128, replace=False).astype(np.int32)
    g, loss = m.grad(D.from_numpy(X), D.from_numpy(T), 300, D.from_numpy(order))
    Xn, Tn = mlp_ref.standardise(X[order], T[order], nm)
    lr, gr = mlp_ref.loss_and_grads(w, Xn, Tn)
    assert abs(loss - lr) <= 1e-4 * lr
    for k, (o, sh) in mlp_ref.OFF.items():
        a, b = g[o:o + int(np.prod(sh))], gr[o:o + int(np.prod(sh))]
        err = np.abs(a - b).max() / np.abs(b).max()
        cos = np.dot(a, b) / np.linalg.norm(a) / np.linalg.norm(b)
        print(k, "relmax", err, "cos", cos)
        assert err <= 1e-3 and cos >= 1 - 1e-6


def solved_pairs(n, seed):
    p, guess, geom = sample_inversion_inputs(n, seed=seed)
    dp, dg, dU = D.from_numpy(p), D.from_numpy(geom), D(p.shape, np.float64)
    H.check(H.lib().harm_prim_to_cons_batch(dp.ptr, dg.ptr, dU.ptr, n, 4.0 / 3.0, None), "p2u")
    dq = D.from_numpy(guess)
    fl, it = D.zeros(n, np.int32), D.zeros(n, np.int32)
    H.check(H.lib().harm_invert_batch(dU.ptr, dg.ptr, dq.ptr, fl.ptr, it.ptr, n, 4.0 / 3.0, 2500.0, None), "inv")
    H.synchronize()
    return p, guess, dg, dU, dq, fl.numpy(), it.numpy()


@pytest.fixture(scope="module")
def trained():
    n = 1 << 16
    _, _, dg, dU, dq, fl, _ = solved_pairs(n, seed=11)
    assert (fl != 0).sum() == 0
    X, T, ok = MLP.pairs(dU, dg, dq, n)
    okh = ok.numpy()
    assert okh.sum() == n
    order = np.nonzero(okh)[0].astype(np.int32)
    np.random.default_rng(0).shuffle(order)
    dorder = D.from_numpy(order)
    models = []
    for rep in range(2):
        m = MLP(seed=1)
        m.fit_norm(X, T, ok, n)
        _, loss0 = m.grad(X, T, n, dorder)  # untrained loss on the first minibatch
        losses = m.train(X, T, n, dorder, 3000, lr=3e-3, lr_end=3e-5, chunk=500)
        models.append((m, [loss0] + losses))
    return models


def test_training_reduces_loss_and_is_reproducible(trained):
    (m1, l1), (m2, l2) = trained
    print("untrained loss and mean loss per 500 steps", l1)
    assert l1[-1] < 0.1 * l1[0]
    w1, n1 = m1.params()
    w2, n2 = m2.params()
    assert np.array_equal(w1, w2) and np.array_equal(n1, n2) and l1 == l2


def test_seeded_newton_converges_like_plain(trained):
    """Every zone converges from the network's W' (no extra failures), to the
    plain solve's accuracy against the true primitives; the mean iteration
    counts of both are reported (the gain is small: the inverter's stopping
    test for ultra-relativistic zones needs |dW'/W'| <= machine epsilon and
    takes 11-13 iterations from any start)."""
    m, _ = trained[0]
    n = 1 << 14
    p, guess, dg, dU, dq, fl_plain, it_plain = solved_pairs(n, seed=7)
    plain = dq.numpy()
    ds = D.from_numpy(guess)
    fl, it, fb = D.zeros(n, np.int32), D.zeros(n, np.int32), D.zeros(1, np.uint64)
    m.invert_batch(dU, dg, ds, fl, it, n, dfallback=fb)
    H.synchronize()
    seeded, fls, its = ds.numpy(), fl.numpy(), it.numpy()
    print("mean iterations plain", it_plain.mean(), "seeded", its.mean(), "fallbacks", int(fb.numpy()[0]))
    assert (fls != 0).sum() <= (fl_plain != 0).sum()
    assert (fls != 0).sum() == 0
    for v in range(5):
        e_plain = np.abs(plain[v] - p[v]).max() / np.abs(p[v]).max()
        e_seed = np.abs(seeded[v] - p[v]).max() / np.abs(p[v]).max()
        print("var", v, "error vs truth: plain", e_plain, "seeded", e_seed)
        assert e_seed <= 10 * e_plain + 1e-12


def test_seeded_step_matches_plain_step():
    """The learned guess inside the step (harm_set_inversion_guess): a network
    trained on zones sampled from a grid after plain steps seeds every
    inversion of the next step; the step's primitives agree with the plain
    step's to 1e-12 (both Newton solves converge to the same root within the
    inverter's tolerance), with no extra inversion failures."""
    from harm_harm_amd.grid import Grid
    cfg = H.default_config(24, 16, 8, recon="weno5", flux="hll", timeorder=3)
    g = Grid(cfg)
    g.init_torus(seed=2)
    ndt = g.fluxcalc("p")
    g.set_time(0.0, 0.5 / sum(1.0 / x for x in ndt))
    g.step(3)
    n = 1 << 15
    dU, dg, dp = g.sample_zones(n, seed=5)
    X, T, ok = MLP.pairs(dU, dg, dp, n, gammafail=cfg.gammafail)
    okh = ok.numpy()
    order = np.nonzero(okh)[0].astype(np.int32)
    np.random.default_rng(0).shuffle(order)
    m = MLP(seed=1)
    m.fit_norm(X, T, ok, n)
    m.train(X, T, n, D.from_numpy(order), 2000, lr=3e-3, lr_end=3e-5, chunk=500)
    p0, (t0, dt0) = g.get("p"), g.time()
    g.set_time(t0, dt0)
    g.step(1)
    plain = g.get("p")
    (ns_plain, nf_plain), _ = g.newton_stats()
    g.set("p", p0)
    g.set_time(t0, dt0)
    g.set_inversion_guess(m)
    g.step(1)
    seeded = g.get("p")
    (ns_seed, nf_seed), _ = g.newton_stats()
    fb, _ = g.guess_stats()
    g.set_inversion_guess(None)
    zones = cfg.n1 * cfg.n2 * cfg.n3 * 3
    err = 0.0
    a, b = g.active(seeded), g.active(plain)
    for v in range(8):
        s = np.abs(b[..., v]).max()
        if s > 0:
            err = max(err, float(np.abs(a[..., v] - b[..., v]).max() / s))
    print("iterations per zone-substep: plain", ns_plain / zones, "seeded", ns_seed / zones,
          "failures", nf_plain, nf_seed, "fallbacks", fb, "relmax", err)
    assert nf_seed <= nf_plain
    assert err <= 1e-12


def test_seeded_step_energy_accounting():
    """The learned-guess inversion's accounting variant keeps the same fix-up
    states as the plain one: with tracking on, one step seeded by an
    (untrained) network and one plain step from the same state give the same
    energy-file record (floor and gamma-limit sums to 1e-10 of their scale;
    the seeded and plain Newton roots agree within the inverter's
    tolerance, so the sums are not bit-identical)."""
    from harm_harm_amd.grid import Grid
    cfg = H.default_config(24, 16, 8, recon="weno5", flux="hll", timeorder=3)
    g = Grid(cfg)
    g.init_torus(seed=2)
    ndt = g.fluxcalc("p")
    g.set_time(0.0, 0.5 / sum(1.0 / x for x in ndt))
    g.step(2)
    p0, (t0, dt0), n0 = g.get("p"), g.time(), g.nstep()
    recs = []
    for guess in (None, MLP(seed=3)):
        g.set("p", p0)
        g.set_time(t0, dt0, n0)
        g.set_inversion_guess(guess)
        g.set_ener_tracking(True)
        g.step(1)
        recs.append(g.ener())
        g.set_ener_tracking(False)
    g.set_inversion_guess(None)
    a, b = recs
    for k in ("U", "pdot", "pcum", "sourceadd", "fladd"):
        s = np.abs(a[k]).max()
        err = np.abs(a[k] - b[k]).max() / (s if s > 0 else 1.0)
        print(k, "scale", s, "relative difference", err)
        assert err <= 1e-10, k
    for c in (1, 2):  # floors, gamma limit
        s = np.abs(a["fladdterms"][c]).max()
        assert np.abs(a["fladdterms"][c] - b["fladdterms"][c]).max() <= 1e-10 * (s if s > 0 else 1.0)
    assert np.any(a["fladdterms"][1] != 0.0)
