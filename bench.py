#!/usr/bin/env python3
"""Benchmark: zone-updates/s of the HARM advance() step on one MI355X.

Workload: the 3-D Fishbone-Moncrief torus in MKS Kerr-Schild coordinates at
256x128x64 (r, theta, phi), FP64, a=0.9375, Gamma=4/3, WENO5BND reconstruction,
HLL fluxes, Toth flux-CT, RK3, Courant 0.8, seed-1 internal-energy noise,
rescaletype-0 floors.  A "step" is one full RK3 step (3 substeps: fluxes, CT,
update, inversion, floors, bounds, failure averaging, dt).  All state is
resident in HBM before the timed region; the timed region is K steps between
two device synchronisations.

Prints one JSON line: the throughput, a roofline record for the dominant kernel
(algorithmic HBM bytes per launch over its event-measured mean duration, with
the measured traffic and FP64 rate from the committed counter summary), the
whole step against HBM, the chosen form's learned inversion guess timed on the
same state (`chosen_form`), and a CPU baseline (the C restatement, OpenMP,
timed on this host's cores on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
FP64_PEAK_TFLOPS = 78.6  # MI355X FP64 (vector and matrix)
GRID = (256, 128, 64)
COUNTERS = os.path.join(ROOT, "profiles", "counters_step_torus256x128x64_weno5bnd_hll_rk3.md")


def measured_counters(path=COUNTERS):
    """Per-launch HBM traffic and FP64 FLOPs of each kernel group, from the
    committed rocprofv3 --pmc summary (FETCH_SIZE / WRITE_SIZE scaled by the
    copy-kernel calibration; SQ_INSTS_VALU_*_F64).  None when absent."""
    try:
        lines = open(path).read().splitlines()
    except OSError:
        return None
    out = {}
    for i, ln in enumerate(lines):
        if ln.startswith("group traffic_bytes"):
            for row in lines[i + 1:]:
                f = row.split()
                if len(f) != 5:
                    break
                out[f[0]] = {"traffic": float(f[1]), "read": float(f[2]), "write": float(f[3]), "flops": float(f[4])}
    return out or None


def algorithmic_bytes(g, kernel):
    """Bytes each kernel group must move per launch at minimum (compulsory
    traffic of its arrays; 2-D geometry tables counted once per (i,j) face or
    zone column)."""
    n1, n2, n3 = g.cfg.n1, g.cfg.n2, g.cfg.n3
    e0, e1, e2 = int(n1 > 1), int(n2 > 1), int(n3 > 1)
    geom = 27 * 8
    if kernel.startswith("flux"):
        d = int(kernel[-1])
        if d == 1:
            ni, nj, nk = n1 + 1, n2 + 2 * e1, n3 + 2 * e2
        elif d == 2:
            ni, nj, nk = n1 + 2 * e0, n2 + 1, n3 + 2 * e2
        else:
            ni, nj, nk = n1 + 2 * e0, n2 + 2 * e1, n3 + 1
        faces = ni * nj * nk
        # PARA: read 8 primitives per cell of the swept block, write 8 fluxes
        # per face; WENO5BND: read the 2 x 8 face values the line pass wrote
        # instead of the primitives
        import harm_harm_amd as H
        return faces * ((128 if g.cfg.recon == H.RECON["weno5"] else 64) + 64) + ni * nj * geom
    zones = n1 * n2 * n3
    if kernel == "update":
        # inside a step every substep moves 448 B per zone: pb 64 + 3 face
        # fluxes 192 + (pi read, ui write | ui read, uf read) 128 + one of
        # uf write / ucum write 64; plus geometry and connection per column
        return zones * 448 + n1 * n2 * (geom + 40 * 8)
    if kernel == "invert":
        # U read, guess read, solution write, flag write
        return zones * (64 + 64 + 64 + 4) + n1 * n2 * geom
    if kernel == "emf_ct":
        return zones * (6 * 8 + 3 * 8 + 3 * 8 + 6 * 8)
    if kernel == "bound":
        # every ghost zone written once from a zone read once (8 doubles each)
        ghosts = (n1 + 2 * 4 * e0) * (n2 + 2 * 4 * e1) * (n3 + 2 * 4 * e2) - zones
        return ghosts * 128
    if kernel == "fixup":
        return zones * 4  # the failure flags; failed zones are few
    if kernel == "recon":
        # WENO5BND per substep, four timed launches: the dP pass (read the
        # primitives 64, write dP of three directions 24) and one line pass
        # per direction (read the primitives 64 and dP 8, write the left and
        # right face values 128); geometry per zone column.  Averaged per launch.
        ndir = e0 + e1 + e2
        per_substep = zones * (88 + ndir * 200) + n1 * n2 * geom
        return per_substep / (1 + ndir)
    return None


def initial_dt(g):
    ndt = g.fluxcalc("p", do_ct=True)
    wavedt = 1.0 / sum(1.0 / x for x in ndt)
    acc = g.update("p", "p", [1.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 0.0], wavedt, 0)
    return min(wavedt, acc)


def cpu_baseline(cfg, nsteps):
    """The oracle (C restatement of the same step, OpenMP) on this host."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    threads = min(16, os.cpu_count() or 1)
    og = _oracle.OGrid(cfg, nthreads=threads)
    p = og.init_torus(seed=1)
    ndt = og.fluxcalc(p, og.zeros(), og.zeros(), og.zeros())
    dt = 1.0 / sum(1.0 / x for x in ndt)
    st = og.new_state(p)
    t0 = time.time()
    for _ in range(nsteps):
        dt = og.step(st, dt)
    el = time.time() - t0
    zones = cfg.n1 * cfg.n2 * cfg.n3
    return {"value": zones * nsteps / el, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "sample": f"{nsteps} full RK3 steps of the same {cfg.n1}x{cfg.n2}x{cfg.n3} WENO5BND/HLL torus "
                      f"({zones * nsteps} zone-updates, {el:.1f} s) with the oracle's C restatement, "
                      f"OpenMP on {threads} host threads"}


def inversion_microbench(args):
    """Batched inversion on seeded synthetic zones (Lorentz factor 1-50,
    magnetisation 1e-4-1e3, u/rho 1e-6-1e2 log-uniform, random Kerr-Schild
    points): the plain Newton solve from a +-20% perturbed guess against the
    same solve seeded by the learned guess (trained first, on the GPU, on pairs
    from plain solves of an independent sample)."""
    import numpy as np
    import harm_harm_amd as H
    from harm_harm_amd.mlp import MLP
    from harm_harm_amd.sampling import sample_inversion_inputs
    D = H.DeviceArray
    n = args.n
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    t0 = time.time()
    p, guess, geom = sample_inversion_inputs(n, seed=7)
    log(f"sampled {n} zones in {time.time() - t0:.1f} s")
    dp, dg, dU = D.from_numpy(p), D.from_numpy(geom), D(p.shape, np.float64)
    H.check(H.lib().harm_prim_to_cons_batch(dp.ptr, dg.ptr, dU.ptr, n, 4.0 / 3.0, None), "p2u")
    dp.free()
    dgs = D.from_numpy(guess)
    dq = D(p.shape, np.float64)
    fl, it = D.zeros(n, np.int32), D.zeros(n, np.int32)
    e0, e1 = H.Event(), H.Event()

    def timed(fn, reps=3):
        best = None
        for _ in range(reps):
            dq.copy_from_device(dgs)
            H.synchronize()
            e0.record()
            fn()
            e1.record()
            H.synchronize()
            ms = e0.elapsed_ms(e1)
            best = ms if best is None else min(best, ms)
        return best

    ms_plain = timed(lambda: H.check(H.lib().harm_invert_batch(dU.ptr, dg.ptr, dq.ptr, fl.ptr, it.ptr, n, 4.0 / 3.0,
                                                               2500.0, None), "invert"))
    q_plain, f_plain, i_plain = dq.numpy(), fl.numpy(), it.numpy()
    # training pairs from an independent sample solved by the plain inversion
    ntr = args.train_n
    pt, gt, geot = sample_inversion_inputs(ntr, seed=11)
    dpt, dgt, dUt = D.from_numpy(pt), D.from_numpy(geot), D(pt.shape, np.float64)
    H.check(H.lib().harm_prim_to_cons_batch(dpt.ptr, dgt.ptr, dUt.ptr, ntr, 4.0 / 3.0, None), "p2u")
    dqt, flt, itt = D.from_numpy(gt), D.zeros(ntr, np.int32), D.zeros(ntr, np.int32)
    H.check(H.lib().harm_invert_batch(dUt.ptr, dgt.ptr, dqt.ptr, flt.ptr, itt.ptr, ntr, 4.0 / 3.0, 2500.0, None), "inv")
    X, T, ok = MLP.pairs(dUt, dgt, dqt, ntr)
    m = MLP(seed=1)
    m.fit_norm(X, T, ok, ntr)
    order = np.nonzero(ok.numpy())[0].astype(np.int32)
    np.random.default_rng(0).shuffle(order)
    dorder = D.from_numpy(order)
    H.synchronize()
    t0 = time.time()
    losses = m.train(X, T, ntr, dorder, args.train_steps, lr=3e-3, lr_end=3e-5, chunk=1000)
    H.synchronize()
    t_train = time.time() - t0
    fb = D.zeros(1, np.uint64)
    ms_mlp = timed(lambda: m.invert_batch(dU, dg, dq, fl, it, n, dfallback=fb))
    q_mlp, f_mlp, i_mlp = dq.numpy(), fl.numpy(), it.numpy()
    nfb = int(fb.numpy()[0]) // 3
    good = (f_plain == 0) & (f_mlp == 0)
    agree = float(max(np.abs(q_mlp[v] - q_plain[v])[good].max() / np.abs(q_plain[v])[good].max() for v in range(5)))
    err_plain = float(max(np.abs(q_plain[v] - p[v]).max() / np.abs(p[v]).max() for v in range(5)))
    err_mlp = float(max(np.abs(q_mlp[v] - p[v]).max() / np.abs(p[v]).max() for v in range(5)))
    cpu = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle
        ns = min(n, 1 << 23)
        threads = min(16, os.cpu_count() or 1)
        t0 = time.time()
        _oracle.invert_batch(_oracle.p2u_batch(p[:, :ns], geom[:, :ns]), geom[:, :ns], guess[:, :ns], nthreads=threads)
        el = time.time() - t0
        cpu = {"value": ns / el, "unit": "inversions/s", "cores": threads, "kind": "port",
               "sample": f"the first {ns} zones of the same sample, the oracle's C restatement of the same solve "
                         f"(incl. its p->U), OpenMP on {threads} host threads, {el:.2f} s"}
    out = {
        "metric": "inversions/s, batched conserved-to-primitive inversion (FP64 Newton), plain vs MLP-seeded",
        "value": n / (ms_plain / 1e3), "unit": "inversions/s", "n_gpus": 1, "steps": 3, "warmup": 0,
        "ms_per_step": ms_plain, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seed 7, log-uniform Lorentz factor / magnetisation / u over rho, random KS points)",
        "config": {"workload": f"batched inversion microbench, {n} zones", "guess": "truth perturbed by +-20%"},
        "plain": {"ms": ms_plain, "inversions_per_s": n / (ms_plain / 1e3), "mean_newton_iterations": float(i_plain.mean()),
                  "failures": int((f_plain != 0).sum()), "max_rel_error_vs_truth": err_plain},
        "mlp_seeded": {"ms": ms_mlp, "inversions_per_s": n / (ms_mlp / 1e3), "mean_newton_iterations": float(i_mlp.mean()),
                       "failures": int((f_mlp != 0).sum()), "fallbacks": nfb, "max_rel_error_vs_truth": err_mlp,
                       "max_rel_difference_vs_plain": agree, "train_pairs": ntr, "train_steps": args.train_steps,
                       "train_seconds": t_train, "final_loss": losses[-1]},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def batch_bench(args):
    """64 independent 2-D problems of uneven sizes (64^2 ... 256^2): Fishbone-
    Moncrief tori (even seeds; spins 0-0.99, beta 30-300) and Bondi flows (odd
    seeds; spin 0, the reference's Bondi boundary set), PARA / LAX-F / RK2,
    stepped together (one launch per stage over all grids, the RK step
    replayed from a hipGraph, tile shapes autotuned per grid) against the same
    grids stepped one after another with the single-grid step; the smallest
    torus and the smallest Bondi flow are each checked against their own CPU
    solve."""
    import numpy as np
    import harm_harm_amd as H
    from harm_harm_amd.batch import Batch
    from harm_harm_amd.grid import Grid
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    rng = np.random.default_rng(2026)
    sizes = [64, 96, 128, 192, 256]
    members = []
    for seed in range(args.ngrids):
        n = int(rng.choice(sizes))
        kind = "torus" if seed % 2 == 0 else "bondi"
        members.append((kind, n, float(rng.uniform(0.0, 0.99)), float(np.exp(rng.uniform(np.log(30), np.log(300)))),
                        seed))
    cfgs = []
    for kind, n, a, beta, seed in members:
        if kind == "torus":
            c = H.default_config(n, n, 1, recon="para", flux="laxf", timeorder=2)
            c.a = a
        else:
            c = H.bondi_config(n, n, recon="para", flux="laxf", timeorder=2)
        cfgs.append(c)
    zones = sum(n * n for _, n, _, _, _ in members)
    B = Batch(cfgs)
    init = []
    for g, (kind, n, a, beta, seed) in zip(B.grids, members):
        if kind == "torus":
            g.init_torus(beta=beta, seed=seed)
        else:
            g.init_bondi()
        ndt = g.fluxcalc("p")
        dt = 0.5 / sum(1.0 / x for x in ndt)
        g.set_time(0.0, dt)
        init.append((g.get("p"), dt))
    t0 = time.time()
    B.tune(reps=3)
    t_tune = time.time() - t0
    # batched, graph-replayed
    B.step(args.warmup, graph=True)
    B.sync()
    for g, (p0, dt) in zip(B.grids, init):
        g.set("p", p0)
        g.set_time(0.0, dt)
    B.sync()
    t0 = time.perf_counter()
    B.step(args.steps, graph=True)
    B.sync()
    el_b = time.perf_counter() - t0
    log(f"batched: {args.steps} steps of {len(members)} grids ({zones} zones) in {el_b:.3f} s")
    # the same grids one after another with the single-grid step
    singles = [Grid(c) for c in cfgs]
    for s_, (p0, dt) in zip(singles, init):
        s_.set("p", p0)
        s_.set_time(0.0, dt)
        s_.step(args.warmup)
        s_.sync()
        s_.set("p", p0)
        s_.set_time(0.0, dt)
    H.synchronize()
    t0 = time.perf_counter()
    for s_ in singles:
        s_.step(args.steps)
        s_.sync()
    el_s = time.perf_counter() - t0
    log(f"one by one: {el_s:.3f} s")
    same = all(np.array_equal(g.get("p"), s_.get("p")) for g, s_ in zip(B.grids, singles))
    # the smallest member of each kind against its own CPU solve
    checks = []
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle
        for kind in ("torus", "bondi"):
            ks = [i for i in range(len(members)) if members[i][0] == kind]
            if not ks:
                continue
            k = min(ks, key=lambda i: members[i][1])
            og = _oracle.OGrid(cfgs[k], nthreads=min(16, os.cpu_count() or 1))
            st = og.new_state(init[k][0].copy())
            dt = init[k][1]
            t0 = time.time()
            for _ in range(args.steps):
                dt = og.step(st, dt)
            el_c = time.time() - t0
            a = B.grids[k].get("p")[:, og.ng2:og.ng2 + og.cfg.n2, og.ng1:og.ng1 + og.cfg.n1]
            b = st["p"][:, og.ng2:og.ng2 + og.cfg.n2, og.ng1:og.ng1 + og.cfg.n1]
            # velocities against the largest velocity component (the Bondi flow's
            # polar velocity is rounding noise), the rest per variable
            vs = max(float(np.abs(b[..., 2:5]).max()), 1e-300)
            rel = max(float(np.abs(a[..., v] - b[..., v]).max()
                            / (vs if 2 <= v <= 4 else max(np.abs(b[..., v]).max(), 1e-300))) for v in range(8))
            n = members[k][1]
            checks.append({"grid": k, "problem": kind, "size": f"{n}x{n}", "steps": args.steps,
                           "max_rel_difference": rel, "cpu_zone_updates_per_s": n * n * args.steps / el_c,
                           "cpu_threads": min(16, os.cpu_count() or 1)})
    nt = sum(1 for m in members if m[0] == "torus")
    out = {
        "metric": "zone-updates/s, 64 uneven 2-D tori and Bondi flows stepped together (batched launches + hipGraph)"
                  " vs one by one",
        "value": zones * args.steps / el_b, "unit": "zone-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": el_b / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded torus initial data; exact Bondi flow)",
        "config": {"workload": f"{nt} 2-D FM tori and {len(members) - nt} Bondi flows, sizes {sizes}, "
                               f"PARA/LAX-F/RK2, {zones} zones", "tune_seconds": t_tune},
        "one_by_one": {"value": zones * args.steps / el_s, "ms_per_step": el_s / args.steps * 1e3,
                       "speedup_of_batch": el_s / el_b, "identical_results": same},
        "oracle_check": checks,
    }
    print(json.dumps(out))


def corner_divb_max(g, geom_cent):
    """Largest Toth corner divergence of gdet B over the interior corners
    (the quantity flux-CT keeps at round-off), relative to max|gdet B| / dx."""
    import numpy as np
    c = g.cfg
    p = g.get("p")  # (ns3, ns2, ns1, 8)
    ng1, ng2, ng3 = g.ng1, g.ng2, g.ng3
    gd = geom_cent[..., 21][None, :, :]
    dx = g.dx
    n1, n2, n3 = c.n1, c.n2, c.n3
    sl = lambda a, di, dj, dk: a[ng3 + dk:ng3 + dk + n3 - 1, ng2 + dj:ng2 + dj + n2 - 1, ng1 + di:ng1 + di + n1 - 1]
    div = 0.0
    scale = 0.0
    for m, s in ((0, dx[1]), (1, dx[2]), (2, dx[3])):
        B = p[..., 5 + m] * gd
        acc = 0.0
        for di in (0, 1):
            for dj in (0, 1):
                for dk in (0, 1):
                    sign = [di, dj, dk][m] * 2 - 1
                    acc = acc + sign * sl(B, di, dj, dk)
        div = div + acc / (4 * s)
        scale = max(scale, float(np.abs(sl(B, 0, 0, 0)).max()))
        del B, acc
    return float(np.abs(div).max()) / (scale / min(dx[1], dx[2], dx[3]))


def large_bench(args):
    """The 3-D torus at 384x192x192 (14.2 M zones) with the plain step:
    zone-updates/s, inversion failures, the corner divergence of B and the
    accumulated accretion sums after the timed steps.  With --mlp, the chosen
    form's learned guess: the network is trained on the GPU from zones of the
    state after the first 200 plain steps, then the same number of steps is
    timed from the same state plainly and with every inversion seeded by the
    network, and one step of each is compared."""
    import numpy as np
    import harm_harm_amd as H
    from harm_harm_amd.grid import Grid
    log = lambda *a: print(*a, file=sys.stderr, flush=True)
    cfg = H.default_config(384, 192, 192, recon="weno5", flux="hll", timeorder=3)
    g = Grid(cfg)
    g.init_torus(seed=1)
    dt0 = initial_dt(g)
    g.set_time(0.0, dt0)
    g.tune(reps=3)
    zones = cfg.n1 * cfg.n2 * cfg.n3
    mlp_out = None
    if args.mlp:
        from harm_harm_amd.mlp import MLP
        t0 = time.time()
        g.step(args.mlp_plain_steps)
        g.sync()
        n = args.mlp_samples
        dU, dgeo, dp = g.sample_zones(n, seed=11)
        X, T, ok = MLP.pairs(dU, dgeo, dp, n, gammafail=cfg.gammafail)
        okh = ok.numpy()
        order = np.nonzero(okh)[0].astype(np.int32)
        np.random.default_rng(0).shuffle(order)
        m = MLP(seed=1)
        m.fit_norm(X, T, ok, n)
        losses = m.train(X, T, n, H.DeviceArray.from_numpy(order), args.mlp_train_steps, lr=3e-3, lr_end=3e-5,
                         chunk=max(1, args.mlp_train_steps // 4))
        t_train = time.time() - t0
        log(f"mlp: {args.mlp_plain_steps} plain steps, {n} sampled zones, training losses {losses}")
    g.step(args.warmup)
    g.sync()
    p_start, (t_start, dt_start) = g.get("p"), g.time()
    n_start = g.nstep()
    g.set_time(t_start, dt_start, n_start)
    t0 = time.perf_counter()
    g.step(args.steps)
    g.sync()
    el = time.perf_counter() - t0
    (_, _), (nstroke, nfail) = g.newton_stats()
    steps_done = g.nstep()
    t_sim, dt_now = g.time()
    divb = corner_divb_max(g, g.geometry(0))
    fa = g.boundary_fluxes()
    if args.mlp:
        # the same steps from the same state with the learned guess
        g.set("p", p_start)
        g.set_time(t_start, dt_start, n_start)
        g.set_inversion_guess(m)
        g.step(1)  # capture
        g.set("p", p_start)
        g.set_time(t_start, dt_start, n_start)
        g.sync()
        t0 = time.perf_counter()
        g.step(args.steps)
        g.sync()
        el_m = time.perf_counter() - t0
        (_, _), (ns_m, nf_m) = g.newton_stats()
        _, fb_m = g.guess_stats()
        # one step of each from the same state
        g.set("p", p_start)
        g.set_time(t_start, dt_start, n_start)
        g.step(1)
        a = g.active(g.get("p"))
        g.set_inversion_guess(None)
        g.set("p", p_start)
        g.set_time(t_start, dt_start, n_start)
        g.step(1)
        b = g.active(g.get("p"))
        rel = max(float(np.abs(a[..., v] - b[..., v]).max() / max(np.abs(b[..., v]).max(), 1e-300)) for v in range(8))
        mlp_out = {"value": zones * args.steps / el_m, "ms_per_step": el_m / args.steps * 1e3,
                   "speedup_over_plain": el / el_m,
                   "newton_iterations_per_zone_update": ns_m / (zones * 3.0 * args.steps),
                   "inversion_failures": nf_m, "fallbacks_to_plain_guess": fb_m,
                   "one_step_max_rel_difference_vs_plain": rel,
                   "training": {"plain_steps_before": args.mlp_plain_steps, "sampled_zones": args.mlp_samples,
                                "adam_steps": args.mlp_train_steps, "final_loss": losses[-1],
                                "seconds": t_train}}
    out = {
        "metric": "zone-updates/s, 3D FM torus 384x192x192 FP64, plain step" + (" vs MLP-guess inversion"
                                                                                if args.mlp else ""),
        "value": zones * args.steps / el,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Fishbone-Moncrief torus initial data)",
        "config": {"workload": "3D FM torus 384x192x192 MKS Kerr-Schild a=0.9375, plain step",
                   "recon": "WENO5BND", "flux": "HLL", "ct": "Toth", "rk": 3, "cour": cfg.cour,
                   "zones": zones, "sim_time_end": t_sim, "dt": dt_now},
        "steps_total": steps_done,
        "newton_iterations_per_zone_update": nstroke / (zones * 3.0 * args.steps),
        "inversion_failures": nfail,
        "divb_max_relative": divb,
        "accretion_sums": {"mdot_in": fa[0], "edot_in": fa[1], "mdot_out": fa[2], "edot_out": fa[3]},
        "mlp_guess": mlp_out,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", default="x".join(map(str, GRID)))
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--prof-steps", type=int, default=10, help="steps of the separate per-kernel timing run")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--workload", choices=["step", "inversion", "batch", "large"], default="step",
                    help="step: the 256x128x64 torus RK3 step (default); inversion: the batched-inversion "
                         "microbench, plain vs MLP-seeded; batch: 64 uneven 2-D tori and Bondi flows stepped together; "
                         "large: the 384x192x192 torus with failures, div B and accretion sums")
    ap.add_argument("--ngrids", type=int, default=64, help="batch: number of grids")
    ap.add_argument("--n", type=int, default=1 << 24, help="inversion microbench: zones")
    ap.add_argument("--train-n", type=int, default=1 << 18, help="inversion microbench: training pairs")
    ap.add_argument("--train-steps", type=int, default=20000, help="inversion microbench: Adam steps")
    ap.add_argument("--mlp", action="store_true", help="large: also time the step with the learned inversion guess")
    ap.add_argument("--mlp-plain-steps", type=int, default=200, help="large --mlp: plain steps before training")
    ap.add_argument("--mlp-samples", type=int, default=1 << 20, help="large --mlp: zones sampled for training")
    ap.add_argument("--mlp-train-steps", type=int, default=3000, help="large --mlp: Adam steps")
    ap.add_argument("--no-chosen", action="store_true", help="step: skip the learned-guess (chosen form) timing")
    ap.add_argument("--host-steps", type=int, default=3,
                    help="step: steps timed with the primitives set from and read back to host memory around each")
    ap.add_argument("--mlp-samples-step", type=int, default=1 << 18, help="step: zones sampled for training")
    ap.add_argument("--mlp-train-steps-step", type=int, default=2000, help="step: Adam steps")
    ap.add_argument("--calibrate-copy-mib", type=int, default=0,
                    help="before the run, one device-to-device copy of this many MiB (the reference dispatch "
                         "for scaling HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs the whole domain on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "inversion":
        return inversion_microbench(args)
    if args.workload == "batch":
        return batch_bench(args)
    if args.workload == "large":
        return large_bench(args)

    import harm_harm_amd as H
    from harm_harm_amd.grid import Grid

    if args.calibrate_copy_mib > 0:
        nb = args.calibrate_copy_mib << 20
        src, dst = H.DeviceArray(nb, "u1"), H.DeviceArray(nb, "u1")
        dst.copy_from_device(src)
        H.synchronize()
        src.free()
        dst.free()

    n1, n2, n3 = (int(x) for x in args.grid.split("x"))
    cfg = H.default_config(n1, n2, n3, recon="weno5", flux="hll", timeorder=3)
    g = Grid(cfg)
    g.init_torus(seed=1)
    dt0 = initial_dt(g)
    g.set_time(0.0, dt0)
    tiles = g.tune(reps=5)  # face-tile depths of directions 2 and 3 (results do not depend on them)
    g.step(args.warmup)
    g.sync()

    # timed region: whole steps replayed from the captured hipGraph
    g.sync()
    t0 = time.perf_counter()
    g.step(args.steps)
    g.sync()
    el = time.perf_counter() - t0
    t_sim, dt_now = g.time()
    # per-kernel times from a separate run with HIP events between the kernel
    # groups (launched one by one so the events can bracket them)
    g.profile(True)
    g.step(args.prof_steps)
    g.sync()
    prof = g.profile_read()
    g.profile(False)
    prof_steps = args.prof_steps
    # the same steps with the energy-file accumulators on (boundary sums every
    # substep, fix-up accounting and the source sample), and the conservation
    # identity of the density over them
    g.set_ener_tracking(True)
    g.step(1)  # capture the tracking step
    g.sync()
    e0 = g.ener()
    t0 = time.perf_counter()
    g.step(args.steps)
    g.sync()
    el_en = time.perf_counter() - t0
    e1 = g.ener()
    g.set_ener_tracking(False)
    pc = e1["pcum"]
    inflow = pc[1] - pc[0] + pc[3] - pc[2] + pc[5] - pc[4]
    resid = e1["U"] - e0["U"] - (inflow - e0["pcum"][1] + e0["pcum"][0] - e0["pcum"][3] + e0["pcum"][2]
                                 - e0["pcum"][5] + e0["pcum"][4]) - (e1["fladd"] - e0["fladd"])
    ener = {"ms_per_step": el_en / args.steps * 1e3, "overhead_vs_plain": el_en / el - 1.0,
            "density_identity_residual": abs(resid[0]) / abs(e1["U"][0]),
            "U_rho": e1["U"][0], "fladd_rho": e1["fladd"][0], "divbmax": e1["divbmax"], "divbavg": e1["divbavg"]}

    zones = n1 * n2 * n3
    # the same step for a caller whose state lives in host memory: the
    # primitives go through harm_set_array before and harm_get_array after
    # every step (HARM's own interleaved layout <-> the device's planes on the
    # host, pageable copies over PCIe).  Reported beside the resident rate,
    # never as the value.
    host = None
    if args.host_steps > 0:
        ph = g.get("p")
        t0 = time.perf_counter()
        for _ in range(args.host_steps):
            g.set("p", ph)
            g.step(1)
            ph = g.get("p")
        el_h = time.perf_counter() - t0
        host = {"steps": args.host_steps, "ms_per_step": el_h / args.host_steps * 1e3,
                "zone_updates_per_s": zones * args.host_steps / el_h,
                "bytes_each_way_per_step": int(ph.nbytes),
                "note": "primitives (storage zones, 8 doubles each) set before and read after every step"}
    value = zones * args.steps / el
    # dominant kernel group by total event time
    dom = max((k for k in prof if prof[k][1] > 0), key=lambda k: prof[k][0])
    ms_tot, nlaunch = prof[dom]
    avg_s = ms_tot / 1e3 / nlaunch
    nbytes = algorithmic_bytes(g, dom)
    achieved = nbytes / avg_s / 1e9 if nbytes else None
    kernels = {k: {"ms_total": round(v[0], 3), "launches": v[1],
                   "ms_avg": round(v[0] / v[1], 4) if v[1] else None,
                   "GBs_alg": round(algorithmic_bytes(g, k) / (v[0] / 1e3 / v[1]) / 1e9, 1)
                   if v[1] and algorithmic_bytes(g, k) else None} for k, v in prof.items() if v[1]}

    meas = measured_counters() if (n1, n2, n3) == GRID else None
    md = meas.get(dom) if meas else None
    fp64 = None
    if md and md["flops"] > 0:
        tf = md["flops"] / avg_s / 1e12
        fp64 = {"bound": "fp64", "kernel": dom, "achieved": tf, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tf / FP64_PEAK_TFLOPS, "flops_per_launch": md["flops"],
                "source": "rocprofv3 --pmc SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64, " + os.path.relpath(COUNTERS, ROOT)}
    # whole step against HBM.  Algorithmic bytes per zone-update of the plain
    # step, independent of how it is cut into kernels: per substep, recon +
    # flux per direction read p 64 and write F 64 (384), CT 144, update 512,
    # inversion 196 = 1236 B, three substeps.  Beside it the compulsory
    # traffic of this build's kernel groups (the WENO5BND face values and dP
    # make the split reconstruction move more than the fused unit would).
    alg_zu = 3 * (3 * 128 + 144 + 512 + 196)
    zps = zones * args.steps / el
    step_bytes = sum((algorithmic_bytes(g, k) or 0) * prof[k][1] / prof_steps for k in prof if prof[k][1])
    step_roof = {"bytes_per_zone_update_alg": alg_zu, "achieved": zps * alg_zu / 1e9, "peak": HBM_PEAK_GBS,
                 "unit": "GB/s", "frac": zps * alg_zu / 1e9 / HBM_PEAK_GBS,
                 "bytes_per_step_kernel_groups": step_bytes,
                 "kernel_groups_GBs": step_bytes / (el / args.steps) / 1e9}

    chosen = None if args.no_chosen else chosen_form(g, cfg, args, el)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(cfg, args.cpu_steps)

    out = {
        "metric": "zone-updates/s, 3D FM torus 256x128x64 FP64: plain vs fused+MLP-guess inversion",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Fishbone-Moncrief torus initial data, random-free geometry)",
        "config": {"workload": f"3D FM torus {n1}x{n2}x{n3} MKS Kerr-Schild a=0.9375, plain step",
                   "recon": "WENO5BND", "flux": "HLL", "ct": "Toth", "rk": 3, "cour": cfg.cour,
                   "form": "plain (multi-kernel step, Newton from previous primitives)",
                   "tile_depth_dir2_dir3": tiles[2:4],
                   "sim_time_end": t_sim, "dt": dt_now},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS if achieved else None,
                     "traffic": md["traffic"] if md else None,
                     "traffic_source": ("rocprofv3 --pmc FETCH_SIZE x2.0 + WRITE_SIZE x1.0 (1 GiB copy calibration), "
                                        + os.path.relpath(COUNTERS, ROOT)) if md else None,
                     "bytes_per_launch_alg": nbytes, "avg_launch_ms": avg_s * 1e3},
        "roofline_fp64": fp64,
        "step_roofline_hbm": step_roof,
        "kernels": kernels,
        "ener_tracking": ener,
        "host_buffers": host,
        "chosen_form": chosen,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def chosen_form(g, cfg, args, el_plain):
    """The step with the learned inversion guess (the chosen form's MLP
    half; the fused persistent step kernel is not built): the network is
    trained on the GPU from zones of the current state, then the same number
    of steps is timed from that state with every inversion seeded by it, and
    one step of each form from the same state is compared."""
    import numpy as np
    import harm_harm_amd as H
    from harm_harm_amd.mlp import MLP
    t0 = time.time()
    n = args.mlp_samples_step
    dU, dgeo, dp = g.sample_zones(n, seed=11)
    X, T, ok = MLP.pairs(dU, dgeo, dp, n, gammafail=cfg.gammafail)
    order = np.nonzero(ok.numpy())[0].astype(np.int32)
    np.random.default_rng(0).shuffle(order)
    m = MLP(seed=1)
    m.fit_norm(X, T, ok, n)
    losses = m.train(X, T, n, H.DeviceArray.from_numpy(order), args.mlp_train_steps_step, lr=3e-3, lr_end=3e-5,
                     chunk=max(1, args.mlp_train_steps_step // 4))
    t_train = time.time() - t0
    p0, (t_s, dt_s), n0 = g.get("p"), g.time(), g.nstep()
    # one step of each form from the same state
    g.step(1)
    b = g.active(g.get("p"))
    (ns_plain1, nf_plain1), _ = g.newton_stats()
    g.set("p", p0)
    g.set_time(t_s, dt_s, n0)
    g.set_inversion_guess(m)
    g.step(1)
    a = g.active(g.get("p"))
    (ns_seed1, nf_seed1), _ = g.newton_stats()
    rel = max(float(np.abs(a[..., v] - b[..., v]).max() / max(np.abs(b[..., v]).max(), 1e-300)) for v in range(8))
    # timed steps with the guess (graph captured by the step above)
    g.set("p", p0)
    g.set_time(t_s, dt_s, n0)
    g.sync()
    t1 = time.perf_counter()
    g.step(args.steps)
    g.sync()
    el = time.perf_counter() - t1
    (_, _), (ns, nf) = g.newton_stats()
    _, fb = g.guess_stats()
    g.set_inversion_guess(None)
    zones = cfg.n1 * cfg.n2 * cfg.n3
    return {"form": "plain multi-kernel step with the MLP inversion guess (8-64-64-2 BF16 on MFMA); "
                    "the fused persistent step kernel is not built",
            "value": zones * args.steps / el, "unit": "zone-updates/s", "ms_per_step": el / args.steps * 1e3,
            "speedup_over_plain": el_plain / el,
            "newton_iterations_per_zone_substep": ns / (zones * 3.0 * args.steps),
            "newton_iterations_per_zone_substep_plain_one_step": ns_plain1 / (zones * 3.0),
            "newton_iterations_per_zone_substep_seeded_one_step": ns_seed1 / (zones * 3.0),
            "inversion_failures": nf, "fallbacks_to_plain_guess": fb,
            "one_step_max_rel_difference_vs_plain": rel,
            "one_step_failures_plain_vs_seeded": [nf_plain1, nf_seed1],
            "training": {"sampled_zones": n, "adam_steps": args.mlp_train_steps_step, "final_loss": losses[-1],
                         "seconds": t_train}}


if __name__ == "__main__":
    main()
