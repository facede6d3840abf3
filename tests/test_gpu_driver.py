"""The host driver (harm_harm_amd.driver.run): HARM's end-of-run rule and
output cadence around the device step (main.c, step_ch.c:200-250,
diag.c:502-700), on a small 2-D torus."""
import numpy as np
import pytest

import harm_harm_amd as H
from harm_harm_amd import driver
from harm_harm_amd.grid import Grid

pytestmark = pytest.mark.gpu


def test_run_to_tf_with_outputs(tmp_path):
    cfg = H.default_config(32, 32, 1, recon="para", flux="laxf", timeorder=2)
    g = Grid(cfg)
    g.init_torus()
    ndt = g.fluxcalc("p")
    dt0 = 0.5 / sum(1.0 / x for x in ndt)
    g.set_time(0.0, dt0)
    tf = 10.3 * dt0
    out = driver.run(g, tf, DTdump=3.0 * dt0, DTener=dt0, outdir=str(tmp_path))
    t, _ = g.time()
    print(out)
    assert t == pytest.approx(tf, rel=1e-14)  # the last step is clipped onto tf
    dumps = sorted((tmp_path / "dumps").glob("dump*"))
    assert len(dumps) == out["dumps"] >= 4  # t = 0, each 3 dt0, and the final time
    assert (tmp_path / "dumps" / "gdump").exists()
    lines = (tmp_path / "ener.out").read_text().splitlines()
    assert len(lines) == out["ener_lines"]
    ts = np.array([float(x.split()[0]) for x in lines])
    assert ts[0] == 0.0 and np.all(np.diff(ts) > 0) and ts[-1] == pytest.approx(tf, rel=1e-14)
    # dump headers carry increasing times, the last one tf
    th = [float(open(d).readline().split()[0]) for d in dumps]
    assert th == sorted(th) and th[-1] == pytest.approx(tf, rel=1e-14)


def test_run_writes_restart_files_that_resume(tmp_path):
    """DTr = 4 steps over 11 steps of a fresh run: rdump-0 at step 0, rdump-1
    at 4, rdump-0 at 8, rdump-1 again at 11 (the final call), and
    rdump--NNNN beside each main dump; a grid read from the last one stands
    where the run ended."""
    cfg = H.default_config(32, 32, 1, recon="para", flux="laxf", timeorder=2)
    g = Grid(cfg)
    g.init_torus()
    g.set_time(0.0, 0.5 * min(g.fluxcalc("p")))
    out = driver.run(g, 1e9, DTdump=None, DTener=None, outdir=str(tmp_path), max_steps=11, DTr=4)
    assert out["steps"] == 11 and out["rdumps"] == 4
    d = tmp_path / "dumps"
    assert (d / "rdump-0").exists() and (d / "rdump-1").exists()
    assert int(open(d / "rdump-0", "rb").readline().split()[5]) == 8
    assert int(open(d / "rdump-1", "rb").readline().split()[5]) == 11
    b = Grid(cfg)
    assert b.read_restart(d / "rdump-1") == 1e9
    assert np.array_equal(b.get("p"), g.get("p")) and b.time() == g.time()
    # a main dump at t = 0 and after each of two steps (the final call repeats
    # no time), each with its numbered restart file
    g = Grid(cfg)
    g.init_torus()
    g.set_time(0.0, 0.5 * min(g.fluxcalc("p")))
    driver.run(g, 1e9, DTdump=1e-9, outdir=str(tmp_path / "m"), gdump=False, max_steps=2, DTr=100)
    names = sorted(x.name for x in (tmp_path / "m" / "dumps").glob("rdump--*"))
    assert names == ["rdump--0000", "rdump--0001", "rdump--0002"]


def test_resumed_run_continues_the_outputs(tmp_path):
    """A run of 6 steps, read back from its last restart file and resumed for
    5 more, leaves the same final state and energy sums as one run of 11
    steps, and the same ener.out lines apart from HARM's restart cadence: the
    restored sums keep accumulating, ener.out is appended to (its line at the
    resume time repeated, as HARM's initial call writes one, and the next
    period's line skipped), main dumps are numbered on."""
    cfg = H.default_config(32, 32, 1, recon="para", flux="laxf", timeorder=2)

    def fresh():
        g = Grid(cfg)
        g.init_torus()
        g.set_time(0.0, 0.5 * min(g.fluxcalc("p")))
        return g

    dt0 = fresh().time()[1]
    kw = dict(DTdump=2.5 * dt0, DTener=dt0, gdump=False, DTr=3)
    one = fresh()
    driver.run(one, 1e9, outdir=str(tmp_path / "one"), max_steps=11, **kw)
    a = fresh()
    out_a = driver.run(a, 1e9, outdir=str(tmp_path / "two"), max_steps=6, **kw)
    d = tmp_path / "two" / "dumps"
    last = max((d / "rdump-0", d / "rdump-1"), key=lambda f: int(open(f, "rb").readline().split()[5]))
    b = Grid(cfg)
    b.set_ener_tracking(True)
    b.read_restart(last)
    assert b.nstep() == 6
    out_b = driver.run(b, 1e9, outdir=str(tmp_path / "two"), max_steps=5, resume=True, **kw)
    assert np.array_equal(b.get("p"), one.get("p")) and b.time() == one.time()
    e1, e2 = one.ener(), b.ener()
    for k in ("pcum", "fladdterms", "sourceadd"):
        assert np.array_equal(e1[k], e2[k]), k
    l1 = (tmp_path / "one" / "ener.out").read_text().splitlines()
    l2 = (tmp_path / "two" / "ener.out").read_text().splitlines()
    # the resumed run's initial call repeats the line at the resume time
    t2 = [float(x.split()[0]) for x in l2]
    dup = [i for i in range(1, len(t2)) if t2[i] == t2[i - 1]]
    assert len(dup) == 1
    t_resume = t2[dup[0]]
    del l2[dup[0]]
    # HARM's post_dump moves the next energy time on by one period after that
    # initial line (diag.c:531-533: dumpc = 1 since t < tdump), so the
    # resumed run skips the first multiple of DTener after the resume time
    # that the continuous run wrote; every other line is the same, bit for bit
    one_by_t = {float(x.split()[0]): x for x in l1}
    for x in l2:
        assert one_by_t.get(float(x.split()[0])) == x, x.split()[:2]
    missing = sorted(set(one_by_t) - {float(x.split()[0]) for x in l2})
    assert len(missing) <= 1
    for tm in missing:
        assert t_resume < tm <= t_resume + 2.0 * kw["DTener"], (tm, t_resume)
    # the main dumps of both parts, numbered on without overwriting, in time order
    n2 = sorted(x.name for x in d.glob("dump[0-9]*"))
    assert n2 == ["dump%04d" % i for i in range(out_a["dumps"] + out_b["dumps"])] and out_b["dumps"] > 0
    th = [float(open(d / n).readline().split()[0]) for n in n2]
    assert all(x < y for x, y in zip(th, th[1:]))
