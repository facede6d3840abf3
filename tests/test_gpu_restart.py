"""Restart files (harm_write_restart / harm_read_restart, restart.c:163-285)
through the C ABI.

* MIXEDOUTPUT (text header, raw doubles): a grid read back from the file
  holds the writer's primitives (ghost zones refilled by the boundary
  conditions), clock and energy accumulators bit for bit, and stepping it
  gives the same bits as the writer stepping on; this rests on the step
  depending only on the primitives and the clock, which the oracle shows
  (a fresh oracle state continues a run identically).
* TEXTOUTPUT: "%21.15g", as the reference: the values come back to 15
  significant digits, not exactly.
* A file of another grid or configuration is refused, naming the field.
"""
import numpy as np
import pytest

import harm_harm_amd as H
from harm_harm_amd.grid import Grid

pytestmark = pytest.mark.gpu

CASES = {
    "torus_3d_weno5_hll_rk3": lambda: H.default_config(24, 16, 8, recon="weno5", flux="hll", timeorder=3),
    "torus_2d_para_laxf_rk2": lambda: H.default_config(32, 32, 1, recon="para", flux="laxf", timeorder=2),
    "bondi_para_laxf_rk2": lambda: H.bondi_config(48, 32, recon="para", flux="laxf", timeorder=2),
}


def _start(cfg):
    g = Grid(cfg)
    if cfg.problem == 1:  # HARM_PROBLEM_BONDI
        g.init_bondi()
    else:
        g.init_torus()
    g.set_time(0.0, 0.5 * min(g.fluxcalc("p")))
    g.set_ener_tracking(True)
    return g


def _state(g):
    e = g.ener()
    return g.get("p"), g.pflag(), g.time(), g.nstep(), e


def _same_ener(a, b):
    for k in ("U", "pdot", "pcum", "fladdterms", "sourceadd"):
        assert np.array_equal(a[k], b[k]), k
    assert a["t"] == b["t"] and a["nstep"] == b["nstep"]


@pytest.mark.parametrize("case", list(CASES))
def test_binary_restart_continues_bit_for_bit(case, tmp_path):
    cfg = CASES[case]()
    a = _start(cfg)
    a.step(5)  # nstep 5: the next source sample (nstep 8) needs the restored clock
    path = tmp_path / "rdump-0"
    a.write_restart(path, binary=True, tf=2.5)
    p_w, _, time_w, n_w, e_w = _state(a)
    ucum_w = a.active(a.get("ucum"))
    b = Grid(cfg)
    b.set_ener_tracking(True)
    assert b.read_restart(path) == 2.5
    p_r, _, time_r, n_r, e_r = _state(b)
    assert np.array_equal(p_w, p_r)  # ghost zones included
    assert np.array_equal(ucum_w, b.active(b.get("ucum")))
    assert time_w == time_r and n_w == n_r == 5
    _same_ener(e_w, e_r)
    a.step(4)
    b.step(4)
    pa, fa, ta, na, ea = _state(a)
    pb, fb, tb, nb, eb = _state(b)
    print(case, "t", ta, "sourceadd[UU]", ea["sourceadd"][1])
    assert np.array_equal(pa, pb) and np.array_equal(fa, fb) and ta == tb and na == nb
    _same_ener(ea, eb)
    assert np.any(ea["sourceadd"] != 0.0)


def test_text_restart_fifteen_digits(tmp_path):
    cfg = CASES["torus_3d_weno5_hll_rk3"]()
    a = _start(cfg)
    a.step(2)
    path = tmp_path / "rdump-1"
    a.write_restart(path, binary=False)
    lines = open(path).read().splitlines()
    nz = cfg.n1 * cfg.n2 * cfg.n3
    assert len(lines) == 1 + nz
    assert all(len(l.split()) == 16 for l in lines[1:])
    head = lines[0].split()
    assert [int(x) for x in head[:3]] == [cfg.n1, cfg.n2, cfg.n3] and int(head[5]) == 2
    assert int(head[13]) == 10 and int(head[16]) == 3 and int(head[17]) == 1  # WENO5, RK3, HLL codes
    b = Grid(cfg)
    b.read_restart(path)
    pa, pb = a.active(a.get("p")), b.active(b.get("p"))
    err = np.abs(pa - pb) / np.maximum(np.abs(pa), 1e-300)
    print("text restart max relative error", err.max())
    assert err.max() <= 1e-14 and np.any(pa != pb)
    assert a.time() == b.time() and a.nstep() == b.nstep()  # header doubles read back exactly


@pytest.mark.parametrize("other,field", [
    (lambda: H.default_config(24, 16, 16, recon="weno5", flux="hll", timeorder=3), "N3"),
    (lambda: H.default_config(24, 16, 8, recon="para", flux="hll", timeorder=3), "lim"),
    (lambda: H.default_config(24, 16, 8, recon="weno5", flux="laxf", timeorder=3), "fluxmethod"),
    (lambda: H.default_config(24, 16, 8, recon="weno5", flux="hll", timeorder=2), "TIMEORDER"),
])
def test_mismatched_restart_is_refused(other, field, tmp_path):
    a = _start(CASES["torus_3d_weno5_hll_rk3"]())
    path = tmp_path / "rdump-0"
    a.write_restart(path)
    b = Grid(other())
    with pytest.raises(Exception, match=field):
        b.read_restart(path)
