"""harm_dump_columns_noalloc: the main dump's columns with no allocation,
through the C ABI, against harm_dump_columns on two grids in a row (the
second smaller), bit for bit; a grid above the static buffer's size is
refused with an error that names the limit."""
import ctypes

import numpy as np
import pytest

import harm_harm_amd as H
from harm_harm_amd.grid import Grid

pytestmark = pytest.mark.gpu


def _rows(g, fn):
    n = g.cfg.n1 * g.cfg.n2 * g.cfg.n3
    out = np.full((n, Grid.DUMP_NCOL), np.nan)
    st = fn(g.h, 0, out.ctypes.data)
    return st, out


def _noalloc():
    L = H.lib()
    f = L.harm_dump_columns_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    f.restype = ctypes.c_int
    return f


def test_noalloc_dump_columns_equal_allocating_ones():
    f = _noalloc()
    L = H.lib()
    for n, recon, flux, to in (((24, 16, 8), "weno5", "hll", 3), ((12, 10, 8), "para", "laxf", 2)):
        cfg = H.default_config(*n, recon=recon, flux=flux, timeorder=to)
        g = Grid(cfg)
        g.init_torus()
        g.set_time(0.0, 0.5 * min(g.fluxcalc("p")))
        g.step(2)
        st0, ref = _rows(g, L.harm_dump_columns)
        st1, got = _rows(g, f)
        H.check(st0, "harm_dump_columns")
        H.check(st1, "harm_dump_columns_noalloc")
        assert np.array_equal(got, ref), n
        g.close()


def test_noalloc_dump_columns_refuse_grids_above_the_limit():
    f = _noalloc()
    cfg = H.default_config(64, 64, 32, recon="para", flux="laxf", timeorder=2)  # 131072 zones
    g = Grid(cfg)
    st = f(g.h, 0, np.zeros((1, 1)).ctypes.data)
    assert st != 0
    msg = H.lib().harm_last_error().decode()
    assert "65536" in msg, msg
    g.close()
