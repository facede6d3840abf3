"""ctypes binding of the oracle (oracle/liboracle.so): the CPU restatement the
HIP path is checked against.  Test infrastructure only."""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_L = ctypes.c_int64

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True)
        L = ctypes.CDLL(ORACLE_SO)
        L.o_invert_batch.argtypes = [_P, _P, _P, _P, _P, _L, _D, _D, _I]
        L.o_invert_batch.restype = None
        L.o_p2u_batch.argtypes = [_P, _P, _P, _L, _D]
        L.o_p2u_batch.restype = None
        L.o_flux_batch.argtypes = [_P, _P, _P, _L, _I, _D]
        L.o_flux_batch.restype = None
        L.o_vchar_batch.argtypes = [_P, _P, _P, _P, _L, _I, _D]
        L.o_vchar_batch.restype = None
        L.o_source_batch.argtypes = [_P, _P, _P, _P, _L, _D]
        L.o_source_batch.restype = None
        _lib = L
    return _lib


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def invert_batch(U, geom, guess, gam=4.0 / 3.0, gammafail=2500.0, nthreads=8):
    U = _c(U)
    geom = _c(geom)
    p = _c(guess).copy()
    n = U.shape[1]
    pflag = np.zeros(n, dtype=np.int32)
    niter = np.zeros(n, dtype=np.int32)
    lib().o_invert_batch(U.ctypes.data, geom.ctypes.data, p.ctypes.data, pflag.ctypes.data, niter.ctypes.data, n,
                         gam, gammafail, nthreads)
    return p, pflag, niter


def p2u_batch(p, geom, gam=4.0 / 3.0):
    p = _c(p)
    geom = _c(geom)
    n = p.shape[1]
    U = np.empty_like(p)
    lib().o_p2u_batch(p.ctypes.data, geom.ctypes.data, U.ctypes.data, n, gam)
    return U


def flux_batch(p, geom, dir, gam=4.0 / 3.0):
    p = _c(p)
    geom = _c(geom)
    n = p.shape[1]
    F = np.empty_like(p)
    lib().o_flux_batch(p.ctypes.data, geom.ctypes.data, F.ctypes.data, n, dir, gam)
    return F


def vchar_batch(p, geom, dir, gam=4.0 / 3.0):
    p = _c(p)
    geom = _c(geom)
    n = p.shape[1]
    vmax = np.empty(n)
    vmin = np.empty(n)
    lib().o_vchar_batch(p.ctypes.data, geom.ctypes.data, vmax.ctypes.data, vmin.ctypes.data, n, dir, gam)
    return vmax, vmin


def source_batch(p, geom, conn, gam=4.0 / 3.0):
    p = _c(p)
    geom = _c(geom)
    conn = _c(conn)
    n = p.shape[1]
    dU = np.empty_like(p)
    lib().o_source_batch(p.ctypes.data, geom.ctypes.data, conn.ctypes.data, dU.ctypes.data, n, gam)
    return dU
