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


def recon_cell(recon, y5):
    """(left, right) edge values of the cell y5[2] from y5 = y[-2..2]."""
    y = _c(y5)
    l, r = _D(), _D()
    L = lib()
    L.o_recon_cell.argtypes = [_I, _P, _P, _P]
    L.o_recon_cell.restype = None
    L.o_recon_cell(int(recon), y.ctypes.data, ctypes.byref(l), ctypes.byref(r))
    return l.value, r.value


def weno5bnd_line(y, dP=None, ng=4):
    """WENO5BND left/right face values of cells ng-1 .. n-ng of the storage
    line y (n values); dP is the line's pressure-jump measure (zeros if None).
    Returns (yl, yr) over the whole line (zeros outside those cells)."""
    y = _c(y)
    n = y.size
    dP = np.zeros(n) if dP is None else _c(dP)
    yl, yr = np.zeros(n), np.zeros(n)
    L = lib()
    L.o_weno5bnd_line.argtypes = [_P, _I, _I, _P, _P, _P]
    L.o_weno5bnd_line.restype = None
    L.o_weno5bnd_line(y.ctypes.data, n, int(ng), dP.ctypes.data, yl.ctypes.data, yr.ctypes.data)
    return yl, yr


def weno5bnd_line_lof(y, dP, ng=4):
    """weno5bnd_line plus each reconstructed cell's final lower-order fraction
    and the fraction before the dP reduction"""
    y = _c(y)
    n = y.size
    dP = _c(dP)
    yl, yr, lof, lof0 = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    L = lib()
    L.o_weno5bnd_line_lof.argtypes = [_P, _I, _I, _P, _P, _P, _P, _P]
    L.o_weno5bnd_line_lof.restype = None
    L.o_weno5bnd_line_lof(y.ctypes.data, n, int(ng), dP.ctypes.data, yl.ctypes.data, yr.ctypes.data,
                          lof.ctypes.data, lof0.ctypes.data)
    return yl, yr, lof, lof0


def smono_indicator(y5, eps=500.0 * np.finfo(float).eps):
    """SMONO indicator of the five values y5 = y[-2..2]."""
    y = _c(y5)
    L = lib()
    L.o_smono_indicator.argtypes = [_P, _D]
    L.o_smono_indicator.restype = ctypes.c_double
    return L.o_smono_indicator(y.ctypes.data + 2 * 8, eps)


def dp_line(cells, geomrec, dir, gam=4.0 / 3.0):
    """dP line of the primitives cells[n, 8] at one 27-field geometry record."""
    L = lib()
    L.o_dp_cell_rec.argtypes = [_P, _P, _D, _I, _P]
    L.o_dp_cell_rec.restype = None
    L.o_dp_line.argtypes = [_P, _I, _P]
    L.o_dp_line.restype = None
    cells = _c(cells)
    rec = _c(geomrec)
    n = cells.shape[0]
    st = np.zeros((n, 4))
    for c in range(n):
        row = np.ascontiguousarray(cells[c])
        out = np.zeros(4)
        L.o_dp_cell_rec(row.ctypes.data, rec.ctypes.data, gam, int(dir), out.ctypes.data)
        st[c] = out
    dP = np.zeros(n)
    L.o_dp_line(st.ctypes.data, n, dP.ctypes.data)
    return dP, st


def face_flux(flux, dir, rec, p_l, p_r, gam=4.0 / 3.0, gammamax=25.0):
    """One face's Riemann flux (0 LAX-F, 1 HLL) at a 27-field geometry record."""
    L = lib()
    L.o_face_flux.argtypes = [_I, _I, _D, _D, _P, _P, _P, _P, _P]
    L.o_face_flux.restype = None
    rec = _c(rec)
    pl = _c(p_l).copy()
    pr = _c(p_r).copy()
    F = np.empty(8)
    ct = _D()
    L.o_face_flux(int(flux), int(dir), gam, gammamax, rec.ctypes.data, pl.ctypes.data, pr.ctypes.data,
                  F.ctypes.data, ctypes.byref(ct))
    return F, ct.value


# ---------------------------------------------------------------------------
# grid-level oracle (o_grid.c)
# ---------------------------------------------------------------------------
class OGridStruct(ctypes.Structure):
    _fields_ = [
        ("n1", ctypes.c_int), ("n2", ctypes.c_int), ("n3", ctypes.c_int), ("ng", ctypes.c_int),
        ("recon", ctypes.c_int), ("flux", ctypes.c_int), ("timeorder", ctypes.c_int), ("nthreads", ctypes.c_int),
        ("gam", ctypes.c_double), ("cour", ctypes.c_double), ("safe", ctypes.c_double),
        ("gammamax", ctypes.c_double), ("gammafail", ctypes.c_double),
        ("dx", ctypes.c_double * 4), ("startx", ctypes.c_double * 4),
        ("rhomin", ctypes.c_double), ("uumin", ctypes.c_double),
        ("geom", ctypes.c_void_p * 4), ("gdet_corn3", ctypes.c_void_p), ("conn", ctypes.c_void_p),
        ("rcent", ctypes.c_void_p), ("problem", ctypes.c_int), ("R0", ctypes.c_double),
        ("ener", ctypes.c_void_p),
        ("hot2cold", ctypes.c_int), ("checkinv", ctypes.c_int), ("resetregion", ctypes.c_int),
        ("failcount", ctypes.c_void_p),
    ]


# energy-file accumulator slots of o_grid.h
EN_PDOT, EN_PCUM, EN_FLTERMS, EN_SRCADD, EN_DT4, EN_WILLBE, EN_NSTEP, EN_N = 0, 48, 96, 192, 200, 201, 202, 256


def _grid_protos(L):
    P = ctypes.POINTER(OGridStruct)
    L.o_bound.argtypes = [P, _P]
    L.o_fluxcalc.argtypes = [P, _P, _P, _P, _P, _P, _I]
    L.o_update.argtypes = [P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _D, _I, _P]
    L.o_invert.argtypes = [P, _P, _P, _P, _P, _I, _P]
    L.o_post_fixup.argtypes = [P, _P, _P, _I]
    L.o_bound_pflag.argtypes = [P, _P]
    L.o_bound_pflag.restype = None
    L.o_step.argtypes = [P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]
    L.o_ener_totals.argtypes = [P, _P, _P]
    L.o_ener_totals.restype = None
    L.o_dump_row.argtypes = [P, _P, _P, _I, _I, _I, _P]
    L.o_dump_row.restype = None
    L.o_current_precalc.argtypes = [P, _P, _P]
    L.o_current_precalc.restype = None
    L.o_current_calc.argtypes = [P, _P, _I, _D, _D, _P]
    L.o_current_calc.restype = None
    L.o_rk_tables.argtypes = [_I, _P, _P]
    L.o_rk_tables.restype = _I
    for n in ("o_bound", "o_fluxcalc", "o_update", "o_invert", "o_post_fixup", "o_step"):
        getattr(L, n).restype = None


class OGrid:
    """The oracle's grid, built from the library's host geometry tables so that
    both sides see identical metric and connection values."""

    def __init__(self, cfg, nthreads=8):
        import harm_harm_amd as H
        L = lib()
        if not hasattr(L, "_grid_ok"):
            _grid_protos(L)
            L._grid_ok = True
        self.cfg = cfg
        self.ng1 = cfg.ng if cfg.n1 > 1 else 0
        self.ng2 = cfg.ng if cfg.n2 > 1 else 0
        self.ng3 = cfg.ng if cfg.n3 > 1 else 0
        self.ns1, self.ns2, self.ns3 = cfg.n1 + 2 * self.ng1, cfg.n2 + 2 * self.ng2, cfg.n3 + 2 * self.ng3
        np_ = self.ns1 * self.ns2
        self.geom = np.empty((4, self.ns2, self.ns1, 27))
        self.conn = np.empty((self.ns2, self.ns1, 40))
        self.gc3 = np.empty((self.ns2, self.ns1))
        self.rc = np.empty(self.ns1)
        sx = (ctypes.c_double * 4)()
        dx = (ctypes.c_double * 4)()
        H.check(H.lib().harm_host_tables(ctypes.byref(cfg), self.geom.ctypes.data, self.conn.ctypes.data,
                                         self.gc3.ctypes.data, self.rc.ctypes.data, sx, dx), "host_tables")
        g = OGridStruct()
        g.n1, g.n2, g.n3, g.ng = cfg.n1, cfg.n2, cfg.n3, cfg.ng
        g.recon, g.flux, g.timeorder, g.nthreads = cfg.recon, cfg.flux, cfg.timeorder, nthreads
        g.gam, g.cour, g.safe, g.gammamax, g.gammafail = cfg.gam, cfg.cour, cfg.safe, cfg.gammamax, cfg.gammafail
        for a in range(4):
            g.dx[a] = dx[a]
            g.startx[a] = sx[a]
        g.rhomin, g.uumin = cfg.rhomin, cfg.uumin
        for loc in range(4):
            g.geom[loc] = self.geom[loc].ctypes.data
        g.gdet_corn3 = self.gc3.ctypes.data
        g.conn = self.conn.ctypes.data
        g.rcent = self.rc.ctypes.data
        g.problem, g.R0 = cfg.problem, cfg.R0
        g.ener = None
        g.hot2cold, g.checkinv, g.resetregion = cfg.hot2cold, cfg.checkinv, cfg.resetregion
        self.failcount = np.zeros(2, dtype=np.int64)
        g.failcount = self.failcount.ctypes.data
        self.en = None
        self.g = g
        self.dx = [dx[a] for a in range(4)]

    @property
    def shape(self):
        return (self.ns3, self.ns2, self.ns1, 8)

    def zeros(self):
        return np.zeros(self.shape)

    def init_torus(self, rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=1):
        import harm_harm_amd as H
        p = self.zeros()
        H.check(H.lib().harm_host_init_torus(ctypes.byref(self.cfg), rin, rmax, kappa, beta, randfact, seed,
                                             p.ctypes.data), "host_init_torus")
        self.bound(p)
        return p

    def init_bondi(self):
        """The Bondi flow on the active zones (the library's host generator,
        pinned against the reference's tables), then bounded by the oracle."""
        import harm_harm_amd as H
        p = self.zeros()
        H.check(H.lib().harm_host_init_bondi(ctypes.byref(self.cfg), p.ctypes.data), "host_init_bondi")
        self.bound(p)
        return p

    def bound(self, p):
        lib().o_bound(ctypes.byref(self.g), p.ctypes.data)

    def fluxcalc(self, pb, F1, F2, F3, do_ct=True):
        ndt = (ctypes.c_double * 4)()
        lib().o_fluxcalc(ctypes.byref(self.g), pb.ctypes.data, F1.ctypes.data, F2.ctypes.data, F3.ctypes.data, ndt,
                         int(do_ct))
        return [ndt[1], ndt[2], ndt[3]]

    def update(self, pi, pb, F1, F2, F3, ui, uf, ucum, CUf, Cunew, dt, timeorder):
        c1 = (ctypes.c_double * 4)(*CUf)
        c2 = (ctypes.c_double * 4)(*Cunew)
        acc = ctypes.c_double()
        lib().o_update(ctypes.byref(self.g), pi.ctypes.data, pb.ctypes.data, F1.ctypes.data, F2.ctypes.data,
                       F3.ctypes.data, ui.ctypes.data, uf.ctypes.data, ucum.ctypes.data, c1, c2, dt, timeorder,
                       ctypes.byref(acc))
        return acc.value

    def invert(self, pb, U, pf, pflag, finalstep):
        ns = ctypes.c_int64()
        lib().o_invert(ctypes.byref(self.g), pb.ctypes.data, U.ctypes.data, pf.ctypes.data, pflag.ctypes.data,
                       int(finalstep), ctypes.byref(ns))
        return ns.value

    def post_fixup(self, pf, pflag, finalstep=True):
        """bounds the flags (one ghost layer), averages the failed zones of pf
        and marks the accounted ones -1, in place"""
        lib().o_post_fixup(ctypes.byref(self.g), pf.ctypes.data, pflag.ctypes.data, int(finalstep))

    def bound_pflag(self, pflag):
        lib().o_bound_pflag(ctypes.byref(self.g), pflag.ctypes.data)

    def new_pflag(self):
        """a flag array over the storage box as a run starts: NANPFLAG (-100)
        everywhere (set_arrays_multidimen.c:404)"""
        return np.full(self.shape[:3], -100, dtype=np.int32)

    def rk_tables(self):
        CUf = np.zeros((4, 4))
        Cunew = np.zeros((4, 4))
        n = lib().o_rk_tables(self.cfg.timeorder, CUf.ctypes.data, Cunew.ctypes.data)
        return n, CUf, Cunew

    def step(self, state, dt, fluxacc=None):
        """state: dict of arrays p, pk0, pk1, ui, uf, ucum, F1, F2, F3, pflag; returns new dt."""
        d = ctypes.c_double(dt)
        fa = fluxacc if fluxacc is not None else None
        lib().o_step(ctypes.byref(self.g), *[state[k].ctypes.data for k in
                                             ("p", "pk0", "pk1", "ui", "uf", "ucum", "F1", "F2", "F3", "pflag")],
                     ctypes.byref(d), fa.ctypes.data if fa is not None else None)
        return d.value

    def track_ener(self, nstep=0):
        """Make step() add to the energy-file accumulators from now on (zeroed;
        `nstep`: the step count of the next step, as the device clock)."""
        self.en = np.zeros(EN_N)
        self.en[EN_NSTEP] = nstep
        self.g.ener = self.en.ctypes.data

    def ener(self, p):
        """dict like the device record: U, pdot, pcum, fladd, sourceadd, divbmax, divbavg"""
        tot = np.zeros(10)
        lib().o_ener_totals(ctypes.byref(self.g), p.ctypes.data, tot.ctypes.data)
        en = self.en if self.en is not None else np.zeros(EN_N)
        nz = self.cfg.n1 * self.cfg.n2 * self.cfg.n3
        terms = en[EN_FLTERMS:EN_FLTERMS + 96].reshape(12, 8).copy()
        return {"U": tot[:8].copy(), "pdot": en[EN_PDOT:EN_PDOT + 48].reshape(6, 8).copy(),
                "pcum": en[EN_PCUM:EN_PCUM + 48].reshape(6, 8).copy(), "fladd": terms.sum(0), "fladdterms": terms,
                "sourceadd": en[EN_SRCADD:EN_SRCADD + 8].copy(), "divbmax": tot[8], "divbavg": tot[9] / nz}

    def current_slots(self):
        """zeroed Faraday samples of the current calculation: (ns3, ns2, ns1, 6, 3)"""
        return np.zeros((self.ns3, self.ns2, self.ns1, 6, 3))

    def current_precalc(self, p, cf):
        lib().o_current_precalc(ctypes.byref(self.g), p.ctypes.data, cf.ctypes.data)

    def current_calc(self, cf, timeterms, dtl, dtr):
        """J^mu of the active zones, (n3, n2, n1, 4)"""
        c = self.cfg
        j = np.zeros((c.n3, c.n2, c.n1, 4))
        lib().o_current_calc(ctypes.byref(self.g), cf.ctypes.data, int(timeterms), float(dtl), float(dtr),
                             j.ctypes.data)
        return j

    def dump_rows(self, p, ucum):
        """dump_content's 55 physics columns (o_dump_row) of every active zone,
        rows in the dump's order (i fastest, then j, then k)"""
        c = self.cfg
        out = np.empty((c.n1 * c.n2 * c.n3, 55))
        row = np.empty(55)
        t = 0
        for k in range(c.n3):
            for j in range(c.n2):
                for i in range(c.n1):
                    lib().o_dump_row(ctypes.byref(self.g), p.ctypes.data, ucum.ctypes.data, i, j, k, row.ctypes.data)
                    out[t] = row
                    t += 1
        return out

    def new_state(self, p):
        st = {k: self.zeros() for k in ("pk0", "pk1", "ui", "uf", "ucum", "F1", "F2", "F3")}
        st["p"] = p
        st["pflag"] = self.new_pflag()
        return st
