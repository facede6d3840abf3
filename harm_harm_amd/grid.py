"""Host-side mirror of HARM's grid driver over the C ABI.

Grid owns one harm_handle (the whole domain resident on one GPU).  Arrays cross
the boundary in HARM's own layout, numpy shape (ns3, ns2, ns1, 8) for
p[k][j][i][pl] over the ghosted storage box (global.storage.h:320-356).
"""
import ctypes

import numpy as np

from . import lib, check, HarmConfig, NewtonStats, HarmEner, ARR, G_NFIELD, NCONN, default_config


# NUMDISSVERSIONS (global.nondepnmemonics.h:417): dissipation columns, -1.0
# placeholders when dissipation is not tracked (DODISS 0, dump_ener.c:405)
NUMDISSVERSIONS = 18


def ener_line(e):
    """The energy-file line of record `e` (Grid.ener) in the reference's
    column order and format (dump_ener.c:391-407, 530): t realnstep U[8]
    pdot[6][8] pcum[6][8] fladd[8] sourceadd[8] divbmax divbavg diss[18]."""
    vals = [e["t"]]
    vals += list(e["U"]) + list(np.ravel(e["pdot"])) + list(np.ravel(e["pcum"]))
    vals += list(e["fladd"]) + list(e["sourceadd"]) + [e["divbmax"], e["divbavg"]]
    vals += [-1.0] * NUMDISSVERSIONS
    out = "%21.15g %ld " % (vals[0], e["nstep"])
    return out + "".join("%21.15g " % v for v in vals[1:]) + "\n"


class Grid:
    def __init__(self, cfg: HarmConfig, handle=None):
        """A grid on one handle; `handle` wraps a handle owned elsewhere (a
        batch member), which close() then leaves alone."""
        self.cfg = cfg
        self.owned = handle is None
        if handle is None:
            h = ctypes.c_void_p()
            check(lib().harm_create(ctypes.byref(cfg), ctypes.byref(h)), "harm_create")
        else:
            h = handle
        self.h = h
        dims = (ctypes.c_int32 * 3)()
        check(lib().harm_storage_dims(self.h, dims), "harm_storage_dims")
        self.ns1, self.ns2, self.ns3 = dims[0], dims[1], dims[2]
        self.ng1 = cfg.ng if cfg.n1 > 1 else 0
        self.ng2 = cfg.ng if cfg.n2 > 1 else 0
        self.ng3 = cfg.ng if cfg.n3 > 1 else 0

    @classmethod
    def torus(cls, n1, n2, n3, **kw):
        return cls(default_config(n1, n2, n3, **kw))

    # ---- layout helpers ----
    @property
    def shape(self):
        return (self.ns3, self.ns2, self.ns1, 8)

    def active(self, a):
        """View of the active zones of a (ns3, ns2, ns1, ...) array."""
        return a[self.ng3:self.ng3 + self.cfg.n3, self.ng2:self.ng2 + self.cfg.n2, self.ng1:self.ng1 + self.cfg.n1]

    def _which(self, which):
        return ARR[which] if isinstance(which, str) else int(which)

    # ---- state exchange ----
    def set(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        check(lib().harm_set_array(self.h, self._which(which), a.ctypes.data), "harm_set_array")

    def get(self, which):
        out = np.empty(self.shape)
        check(lib().harm_get_array(self.h, self._which(which), out.ctypes.data), "harm_get_array")
        return out

    def pflag(self):
        out = np.empty(self.shape[:3], dtype=np.int32)
        check(lib().harm_get_pflag(self.h, out.ctypes.data), "harm_get_pflag")
        return out

    def geometry(self, loc):
        out = np.empty((self.ns2, self.ns1, G_NFIELD))
        check(lib().harm_get_geometry(self.h, loc, out.ctypes.data), "harm_get_geometry")
        return out

    def connection(self):
        out = np.empty((self.ns2, self.ns1, NCONN))
        check(lib().harm_get_connection(self.h, out.ctypes.data), "harm_get_connection")
        return out

    def corner_gdet(self):
        out = np.empty((self.ns2, self.ns1))
        check(lib().harm_get_corner_gdet(self.h, out.ctypes.data), "harm_get_corner_gdet")
        return out

    @property
    def dx(self):
        """[0, dx1, dx2, dx3] of the uniform internal coordinates"""
        X0, X1, V = (ctypes.c_double * 4)(), (ctypes.c_double * 4)(), (ctypes.c_double * 4)()
        check(lib().harm_get_coords(self.h, 0, 0, 0, X0, V), "harm_get_coords")
        check(lib().harm_get_coords(self.h, 1, 1, 1, X1, V), "harm_get_coords")
        return [0.0] + [X1[a] - X0[a] for a in (1, 2, 3)]

    def rcent(self):
        r = np.empty(self.ns1)
        X = (ctypes.c_double * 4)()
        V = (ctypes.c_double * 4)()
        for ii in range(self.ns1):
            check(lib().harm_get_coords(self.h, ii - self.ng1, 0, 0, X, V), "harm_get_coords")
            r[ii] = V[1]
        return r

    def time(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        check(lib().harm_get_time(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), "harm_get_time")
        return t.value, dt.value

    def set_time(self, t, dt, nstep=0):
        check(lib().harm_set_time(self.h, t, dt, nstep), "harm_set_time")

    def nstep(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int64()
        check(lib().harm_get_time(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)), "harm_get_time")
        return n.value

    def set_inversion_guess(self, mlp):
        """Start every inversion of this grid's steps from the network's W'
        (an MLP), or from the previous primitives again (None)."""
        self._guess = mlp  # kept alive while in use
        check(lib().harm_set_inversion_guess(self.h, mlp.h if mlp is not None else None),
              "harm_set_inversion_guess")

    def guess_stats(self):
        """(fallbacks in the last step, fallbacks since set_time)"""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        check(lib().harm_get_guess_stats(self.h, ctypes.byref(a), ctypes.byref(b)), "harm_get_guess_stats")
        return a.value, b.value

    def sample_zones(self, n, seed=1, which_U="ucum", which_p="p"):
        """n active zones (seeded draw): device SoA arrays U (8, n), geometry
        (27, n), p (8, n) for MLP training pairs."""
        from . import DeviceArray
        U, geom, p = DeviceArray((8, n), np.float64), DeviceArray((G_NFIELD, n), np.float64), DeviceArray((8, n), np.float64)
        check(lib().harm_grid_sample_zones(self.h, self._which(which_U), self._which(which_p), n, seed, U.ptr,
                                           geom.ptr, p.ptr), "harm_grid_sample_zones")
        return U, geom, p

    def newton_stats(self):
        """((nstroke, nfail) of the last step, (nstroke, nfail) summed since set_time)"""
        last, tot = NewtonStats(), NewtonStats()
        check(lib().harm_get_newton_stats(self.h, ctypes.byref(last), ctypes.byref(tot)), "harm_get_newton_stats")
        return (last.nstroke, last.nfail), (tot.nstroke, tot.nfail)

    # ---- problem setup ----
    def init_bondi(self):
        check(lib().harm_init_bondi(self.h), "harm_init_bondi")

    def init_torus(self, rin=6.0, rmax=12.0, kappa=1e-3, beta=100.0, randfact=4e-2, seed=1):
        check(lib().harm_init_torus(self.h, rin, rmax, kappa, beta, randfact, seed), "harm_init_torus")

    # ---- hot path ----
    def fluxcalc(self, which="p", do_ct=True):
        ndt = (ctypes.c_double * 3)()
        check(lib().harm_fluxcalc(self.h, self._which(which), int(do_ct), ndt), "harm_fluxcalc")
        return list(ndt)

    def update(self, which_pi, which_pb, CUf, Cunew, dt, timeorder):
        c1 = (ctypes.c_double * 4)(*CUf)
        c2 = (ctypes.c_double * 4)(*Cunew)
        acc = ctypes.c_double()
        check(lib().harm_update(self.h, self._which(which_pi), self._which(which_pb), c1, c2, dt, timeorder,
                                ctypes.byref(acc)), "harm_update")
        return acc.value

    def invert(self, which_pb, which_pf, finalstep):
        st = NewtonStats()
        check(lib().harm_invert_grid(self.h, self._which(which_pb), self._which(which_pf), int(finalstep),
                                     ctypes.byref(st)), "harm_invert_grid")
        return st.nstroke, st.nfail

    def bound(self, which="p"):
        check(lib().harm_bound(self.h, self._which(which)), "harm_bound")

    def post_fixup(self, which="p", finalstep=True):
        check(lib().harm_post_fixup(self.h, self._which(which), int(finalstep)), "harm_post_fixup")

    def step(self, nsteps=1):
        check(lib().harm_step(self.h, nsteps), "harm_step")

    def sync(self):
        check(lib().harm_sync(self.h), "harm_sync")

    def stream(self):
        s = ctypes.c_void_p()
        check(lib().harm_get_stream(self.h, ctypes.byref(s)), "harm_get_stream")
        return s.value

    # ---- energy-file diagnostics (dump_ener.c) ----
    def set_ener_tracking(self, on=True):
        """Steps of this grid add to the energy-file accumulators (cleared now)."""
        check(lib().harm_set_ener_tracking(self.h, int(bool(on))), "harm_set_ener_tracking")

    def reset_ener(self):
        check(lib().harm_reset_ener(self.h), "harm_reset_ener")

    def ener(self, which="p"):
        """One energy-file record as a dict of numpy arrays: t, nstep, U (8),
        pdot and pcum (6 boundaries x 8), fladd, fladdterms (12 categories x
        8), sourceadd, divbmax, divbavg."""
        e = HarmEner()
        check(lib().harm_get_ener(self.h, self._which(which), ctypes.byref(e)), "harm_get_ener")
        return {"t": e.t, "nstep": e.nstep, "U": np.array(e.U[:]), "pdot": np.array([r[:] for r in e.pdot]),
                "pcum": np.array([r[:] for r in e.pcum]), "fladd": np.array(e.fladd[:]),
                "fladdterms": np.array([r[:] for r in e.fladdterms]),
                "sourceadd": np.array(e.sourceadd[:]), "divbmax": e.divbmax, "divbavg": e.divbavg}

    # ---- HARM-format dumps (dump.c) ----
    DUMP_NCOL, GDUMP_NCOL = 72, 126

    def currents(self, at_start=False):
        """The four-current the main dump carries (harm_currents): at the start
        of a run, or after a step that ends at a main dump."""
        check(lib().harm_currents(self.h, int(bool(at_start))), "harm_currents")

    def dump_columns(self, which="p"):
        """the main dump's rows (dump_content): (n1*n2*n3, 72), i fastest"""
        c = self.cfg
        out = np.empty((c.n1 * c.n2 * c.n3, self.DUMP_NCOL))
        check(lib().harm_dump_columns(self.h, self._which(which), out.ctypes.data), "harm_dump_columns")
        return out

    def gdump_columns(self):
        """the grid dump's rows (gdump_content): (n1*n2*n3, 126)"""
        c = self.cfg
        out = np.empty((c.n1 * c.n2 * c.n3, self.GDUMP_NCOL))
        check(lib().harm_gdump_columns(self.h, out.ctypes.data), "harm_gdump_columns")
        return out

    def write_dump(self, path, which="p"):
        check(lib().harm_write_dump(self.h, self._which(which), str(path).encode()), "harm_write_dump")

    def write_gdump(self, path):
        check(lib().harm_write_gdump(self.h, str(path).encode()), "harm_write_gdump")

    # ---- restart files (restart.c) ----
    def write_restart(self, path, binary=True, tf=0.0):
        """rdump: binary=True writes MIXEDOUTPUT (text header, raw doubles; an
        exact restart), False TEXTOUTPUT (15 digits, as the reference)."""
        check(lib().harm_write_restart(self.h, str(path).encode(), 2 if binary else 0, float(tf)),
              "harm_write_restart")

    def read_restart(self, path):
        """Loads an rdump of the same grid and configuration; returns its tf."""
        tf = ctypes.c_double()
        check(lib().harm_read_restart(self.h, str(path).encode(), ctypes.byref(tf)), "harm_read_restart")
        return tf.value

    def boundary_fluxes(self):
        out = (ctypes.c_double * 4)()
        check(lib().harm_get_boundary_fluxes(self.h, out), "harm_get_boundary_fluxes")
        return list(out)

    KERNELS = ("flux1", "flux2", "flux3", "emf_ct", "update", "invert", "bound", "fixup", "other", "recon")

    def profile(self, enable=True):
        check(lib().harm_profile(self.h, int(enable)), "harm_profile")

    def profile_read(self):
        ms = np.zeros(len(self.KERNELS))
        n = np.zeros(len(self.KERNELS), dtype=np.int64)
        check(lib().harm_profile_read(self.h, ms.ctypes.data, n.ctypes.data), "harm_profile_read")
        return {k: (ms[i], int(n[i])) for i, k in enumerate(self.KERNELS)}

    def tune(self, reps=5):
        td = (ctypes.c_int32 * 4)()
        check(lib().harm_tune(self.h, reps, td), "harm_tune")
        return list(td)

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                lib().harm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
