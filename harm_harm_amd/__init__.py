"""harm_harm_amd -- MI355X-native HARM advance() step behind a C ABI.

The compute path is libharmstep.so (hand-written gfx950 FP64 kernels plus a C
host driver, declared in include/harmstep.h).  This module only binds that
library with ctypes; it has no compute fallback.  Loading fails loudly when the
library has not been built.
"""
import ctypes
import os

__all__ = ["lib", "HarmConfig", "NewtonStats", "HarmError", "check", "default_config", "LIB_PATH", "DeviceArray",
           "device_count", "synchronize", "Event"]

_HERE = os.path.dirname(os.path.abspath(__file__))
# HARM_HARM_AMD_LIB points at another build of the library (same ABI)
LIB_PATH = os.environ.get("HARM_HARM_AMD_LIB") or os.path.join(_HERE, "libharmstep.so")

NPR = 8
RHO, UU, U1, U2, U3, B1, B2, B3 = range(8)
G_NFIELD = 27
NCONN = 40
RECON = {"donor": 0, "mc": 1, "para": 2, "weno5": 3}
FLUX = {"laxf": 0, "hll": 1}
ARR = {"p": 0, "pk0": 1, "pk1": 2, "ui": 3, "uf": 4, "ucum": 5, "F1": 6, "F2": 7, "F3": 8}


class HarmError(RuntimeError):
    pass


class HarmConfig(ctypes.Structure):
    """Mirror of harm_config (include/harmstep.h)."""
    _fields_ = [
        ("n1", ctypes.c_int32), ("n2", ctypes.c_int32), ("n3", ctypes.c_int32), ("ng", ctypes.c_int32),
        ("recon", ctypes.c_int32), ("flux", ctypes.c_int32), ("timeorder", ctypes.c_int32),
        ("problem", ctypes.c_int32),
        ("a", ctypes.c_double), ("gam", ctypes.c_double), ("R0", ctypes.c_double), ("Rin", ctypes.c_double),
        ("Rout", ctypes.c_double), ("hslope", ctypes.c_double), ("cour", ctypes.c_double),
        ("safe", ctypes.c_double), ("rhomin", ctypes.c_double), ("uumin", ctypes.c_double),
        ("gammamax", ctypes.c_double), ("gammafail", ctypes.c_double),
        ("hot2cold", ctypes.c_int32), ("checkinv", ctypes.c_int32), ("resetregion", ctypes.c_int32),
    ]


class NewtonStats(ctypes.Structure):
    _fields_ = [("nstroke", ctypes.c_int64), ("nfail", ctypes.c_int64)]


class HarmEner(ctypes.Structure):
    """harm_ener: one energy-file record (dump_ener.c:391-407)."""
    _fields_ = [("t", ctypes.c_double), ("nstep", ctypes.c_int64), ("U", ctypes.c_double * 8),
                ("pdot", (ctypes.c_double * 8) * 6), ("pcum", (ctypes.c_double * 8) * 6),
                ("fladd", ctypes.c_double * 8), ("fladdterms", (ctypes.c_double * 8) * 12),
                ("sourceadd", ctypes.c_double * 8),
                ("divbmax", ctypes.c_double), ("divbavg", ctypes.c_double)]


_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int32
_L = ctypes.c_int64
_DP = ctypes.POINTER(ctypes.c_double)

# (name, restype, argtypes); harm_status is an int.
_PROTOS = [
    ("harm_last_error", ctypes.c_char_p, []),
    ("harm_default_config", None, [ctypes.POINTER(HarmConfig), _I, _I, _I]),
    ("harm_bondi_config", None, [ctypes.POINTER(HarmConfig), _I, _I]),
    ("harm_geom_ks_record", None, [_D, _D, _D, _P]),
    ("harm_geom_mks_record", None, [ctypes.POINTER(HarmConfig), _P, _P]),
    ("harm_geom_ks_batch", None, [_L, _P, _P, _P, _P]),
    ("harm_invert_batch", _I, [_P, _P, _P, _P, _P, _L, _D, _D, _P]),
    ("harm_prim_to_cons_batch", _I, [_P, _P, _P, _L, _D, _P]),
    ("harm_host_tables", _I, [ctypes.POINTER(HarmConfig), _P, _P, _P, _P, _P, _P]),
    ("harm_host_init_torus", _I, [ctypes.POINTER(HarmConfig), _D, _D, _D, _D, _D, ctypes.c_uint32, _P]),
    ("harm_device_count", _I, [ctypes.POINTER(_I)]),
    ("harm_dev_malloc", _I, [ctypes.POINTER(_P), ctypes.c_size_t]),
    ("harm_dev_free", _I, [_P]),
    ("harm_memcpy_htod", _I, [_P, _P, ctypes.c_size_t]),
    ("harm_memcpy_dtoh", _I, [_P, _P, ctypes.c_size_t]),
    ("harm_memcpy_dtod", _I, [_P, _P, ctypes.c_size_t]),
    ("harm_device_synchronize", _I, []),
    ("harm_event_create", _I, [ctypes.POINTER(_P)]),
    ("harm_event_destroy", _I, [_P]),
    ("harm_event_record", _I, [_P, _P]),
    ("harm_event_elapsed_ms", _I, [_P, _P, ctypes.POINTER(ctypes.c_float)]),
]

_OPTIONAL_PROTOS = [
    ("harm_create", _I, [ctypes.POINTER(HarmConfig), ctypes.POINTER(_P)]),
    ("harm_destroy", _I, [_P]),
    ("harm_sync", _I, [_P]),
    ("harm_storage_dims", _I, [_P, _P]),
    ("harm_get_geometry", _I, [_P, _I, _P]),
    ("harm_get_connection", _I, [_P, _P]),
    ("harm_get_corner_gdet", _I, [_P, _P]),
    ("harm_get_coords", _I, [_P, _I, _I, _I, _P, _P]),
    ("harm_set_array", _I, [_P, _I, _P]),
    ("harm_get_array", _I, [_P, _I, _P]),
    ("harm_get_pflag", _I, [_P, _P]),
    ("harm_set_time", _I, [_P, _D, _D, _L]),
    ("harm_get_time", _I, [_P, _P, _P, _P]),
    ("harm_init_bondi", _I, [_P]),
    ("harm_host_init_bondi", _I, [_P, _P]),
    ("harm_bondi_state", _I, [_D, _D, _P]),
    ("harm_set_inversion_guess", _I, [_P, _P]),
    ("harm_get_guess_stats", _I, [_P, _P, _P]),
    ("harm_grid_sample_zones", _I, [_P, _I, _I, _L, ctypes.c_uint64, _P, _P, _P]),
    ("harm_get_newton_stats", _I, [_P, _P, _P]),
    ("harm_init_torus", _I, [_P, _D, _D, _D, _D, _D, ctypes.c_uint32]),
    ("harm_fluxcalc", _I, [_P, _I, _I, _P]),
    ("harm_update", _I, [_P, _I, _I, _P, _P, _D, _I, _P]),
    ("harm_invert_grid", _I, [_P, _I, _I, _I, ctypes.POINTER(NewtonStats)]),
    ("harm_bound", _I, [_P, _I]),
    ("harm_post_fixup", _I, [_P, _I, _I]),
    ("harm_advance", _I, [_P, _I, _P]),
    ("harm_step", _I, [_P, _I]),
    ("harm_get_boundary_fluxes", _I, [_P, _P]),
    ("harm_set_ener_tracking", _I, [_P, _I]),
    ("harm_reset_ener", _I, [_P]),
    ("harm_get_ener", _I, [_P, _I, _P]),
    ("harm_dump_columns", _I, [_P, _I, _P]),
    ("harm_currents", _I, [_P, _I]),
    ("harm_gdump_columns", _I, [_P, _P]),
    ("harm_write_dump", _I, [_P, _I, ctypes.c_char_p]),
    ("harm_write_gdump", _I, [_P, ctypes.c_char_p]),
    ("harm_write_restart", _I, [_P, ctypes.c_char_p, _I, ctypes.c_double]),
    ("harm_read_restart", _I, [_P, ctypes.c_char_p, _P]),
    ("harm_get_stream", _I, [_P, ctypes.POINTER(_P)]),
    ("harm_profile", _I, [_P, _I]),
    ("harm_profile_read", _I, [_P, _P, _P]),
    ("harm_tune", _I, [_P, _I, _P]),
    ("harm_batch_create", _I, [_P, _I, ctypes.POINTER(_P)]),
    ("harm_batch_destroy", _I, [_P]),
    ("harm_batch_grid", _I, [_P, _I, ctypes.POINTER(_P)]),
    ("harm_batch_tune", _I, [_P, _I]),
    ("harm_batch_step", _I, [_P, _I, _I]),
    ("harm_batch_sync", _I, [_P]),
    ("harm_mlp_create", _I, [ctypes.c_uint64, ctypes.POINTER(_P)]),
    ("harm_mlp_destroy", _I, [_P]),
    ("harm_mlp_get_params", _I, [_P, _P, _P]),
    ("harm_mlp_set_params", _I, [_P, _P, _P]),
    ("harm_mlp_pairs", _I, [_P, _P, _P, _L, _D, _D, _P, _P, _P, _P]),
    ("harm_mlp_fit_norm", _I, [_P, _P, _P, _P, _L]),
    ("harm_mlp_train", _I, [_P, _P, _P, _L, _P, _L, _L, _I, ctypes.c_float, ctypes.POINTER(ctypes.c_float), _P]),
    ("harm_mlp_grad", _I, [_P, _P, _P, _L, _P, _P, ctypes.POINTER(ctypes.c_float)]),
    ("harm_mlp_forward", _I, [_P, _P, _P, _L, _P]),
    ("harm_invert_batch_mlp", _I, [_P, _P, _P, _P, _P, _P, _L, _D, _D, _P, _P]),
]

_lib = None


def lib():
    """Return the loaded libharmstep.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (python -c 'import __graft_entry__ as g; g.build()')")
    L = ctypes.CDLL(LIB_PATH)
    for name, res, args in _PROTOS + _OPTIONAL_PROTOS:
        if not hasattr(L, name):
            continue
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status, what=""):
    if status != 0:
        msg = lib().harm_last_error()
        raise HarmError(f"{what}: status {status}: {msg.decode() if msg else ''}")


def default_config(n1, n2, n3, **kw):
    c = HarmConfig()
    lib().harm_default_config(ctypes.byref(c), n1, n2, n3)
    for k, v in kw.items():
        if k == "recon" and isinstance(v, str):
            v = RECON[v]
        if k == "flux" and isinstance(v, str):
            v = FLUX[v]
        setattr(c, k, v)
    return c


def bondi_config(n1, n2, **kw):
    """The Bondi problem's setup on an n1 x n2 (r, theta) grid; keyword
    arguments override fields as in default_config."""
    c = HarmConfig()
    lib().harm_bondi_config(ctypes.byref(c), n1, n2)
    for k, v in kw.items():
        if k == "recon" and isinstance(v, str):
            v = RECON[v]
        if k == "flux" and isinstance(v, str):
            v = FLUX[v]
        setattr(c, k, v)
    return c


def bondi_state(r, gam=4.0 / 3.0):
    """(rho, u, u^r) of the Bondi flow at radius r (host computation)."""
    out = (ctypes.c_double * 3)()
    check(lib().harm_bondi_state(gam, r, out), "harm_bondi_state")
    return tuple(out)


def device_count():
    n = ctypes.c_int32(0)
    lib().harm_device_count(ctypes.byref(n))
    return n.value


def synchronize():
    check(lib().harm_device_synchronize(), "hipDeviceSynchronize")


class DeviceArray:
    """A device allocation owned by this object (hipMalloc through the library)."""

    def __init__(self, shape, dtype):
        import numpy as np
        self.shape = tuple(shape) if hasattr(shape, "__len__") else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        ptr = ctypes.c_void_p()
        check(lib().harm_dev_malloc(ctypes.byref(ptr), self.nbytes), "dev_malloc")
        self.ptr = ptr.value

    @classmethod
    def from_numpy(cls, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        check(lib().harm_memcpy_htod(d.ptr, a.ctypes.data, d.nbytes), "memcpy_htod")
        return d

    @classmethod
    def zeros(cls, shape, dtype):
        import numpy as np
        return cls.from_numpy(np.zeros(shape, dtype=dtype))

    def copy_from(self, a):
        import numpy as np
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.nbytes == self.nbytes
        check(lib().harm_memcpy_htod(self.ptr, a.ctypes.data, self.nbytes), "memcpy_htod")

    def copy_from_device(self, other):
        assert other.nbytes == self.nbytes
        check(lib().harm_memcpy_dtod(self.ptr, other.ptr, self.nbytes), "memcpy_dtod")

    def numpy(self):
        import numpy as np
        out = np.empty(self.shape, dtype=self.dtype)
        check(lib().harm_memcpy_dtoh(out.ctypes.data, self.ptr, self.nbytes), "memcpy_dtoh")
        return out

    def free(self):
        if getattr(self, "ptr", None):
            lib().harm_dev_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    """A HIP event for timing work on a given stream."""

    def __init__(self):
        e = ctypes.c_void_p()
        check(lib().harm_event_create(ctypes.byref(e)), "event_create")
        self.ev = e.value

    def record(self, stream=None):
        check(lib().harm_event_record(self.ev, stream), "event_record")

    def elapsed_ms(self, later):
        ms = ctypes.c_float()
        check(lib().harm_event_elapsed_ms(self.ev, later.ev, ctypes.byref(ms)), "event_elapsed")
        return ms.value
