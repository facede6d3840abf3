"""Batches of independent grids stepped together (binding of harm_batch_* in
libharmstep.so).  Each member is a Grid on a handle the batch owns, so every
single-grid call (set/get, init_torus, time, the hot-path pieces) works on it."""
import ctypes

from . import lib, check, HarmConfig
from .grid import Grid


class Batch:
    def __init__(self, cfgs):
        arr = (HarmConfig * len(cfgs))(*cfgs)
        b = ctypes.c_void_p()
        check(lib().harm_batch_create(arr, len(cfgs), ctypes.byref(b)), "harm_batch_create")
        self.b = b
        self.grids = []
        for i, c in enumerate(cfgs):
            h = ctypes.c_void_p()
            check(lib().harm_batch_grid(self.b, i, ctypes.byref(h)), "harm_batch_grid")
            self.grids.append(Grid(c, handle=h))

    def step(self, nsteps=1, graph=True):
        check(lib().harm_batch_step(self.b, nsteps, int(graph)), "harm_batch_step")

    def sync(self):
        check(lib().harm_batch_sync(self.b), "harm_batch_sync")

    def tune(self, reps=5):
        check(lib().harm_batch_tune(self.b, reps), "harm_batch_tune")

    def close(self):
        if getattr(self, "b", None):
            for g in self.grids:
                g.close()
            lib().harm_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
