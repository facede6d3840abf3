"""Learned initial guess for the batched inversion (binding of harm_mlp_* in
libharmstep.so; all compute runs in the library's kernels)."""
import ctypes

import numpy as np

from . import lib, check, DeviceArray

NPARAM, NNORM, NIN, NOUT, BATCH = 4866, 20, 8, 2, 128
# parameter offsets in the FP32 block: W1[8][64] b1[64] W2[64][64] b2[64] W3[64][2] b3[2]
OFF = {"W1": (0, (8, 64)), "b1": (512, (64,)), "W2": (576, (64, 64)), "b2": (4672, (64,)),
       "W3": (4736, (64, 2)), "b3": (4864, (2,))}


def unpack(params):
    return {k: params[o:o + int(np.prod(sh))].reshape(sh) for k, (o, sh) in OFF.items()}


class MLP:
    def __init__(self, seed=1):
        h = ctypes.c_void_p()
        check(lib().harm_mlp_create(seed, ctypes.byref(h)), "harm_mlp_create")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().harm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        w = np.empty(NPARAM, dtype=np.float32)
        nm = np.empty(NNORM, dtype=np.float32)
        check(lib().harm_mlp_get_params(self.h, w.ctypes.data, nm.ctypes.data), "harm_mlp_get_params")
        return w, nm

    def set_params(self, w=None, norm=None):
        w = None if w is None else np.ascontiguousarray(w, dtype=np.float32)
        norm = None if norm is None else np.ascontiguousarray(norm, dtype=np.float32)
        check(lib().harm_mlp_set_params(self.h, w.ctypes.data if w is not None else None,
                                        norm.ctypes.data if norm is not None else None), "harm_mlp_set_params")

    @staticmethod
    def pairs(dU, dgeom, dp, n, gam=4.0 / 3.0, gammafail=2500.0):
        """Device (X[n][8], T[n][2], ok[n]) from solved inversions."""
        X = DeviceArray((n, NIN), np.float32)
        T = DeviceArray((n, NOUT), np.float32)
        ok = DeviceArray((n,), np.int32)
        check(lib().harm_mlp_pairs(dU.ptr, dgeom.ptr, dp.ptr, n, gam, gammafail, X.ptr, T.ptr, ok.ptr, None),
              "harm_mlp_pairs")
        return X, T, ok

    def fit_norm(self, X, T, ok, n):
        check(lib().harm_mlp_fit_norm(self.h, X.ptr, T.ptr, ok.ptr, n), "harm_mlp_fit_norm")

    def train(self, X, T, nx, order, steps, lr=2e-3, lr_end=2e-5, chunk=500, log=None):
        """Adam over `steps` minibatches taken from the device index array
        `order`; the learning rate decays geometrically from lr to lr_end.
        Returns the mean loss of each chunk of steps."""
        losses = []
        done = 0
        norder = order.shape[0]
        while done < steps:
            k = min(chunk, steps - done)
            frac = (done + 0.5 * k) / steps
            rate = lr * (lr_end / lr) ** frac
            loss = ctypes.c_float()
            check(lib().harm_mlp_train(self.h, X.ptr, T.ptr, nx, order.ptr, norder, done * BATCH, k,
                                       ctypes.c_float(rate), ctypes.byref(loss), None), "harm_mlp_train")
            losses.append(loss.value)
            done += k
            if log:
                log(f"mlp train step {done}/{steps} lr {rate:.2e} loss {loss.value:.4e}")
        return losses

    def grad(self, X, T, nx, order):
        g = DeviceArray((NPARAM,), np.float32)
        loss = ctypes.c_float()
        check(lib().harm_mlp_grad(self.h, X.ptr, T.ptr, nx, order.ptr, g.ptr, ctypes.byref(loss)), "harm_mlp_grad")
        out = g.numpy()
        g.free()
        return out, loss.value

    def forward(self, X, n):
        Y = DeviceArray((n, NOUT), np.float32)
        check(lib().harm_mlp_forward(self.h, X.ptr, Y.ptr, n, None), "harm_mlp_forward")
        return Y

    def invert_batch(self, dU, dgeom, dp, dflag, dniter, n, gam=4.0 / 3.0, gammafail=2500.0, dfallback=None):
        check(lib().harm_invert_batch_mlp(self.h, dU.ptr, dgeom.ptr, dp.ptr, dflag.ptr,
                                          dniter.ptr if dniter is not None else None, n, gam, gammafail,
                                          dfallback.ptr if dfallback is not None else None, None),
              "harm_invert_batch_mlp")
