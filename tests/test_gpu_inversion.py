"""GPU parity of the per-zone kernels against the oracle (batched form).

The HIP kernels are compiled without FMA contraction and follow the
reference's operation order, so results are expected to be bit-identical to
the CPU restatement; the assertion bar is 1e-12 relative (the tolerance the
step's specification states for FP64), flags and Newton iteration counts must
match exactly.
"""
import numpy as np
import pytest

import _oracle
import harm_harm_amd as H
from harm_harm_amd.sampling import sample_inversion_inputs

pytestmark = pytest.mark.gpu


D = H.DeviceArray


def _gpu_p2u(p, geom):
    dp, dg = D.from_numpy(p), D.from_numpy(geom)
    dU = D(p.shape, np.float64)
    H.check(H.lib().harm_prim_to_cons_batch(dp.ptr, dg.ptr, dU.ptr, p.shape[1], 4.0 / 3.0, None), "p2u")
    H.synchronize()
    return dU.numpy()


def _gpu_invert(U, geom, guess):
    n = U.shape[1]
    dU, dg, dp = D.from_numpy(U), D.from_numpy(geom), D.from_numpy(guess)
    fl = D.zeros(n, np.int32)
    it = D.zeros(n, np.int32)
    H.check(H.lib().harm_invert_batch(dU.ptr, dg.ptr, dp.ptr, fl.ptr, it.ptr, n, 4.0 / 3.0, 2500.0, None), "invert")
    H.synchronize()
    return dp.numpy(), fl.numpy(), it.numpy()


def _rel(a, b):
    scale = np.maximum(np.abs(a), np.abs(b))
    scale[scale == 0] = 1.0
    return np.abs(a - b) / scale


@pytest.mark.parametrize("seed", [7, 8])
def test_prim_to_cons_matches_oracle(seed):
    p, _, geom = sample_inversion_inputs(1 << 15, seed=seed)
    Ug = _gpu_p2u(p, geom)
    Uo = _oracle.p2u_batch(p, geom)
    assert _rel(Ug, Uo).max() <= 1e-12


@pytest.mark.parametrize("n,seed", [(1 << 15, 7), (12345, 3)])
def test_invert_batch_matches_oracle(n, seed):
    p, guess, geom = sample_inversion_inputs(n, seed=seed)
    U = _oracle.p2u_batch(p, geom)
    pg, fg, ig = _gpu_invert(U, geom, guess)
    po, fo, io = _oracle.invert_batch(U, geom, guess)
    print(f"n={n} seed={seed} exact={np.mean(pg == po):.6f} fail={np.mean(fo != 0):.4f} iters={io.mean():.3f}")
    assert np.array_equal(fg, fo)
    assert np.array_equal(ig, io)
    assert _rel(pg, po).max() <= 1e-12


def test_invert_batch_edge_cases():
    """Tiny batches (one partial wavefront), an unphysical zone mixed in."""
    p, guess, geom = sample_inversion_inputs(37, seed=5)
    U = _oracle.p2u_batch(p, geom)
    U[1, 3] = 1.0 * geom[21, 3]
    pg, fg, ig = _gpu_invert(U, geom, guess)
    po, fo, io = _oracle.invert_batch(U, geom, guess)
    assert fo[3] != 0
    assert np.array_equal(fg, fo) and np.array_equal(ig, io)
    assert _rel(pg, po).max() <= 1e-12
