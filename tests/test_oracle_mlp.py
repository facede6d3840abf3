"""Pins the CPU reference of the inversion-guess network (no GPU): its
backpropagation against central finite differences in FP64, its BF16
rounding against known answers, and the rounding-free path of the
BF16-emulating version against the FP64 gradients."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_ref  # noqa: E402


def rand_params(rng, scale=0.3):
    return rng.normal(0, scale, mlp_ref.NPARAM)


def test_bf16_rounding_known_answers():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -7, -2.5, 0.0, 65504.0], dtype=np.float32)
    # 1 + 2^-8 is halfway between 1 and 1 + 2^-7: ties go to the even mantissa (1.0)
    want = np.array([1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -7, -2.5, 0.0, 65536.0], dtype=np.float32)
    assert np.array_equal(mlp_ref.bf16(x), want)
    y = np.random.default_rng(0).normal(size=1000).astype(np.float32)
    b = mlp_ref.bf16(y)
    assert np.all(b.view(np.uint32) & 0xFFFF == 0)
    assert np.max(np.abs(b - y) / np.abs(y)) <= 2 ** -8


def test_backprop_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = rand_params(rng)
    X = rng.normal(size=(16, 8))
    T = rng.normal(size=(16, 2))
    g = mlp_ref.grads_fp64(p, X, T)
    idx = rng.choice(mlp_ref.NPARAM, 300, replace=False)
    idx = np.concatenate([idx, [0, 511, 512, 575, 576, 4671, 4672, 4735, 4736, 4863, 4864, 4865]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mlp_ref.loss_fp64(p + e, X, T) - mlp_ref.loss_fp64(p - e, X, T)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_emulating_path_without_rounding_equals_fp64():
    rng = np.random.default_rng(2)
    p = rand_params(rng)
    X = rng.normal(size=(128, 8))
    T = rng.normal(size=(128, 2))
    _, g = mlp_ref.loss_and_grads(p, X, T, bf16_ops=False)
    g64 = mlp_ref.grads_fp64(p, X, T)
    # the emulating path keeps FP32 pre-activations and FP32 gradient rounding
    assert np.max(np.abs(g - g64)) <= 1e-5 * np.max(np.abs(g64))
