"""CPU reference of the inversion-guess network (8 -> 64 -> 64 -> 2, tanh).

TEST INFRASTRUCTURE ONLY: the checker the GPU forward and backward kernels
are compared with; nothing in the library uses it.  The network has no
counterpart in the reference (HARM starts every Newton solve from the previous
primitives), so there is no file:line to cite; the formulas are those of a
plain multilayer perceptron with a mean-squared loss, and `grads` is pinned by
central finite differences in FP64 (tests/test_oracle_mlp.py).

`bf16=True` reproduces the GPU's rounding points: matrix-product operands
(standardised inputs, weights, hidden activations, output and hidden-layer
gradients) are rounded to BF16 (round to nearest even), products accumulate
in FP32 or better, biases, tanh and the loss stay FP32.
"""
import numpy as np

NIN, NH, NOUT = 8, 64, 2
OFF = {"W1": (0, (8, 64)), "b1": (512, (64,)), "W2": (576, (64, 64)), "b2": (4672, (64,)),
       "W3": (4736, (64, 2)), "b3": (4864, (2,))}
NPARAM = 4866


def bf16(x):
    """Round float32 values to the nearest BF16 value (ties to even)."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u >> 16) & 1) + 0x7FFF
    out = (((u + r) >> 16) << 16).astype(np.uint32)
    nan = np.isnan(a)
    out = np.where(nan, (a.view(np.uint32) | 0x00400000) & 0xFFFF0000, out).astype(np.uint32)
    return out.view(np.float32)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    return {k: p[o:o + int(np.prod(sh))].reshape(sh) for k, (o, sh) in OFF.items()}


def standardise(X, T, norm):
    norm = np.asarray(norm, dtype=np.float32)
    Xn = ((np.asarray(X, np.float32) - norm[0:8]) * norm[8:16]).astype(np.float32)
    Tn = None if T is None else ((np.asarray(T, np.float32) - norm[16:18]) / norm[18:20]).astype(np.float32)
    return Xn, Tn


def _r(x, on):
    return bf16(x).astype(np.float64) if on else np.asarray(x, np.float64)


def forward(params, Xn, bf16_ops=True):
    """Standardised inputs (n, 8) -> standardised outputs (n, 2) and the
    intermediates the backward pass needs."""
    P = unpack(params)
    c = {}
    c["X"] = _r(Xn, bf16_ops)
    W1, W2, W3 = (_r(P[k].astype(np.float32), bf16_ops) for k in ("W1", "W2", "W3"))
    z1 = (c["X"] @ W1).astype(np.float32) + P["b1"].astype(np.float32)
    c["H1"] = _r(np.tanh(z1.astype(np.float32)), bf16_ops)
    z2 = (c["H1"] @ W2).astype(np.float32) + P["b2"].astype(np.float32)
    c["H2"] = _r(np.tanh(z2.astype(np.float32)), bf16_ops)
    Y = (c["H2"] @ W3).astype(np.float32) + P["b3"].astype(np.float32)
    c["W1"], c["W2"], c["W3"] = W1, W2, W3
    return Y.astype(np.float64), c


def loss_and_grads(params, Xn, Tn, bf16_ops=True):
    """Mean-squared loss over the batch and both outputs, and its gradient
    with respect to the FP32 parameter block (same layout as the library)."""
    Y, c = forward(params, Xn, bf16_ops)
    B = Xn.shape[0]
    d = (Y - np.asarray(Tn, np.float64))
    loss = float(np.mean(d * d))
    dY = (d * (2.0 / (B * NOUT))).astype(np.float32).astype(np.float64)
    dYr = _r(dY, bf16_ops)
    g = {}
    g["W3"] = c["H2"].T @ dYr
    g["b3"] = dY.sum(axis=0)
    dH2 = dYr @ c["W3"].T
    G2 = _r((dH2 * (1.0 - c["H2"] ** 2)).astype(np.float32), bf16_ops)
    g["W2"] = c["H1"].T @ G2
    g["b2"] = G2.sum(axis=0)
    dH1 = G2 @ c["W2"].T
    G1 = _r((dH1 * (1.0 - c["H1"] ** 2)).astype(np.float32), bf16_ops)
    g["W1"] = c["X"].T @ G1
    g["b1"] = G1.sum(axis=0)
    out = np.zeros(NPARAM)
    for k, (o, sh) in OFF.items():
        out[o:o + int(np.prod(sh))] = g[k].reshape(-1)
    return loss, out


def loss_fp64(params, Xn, Tn):
    """The same loss in plain FP64 (no BF16 rounding): the function that the
    finite-difference check differentiates."""
    P = unpack(params)
    X = np.asarray(Xn, np.float64)
    H1 = np.tanh(X @ P["W1"] + P["b1"])
    H2 = np.tanh(H1 @ P["W2"] + P["b2"])
    Y = H2 @ P["W3"] + P["b3"]
    d = Y - np.asarray(Tn, np.float64)
    return float(np.mean(d * d))


def grads_fp64(params, Xn, Tn):
    """Backpropagation in FP64 without rounding (pinned by finite differences)."""
    P = unpack(params)
    X = np.asarray(Xn, np.float64)
    H1 = np.tanh(X @ P["W1"] + P["b1"])
    H2 = np.tanh(H1 @ P["W2"] + P["b2"])
    Y = H2 @ P["W3"] + P["b3"]
    B = X.shape[0]
    dY = (Y - np.asarray(Tn, np.float64)) * (2.0 / (B * NOUT))
    g = {"W3": H2.T @ dY, "b3": dY.sum(0)}
    G2 = (dY @ P["W3"].T) * (1 - H2 ** 2)
    g["W2"], g["b2"] = H1.T @ G2, G2.sum(0)
    G1 = (G2 @ P["W2"].T) * (1 - H1 ** 2)
    g["W1"], g["b1"] = X.T @ G1, G1.sum(0)
    out = np.zeros(NPARAM)
    for k, (o, sh) in OFF.items():
        out[o:o + int(np.prod(sh))] = g[k].reshape(-1)
    return out
