"""NumPy reference of the proposal MLP (16 -> 64 -> 64 -> 2, ReLU) and of its
pinball-loss gradient, for checking the GPU inference and training kernels.

`quantize=True` rounds what the GPU kernels round to BF16 before each matrix
product (normalised inputs, hidden activations, weights); the sums run in
FP64 here and in FP32 on the matrix cores, so the comparison tolerances allow
for accumulation order.  `quantize=False` is the plain FP64 network, used to
pin the gradient formulas against finite differences.
"""
from __future__ import annotations

import numpy as np

NF, NH, NO = 16, 64, 2


def bf16(x) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def _q(x, quantize):
    return bf16(x).astype(np.float64) if quantize else np.asarray(x, np.float64)


def forward(W: dict, F: np.ndarray, quantize: bool = True):
    """Returns (Y, cache).  W: mu, sigma, w1 (16x64), b1, w2, b2, w3 (64x2), b3."""
    X = ((np.asarray(F, np.float32) - W["mu"].astype(np.float32))
         / W["sigma"].astype(np.float32))
    a0 = _q(X, quantize)
    w1, w2, w3 = (_q(W[k], quantize) for k in ("w1", "w2", "w3"))
    Z1 = a0 @ w1 + W["b1"].astype(np.float64)
    A1 = np.maximum(Z1, 0.0)
    a1 = _q(A1, quantize)
    Z2 = a1 @ w2 + W["b2"].astype(np.float64)
    A2 = np.maximum(Z2, 0.0)
    a2 = _q(A2, quantize)
    Y = a2 @ w3 + W["b3"].astype(np.float64)
    return Y, dict(a0=a0, a1=a1, a2=a2, Z1=Z1, Z2=Z2, w1=w1, w2=w2, w3=w3)


def pinball(Y, T, tau):
    e = np.asarray(T, np.float64) - Y
    return float(np.mean(np.sum(np.where(e > 0, tau * e, (tau - 1.0) * e), axis=1)))


def loss_grad(W: dict, F, T, tau: float, quantize: bool = True):
    """Mean pinball loss and the gradient, flattened as w1, b1, w2, b2, w3, b3."""
    Y, c = forward(W, F, quantize)
    B = Y.shape[0]
    e = np.asarray(T, np.float64) - Y
    dY = np.where(e > 0, -tau, 1.0 - tau) / B
    dw3 = c["a2"].T @ dY
    db3 = dY.sum(0)
    dZ2 = (dY @ c["w3"].T) * (c["Z2"] > 0)
    dw2 = c["a1"].T @ dZ2
    db2 = dZ2.sum(0)
    dZ1 = (dZ2 @ c["w2"].T) * (c["Z1"] > 0)
    dw1 = c["a0"].T @ dZ1
    db1 = dZ1.sum(0)
    g = np.concatenate([dw1.ravel(), db1, dw2.ravel(), db2, dw3.ravel(), db3])
    return pinball(Y, T, tau), g


SLICES = {"w1": (0, NF * NH), "b1": (NF * NH, NF * NH + NH)}
_o = NF * NH + NH
SLICES["w2"] = (_o, _o + NH * NH)
_o += NH * NH
SLICES["b2"] = (_o, _o + NH)
_o += NH
SLICES["w3"] = (_o, _o + NH * NO)
_o += NH * NO
SLICES["b3"] = (_o, _o + NO)
NPARAM = _o + NO


def unflatten(W: dict, theta: np.ndarray) -> dict:
    out = dict(W)
    shapes = {"w1": (NF, NH), "b1": (NH,), "w2": (NH, NH), "b2": (NH,), "w3": (NH, NO),
              "b3": (NO,)}
    for k, (a, b) in SLICES.items():
        out[k] = theta[a:b].reshape(shapes[k])
    return out


def flatten(W: dict) -> np.ndarray:
    return np.concatenate([np.asarray(W[k], np.float64).ravel() for k in SLICES])


def random_weights(seed: int) -> dict:
    r = np.random.default_rng(seed)
    return {"mu": r.normal(0, 1, NF).astype(np.float32),
            "sigma": r.uniform(0.5, 2.0, NF).astype(np.float32),
            "w1": r.normal(0, np.sqrt(2 / NF), (NF, NH)).astype(np.float32),
            "b1": r.normal(0, 0.1, NH).astype(np.float32),
            "w2": r.normal(0, np.sqrt(2 / NH), (NH, NH)).astype(np.float32),
            "b2": r.normal(0, 0.1, NH).astype(np.float32),
            "w3": r.normal(0, np.sqrt(1 / NH), (NH, NO)).astype(np.float32),
            "b3": r.normal(0, 0.1, NO).astype(np.float32)}
