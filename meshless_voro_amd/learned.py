"""Learned neighbour-count proposal: features, training, inference and the
checkpoint format of the MLP (16 features -> 64 -> 64 -> 2, BF16 weights on
MFMA, FP32 accumulation) that proposes, per generator, the neighbour-list
length of the first pass (output 0: log k) and the start radius of its
neighbour gather (output 1: log(security radius / local spacing)).

The proposal only sizes the work: a cell whose security radius is not reached
by its list is redone with a doubled list, exactly as in the plain mode, so a
learned build returns the same cells and faces as a plain one.

All compute runs in libmvoro (mv_model_* of include/mvoro.h); there is no CPU
path.  The reference has no counterpart: it grows the neighbour search
lazily (src/voronoi/convex_cell.rs:316-356 asks its k-d tree for one more
neighbour at a time).
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Tuple

import numpy as np

from . import _native

NF, NH, NO = 16, 64, 2
CHECKPOINT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "learned_k.mvlk")


class ModelError(RuntimeError):
    pass


def _check(st: int, what: str) -> None:
    if st != 0:
        msg = _native.lib().mv_status_string(st).decode()
        raise ModelError(f"{what}: {msg} ({st})")


class Model:
    """A proposal checkpoint (the mv_model struct).  File format: the struct's
    bytes, little-endian, magic "MVLK", version 1."""

    def __init__(self, raw: _native.mv_model):
        if raw.magic != _native.MV_MODEL_MAGIC or raw.version != 1:
            raise ModelError("not a version-1 mv_model")
        self.raw = raw

    @classmethod
    def from_weights(cls, W: dict, tau: float = 0.95, kscale: float = 1.0,
                     rscale: float = 0.0) -> "Model":
        """Checkpoint from FP32 arrays (mu, sigma, w1 16x64, b1, w2 64x64, b2,
        w3 64x2, b3); weights are rounded to BF16 (nearest, ties to even)."""
        def bf(a, n):
            u = np.ascontiguousarray(a, np.float32).ravel().view(np.uint32).astype(np.uint64)
            assert u.size == n
            return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
        r = _native.mv_model()
        r.magic, r.version = _native.MV_MODEL_MAGIC, 1
        r.mu[:] = [float(v) for v in np.asarray(W["mu"], np.float32)]
        r.sigma[:] = [float(v) for v in np.asarray(W["sigma"], np.float32)]
        r.w1[:] = bf(W["w1"], NF * NH).tolist()
        r.w2[:] = bf(W["w2"], NH * NH).tolist()
        r.w3[:] = bf(W["w3"], NH * NO).tolist()
        r.b1[:] = [float(v) for v in np.asarray(W["b1"], np.float32)]
        r.b2[:] = [float(v) for v in np.asarray(W["b2"], np.float32)]
        r.b3[:] = [float(v) for v in np.asarray(W["b3"], np.float32)]
        r.tau, r.kscale, r.rscale = tau, kscale, rscale
        return cls(r)

    # -- serialisation ----------------------------------------------------
    @classmethod
    def load(cls, path: str = CHECKPOINT) -> "Model":
        data = open(path, "rb").read()
        if len(data) != C.sizeof(_native.mv_model):
            raise ModelError(f"{path}: {len(data)} bytes, expected "
                             f"{C.sizeof(_native.mv_model)}")
        return cls(_native.mv_model.from_buffer_copy(data))

    def save(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(bytes(self.raw))

    @property
    def tau(self) -> float:
        return float(self.raw.tau)

    @property
    def kscale(self) -> float:
        return float(self.raw.kscale)

    @kscale.setter
    def kscale(self, v: float) -> None:
        self.raw.kscale = float(v)

    @property
    def rscale(self) -> float:
        return float(self.raw.rscale)

    @rscale.setter
    def rscale(self, v: float) -> None:
        self.raw.rscale = float(v)

    def weights(self) -> dict:
        """FP32 views of the (BF16) weights: w1 16x64, w2 64x64, w3 64x2."""
        def bf(a, shape):
            u = np.frombuffer(bytes(a), dtype=np.uint16).astype(np.uint32) << 16
            return u.view(np.float32).reshape(shape).copy()
        r = self.raw
        return {"mu": np.array(r.mu[:], np.float32), "sigma": np.array(r.sigma[:], np.float32),
                "w1": bf(r.w1, (NF, NH)), "w2": bf(r.w2, (NH, NH)), "w3": bf(r.w3, (NH, NO)),
                "b1": np.array(r.b1[:], np.float32), "b2": np.array(r.b2[:], np.float32),
                "b3": np.array(r.b3[:], np.float32)}

    # -- compute -------------------------------------------------------------
    def predict(self, features: np.ndarray) -> np.ndarray:
        """Raw outputs (n x 2) of the GPU inference kernel."""
        F = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, NF)
        out = np.zeros((F.shape[0], NO), np.float32)
        _check(_native.lib().mv_model_predict(C.byref(self.raw), F.ctypes.data, F.shape[0],
                                              out.ctypes.data), "mv_model_predict")
        return out

    def loss_grad(self, features: np.ndarray, targets: np.ndarray,
                  tau: Optional[float] = None) -> Tuple[float, np.ndarray]:
        """Mean pinball loss and its gradient (5378 floats: w1, b1, w2, b2, w3,
        b3) from one pass of the training kernels."""
        F = np.ascontiguousarray(features, dtype=np.float32).reshape(-1, NF)
        T = np.ascontiguousarray(targets, dtype=np.float32).reshape(-1, NO)
        loss = C.c_float()
        g = np.zeros(NF * NH + NH + NH * NH + NH + NH * NO + NO, np.float32)
        _check(_native.lib().mv_model_loss_grad(
            C.byref(self.raw), F.ctypes.data, T.ctypes.data, F.shape[0],
            float(self.tau if tau is None else tau), C.byref(loss), g.ctypes.data),
            "mv_model_loss_grad")
        return float(loss.value), g


def features(generators, anchor, width, dimensionality: int = 3,
             periodic: bool = False) -> Tuple[np.ndarray, np.ndarray]:
    """The 16 proposal features (n x 16) and local spacing (n) of every
    generator, computed by the same GPU kernel a learned build runs."""
    xyz = np.ascontiguousarray(np.asarray(generators, dtype=np.float64).reshape(-1, 3))
    n = xyz.shape[0]
    F = np.zeros((n, NF), np.float32)
    h = np.zeros(n, np.float32)
    a = np.ascontiguousarray(np.asarray(anchor, np.float64).reshape(3))
    w = np.ascontiguousarray(np.asarray(width, np.float64).reshape(3))
    _check(_native.lib().mv_model_features(xyz.ctypes.data, n, a.ctypes.data, w.ctypes.data,
                                           int(dimensionality), int(bool(periodic)),
                                           F.ctypes.data, h.ctypes.data), "mv_model_features")
    return F, h


def targets(voronoi, hloc: np.ndarray) -> np.ndarray:
    """Training targets from an exact build: log of the list length a cell
    needed (entries processed, self included: popped - 1) and log of its
    final security radius over the local spacing."""
    k = np.maximum(voronoi.cell_n_popped.astype(np.float64) - 1.0, 1.0)
    sr = np.maximum(voronoi.cell_safety_radius, 1e-300)
    return np.stack([np.log(k), np.log(sr / np.maximum(hloc.astype(np.float64), 1e-300))],
                    axis=1).astype(np.float32)


def train(features_: np.ndarray, targets_: np.ndarray, steps: int = 3000, batch: int = 4096,
          seed: int = 1, lr: float = 2e-3, tau: float = 0.95) -> Tuple[Model, float]:
    """Train on the GPU; returns the model and its mean pinball loss over every
    row (BF16 checkpoint through the inference kernel)."""
    F = np.ascontiguousarray(features_, dtype=np.float32).reshape(-1, NF)
    T = np.ascontiguousarray(targets_, dtype=np.float32).reshape(-1, NO)
    cfg = _native.mv_train_config(steps=steps, batch=batch, seed=seed, lr=lr, tau=tau)
    raw = _native.mv_model()
    loss = C.c_float()
    _check(_native.lib().mv_model_train(F.ctypes.data, T.ctypes.data, F.shape[0],
                                        C.byref(cfg), C.byref(raw), C.byref(loss)),
           "mv_model_train")
    return Model(raw), float(loss.value)


_default: Optional[Model] = None


def default_model() -> Model:
    """The committed checkpoint (learned_k.mvlk beside this file)."""
    global _default
    if _default is None:
        if not os.path.exists(CHECKPOINT):
            raise ModelError(f"{CHECKPOINT} is missing: train one with "
                             "meshless_voro_amd.learned.train and save it there")
        _default = Model.load(CHECKPOINT)
    return _default
