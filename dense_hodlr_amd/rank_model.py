"""Learned rank predictor: Python side of hodlr_rank_model_* (include/dense_hodlr.h).

RankModel wraps the GPU network (8 -> 64 -> 64 -> 1, GELU; FP32 parameters,
BF16 inference) that predicts log(ACA rank) of an off-diagonal block from its
geometry.  collect_samples() runs the plain adaptive ACA on seeded kernel
matrices and pairs each block's features with the rank ACA produced; the
network trains on those.  HODLR_Matrix.set_rank_predictor(model) makes
compression size its panels from the predictions, with ACA still deciding
every rank (the predictor only saves the adaptive panel growth).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import check, lib

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "rank_mlp.npz")
FAMILY_CODE = {"gaussian": 0, "matern32": 1, "se": 2, "imq": 3, "quadratic": 4, "exponential": 5}


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _f32(a, shape=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
    if shape is not None and a.shape[1:] != shape:
        raise ValueError(f"expected trailing shape {shape}, got {a.shape}")
    return a


class RankModel:
    def __init__(self, seed: int = 0):
        h = C.c_void_p()
        check(lib().hodlr_rank_model_create(C.byref(h), int(seed)))
        self._h = h

    def __del__(self):
        if getattr(self, "_h", None) is not None and _lib._lib is not None:
            _lib._lib.hodlr_rank_model_destroy(self._h)
            self._h = None

    @staticmethod
    def num_params() -> int:
        return lib().hodlr_rank_model_num_params()

    @property
    def params(self) -> np.ndarray:
        out = np.zeros(self.num_params(), dtype=np.float32)
        check(lib().hodlr_rank_model_get_params(self._h, _fp(out)))
        return out

    @params.setter
    def params(self, p):
        p = _f32(p)
        if p.shape != (self.num_params(),):
            raise ValueError("wrong number of parameters")
        check(lib().hodlr_rank_model_set_params(self._h, _fp(p)))

    def forward(self, X):
        X = _f32(X, (8,))
        y = np.zeros(len(X), dtype=np.float32)
        check(lib().hodlr_rank_model_forward(self._h, _fp(X), len(X), _fp(y)))
        return y

    def predict_bf16(self, X):
        X = _f32(X, (8,))
        y = np.zeros(len(X), dtype=np.float32)
        check(lib().hodlr_rank_model_predict_bf16(self._h, _fp(X), len(X), _fp(y)))
        return y

    def gradient(self, X, t):
        X, t = _f32(X, (8,)), _f32(t)
        g = np.zeros(self.num_params(), dtype=np.float32)
        loss = C.c_float()
        check(lib().hodlr_rank_model_gradient(self._h, _fp(X), _fp(t), len(X), _fp(g),
                                              C.byref(loss)))
        return float(loss.value), g

    def train(self, X, t, epochs=100, batch=512, lr=1e-3, seed=0):
        X, t = _f32(X, (8,)), _f32(t)
        losses = np.zeros(epochs, dtype=np.float32)
        check(lib().hodlr_rank_model_train(self._h, _fp(X), _fp(t), len(X), int(epochs),
                                           int(batch), float(lr), int(seed), _fp(losses)))
        return losses

    def save(self, path):
        np.savez(path, params=self.params)

    @classmethod
    def load(cls, path=WEIGHTS):
        m = cls()
        with np.load(path, allow_pickle=False) as f:
            m.params = f["params"]
        return m


def block_features(g, tol: float) -> np.ndarray:
    """Features of every off-diagonal block of a kernel HODLR_Matrix (2 per
    node, in nodes() order: top block, then bottom block)."""
    nb = 2 * lib().hodlr_num_nodes(g._h)
    out = np.zeros((max(nb, 1), 8), dtype=np.float32)
    check(lib().hodlr_block_features(g._h, float(tol), _fp(out)), g._h)
    return out[:nb]


def block_ranks(g) -> np.ndarray:
    """ACA ranks in the order of block_features()."""
    nd = g.nodes()
    return np.stack([nd[:, 4], nd[:, 5]], axis=1).ravel()


def collect_samples(runs, leaf=256):
    """runs: iterable of dicts (points, family, length_scale, diag, tol[, leaf]).
    Compresses each with plain adaptive ACA on the GPU and returns features
    X (n x 8) and targets t = log(rank)."""
    from .hodlr_matrix import HODLR_Matrix
    Xs, ts = [], []
    for r in runs:
        g = HODLR_Matrix.from_kernel(r["points"], r["family"], r["length_scale"], r["diag"],
                                     sizeThreshold=r.get("leaf", leaf))
        g.set_LRTolerance(r["tol"])
        g.storeLRinTree()
        Xs.append(block_features(g, r["tol"]))
        ts.append(np.log(np.maximum(block_ranks(g), 1)).astype(np.float32))
        del g
    return np.concatenate(Xs), np.concatenate(ts)


__all__ = ["RankModel", "block_features", "block_ranks", "collect_samples", "WEIGHTS"]
