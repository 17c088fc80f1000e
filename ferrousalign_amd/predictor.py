"""Learned work predictor: packs extension pairs by predicted DP cells.

A 6-32-32-1 MLP (FP16 storage, FP32 accumulation) trained with the
library's own HIP kernels (features, forward, backward, reduce + Adam) on the
cell counts that the plain extension kernel records for every pair.  Its
prediction only orders pairs inside each kernel's bucket (FA_PACK_COST_ORDER:
descending predicted cells where a lane group holds 64 or more columns, so
that the lane groups of one wavefront hold pairs that retire together, and
their bands shrink together; FA_PACK_COST: length order with predicted early
stoppers apart); it never changes an alignment result.  Length-sorted packing (FA_PACK_LENGTH) is
the fallback when no trained network is available.

Features (computed on the device from the resident 2-bit sequences): qlen,
tlen, h0, w after end-bonus capping, mismatches in the first 32 bp of the
best diagonal within three of the main one (an indel near the start leaves
the main diagonal looking like a random target), query GC fraction.
"""
import ctypes
import os

import numpy as np

from . import _native as N

NIN, NH = 6, 32
NPARAM = NH * NIN + NH + NH * NH + NH + NH + 1  # 1,313


def _declare(lib):
    if getattr(lib, "_mlp_declared", False):
        return
    P, I32, I64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
    lib.fa_mlp_create.argtypes = [P, ctypes.POINTER(P)]
    lib.fa_mlp_destroy.argtypes = [P, P]
    lib.fa_mlp_set_params.argtypes = [P, P, P, P]
    lib.fa_mlp_get_params.argtypes = [P, P, P]
    lib.fa_batch_features.argtypes = [P, P, P]
    lib.fa_mlp_predict.argtypes = [P, P, P, P, I32]
    lib.fa_mlp_gradient.argtypes = [P, P, P, P, I64, P, ctypes.POINTER(ctypes.c_float)]
    lib.fa_mlp_train.argtypes = [P, P, P, P, I64, I32, ctypes.c_float, ctypes.POINTER(ctypes.c_float)]
    for name in ("fa_mlp_create", "fa_mlp_destroy", "fa_mlp_set_params", "fa_mlp_get_params",
                 "fa_batch_features", "fa_mlp_predict", "fa_mlp_gradient", "fa_mlp_train"):
        getattr(lib, name).restype = ctypes.c_int
    lib._mlp_declared = True


def init_params(seed=0):
    """He-normal weights, zero biases; the output bias starts at log(1+5000)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, np.float32)
    o = 0
    p[o:o + NH * NIN] = rng.normal(0, np.sqrt(2.0 / NIN), NH * NIN)
    o += NH * NIN + NH
    p[o:o + NH * NH] = rng.normal(0, np.sqrt(2.0 / NH), NH * NH)
    o += NH * NH + NH
    p[o:o + NH] = rng.normal(0, np.sqrt(1.0 / NH), NH)
    p[-1] = np.log1p(5000.0)
    return p


def norm_from_features(feat):
    mean = feat.mean(axis=0).astype(np.float32)
    std = feat.std(axis=0).astype(np.float32)
    inv = np.where(std > 1e-6, 1.0 / np.maximum(std, 1e-6), 1.0).astype(np.float32)
    return np.concatenate([mean, inv]).astype(np.float32)


class WorkPredictor:
    def __init__(self, ctx, params=None, norm=None, seed=0):
        self.ctx = ctx
        self.lib = N.lib()
        _declare(self.lib)
        self._h = ctypes.c_void_p()
        N.check(self.lib.fa_mlp_create(ctx.handle, ctypes.byref(self._h)), "fa_mlp_create")
        self.params = init_params(seed) if params is None else np.ascontiguousarray(params, np.float32)
        self.norm = np.concatenate([np.zeros(NIN), np.ones(NIN)]).astype(np.float32) if norm is None \
            else np.ascontiguousarray(norm, np.float32)
        self.upload()

    # -- parameters ---------------------------------------------------------------
    def upload(self):
        N.check(self.lib.fa_mlp_set_params(self.ctx.handle, self._h, N.ptr(self.params), N.ptr(self.norm)),
                "fa_mlp_set_params")

    def download(self):
        p = np.zeros(NPARAM, np.float32)
        N.check(self.lib.fa_mlp_get_params(self.ctx.handle, self._h, N.ptr(p)), "fa_mlp_get_params")
        self.params = p
        return p

    def save(self, path):
        self.download()
        np.savez(path, params=self.params, norm=self.norm)

    @classmethod
    def load(cls, ctx, path):
        z = np.load(path, allow_pickle=False)
        return cls(ctx, params=z["params"], norm=z["norm"])

    # -- device work ----------------------------------------------------------------
    def features(self, db):
        f = np.zeros((db.n, NIN), np.float32)
        N.check(self.lib.fa_batch_features(self.ctx.handle, db._h, N.ptr(f)), "fa_batch_features")
        return f

    def predict(self, db, set_cost=False, fresh_features=False, download=True):
        """Predicted cells of every pair of a device batch; set_cost makes them
        the batch's packing cost (kept in device memory), fresh_features
        computes the features again first (the per-batch cost of a new batch);
        download=False skips the copy to the host and returns None."""
        out = np.zeros(db.n, np.float32) if download else None
        flags = (1 if set_cost else 0) | (2 if fresh_features else 0)
        N.check(self.lib.fa_mlp_predict(self.ctx.handle, self._h, db._h, N.ptr(out), flags), "fa_mlp_predict")
        return out

    def gradient(self, db, pairs):
        pairs = np.ascontiguousarray(pairs, np.uint32)
        g = np.zeros(NPARAM, np.float32)
        loss = ctypes.c_float()
        N.check(self.lib.fa_mlp_gradient(self.ctx.handle, self._h, db._h, N.ptr(pairs), len(pairs), N.ptr(g),
                                         ctypes.byref(loss)), "fa_mlp_gradient")
        return g, loss.value

    def fit(self, db, epochs=3, minibatch=16384, lr=3e-3, seed=0, log=None, holdout=None):
        """Train on the cells recorded by db's last run.  The normalisation
        comes from the training features.  Returns per-epoch mean losses."""
        feat = self.features(db)
        self.norm = norm_from_features(feat)
        self.upload()
        rng = np.random.default_rng(seed)
        losses = []
        for ep in range(epochs):
            order = rng.permutation(db.n).astype(np.uint32)
            loss = ctypes.c_float()
            rate = lr * (0.5 ** ep)
            N.check(self.lib.fa_mlp_train(self.ctx.handle, self._h, db._h, N.ptr(order), db.n, int(minibatch),
                                          ctypes.c_float(rate), ctypes.byref(loss)), "fa_mlp_train")
            losses.append(loss.value)
            if log:
                log("predictor epoch %d: train loss %.4f (lr %.2g)" % (ep, loss.value, rate))
        self.download()
        return losses

    def close(self):
        if self._h:
            self.lib.fa_mlp_destroy(self.ctx.handle, self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mae(pred, cells):
    return float(np.mean(np.abs(pred.astype(np.float64) - cells.astype(np.float64))))
