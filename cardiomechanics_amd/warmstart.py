"""Learned Newton warm start for Newmark steps (cmg_mlp_* in include/cmg.h).

A per-node MLP (8 -> 128 -> 128 -> 3, tanh) predicts each Newmark step's
displacement increment from the two previous increments, the pressure rate and the
node's activation phase.  Training data are the context's own converged steps; the
forward pass, the backward pass and Adam run in the library's HIP kernels, the data
stay in device memory.  The proposed guess is used by the step only when its residual
is below the Newmark predictor's, so Newton always converges to the same tolerance
and the predictor remains the fallback.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .context import CmgError, MechanicsContext
from .hip import DeviceArray


class WarmStartNet:
    HIDDEN = 128

    def __init__(self, ctx: MechanicsContext, seed=0, inc_scale=1e4, pressure_rate_scale=1e-5, phase_scale=1.0 / 0.3):
        self.ctx = ctx
        self.L = ctx.L
        h = ctypes.c_void_p()
        rc = self.L.cmg_mlp_create(ctx.h, self.HIDDEN, seed, ctypes.byref(h))
        if rc != _lib.SUCCESS:
            raise CmgError(rc, "cmg_mlp_create")
        self.h = h
        ctx._add_dependent(self)
        self.n_weights = self.L.cmg_mlp_num_weights(self.HIDDEN)
        self.N = ctx.n_nodes
        self.set_scales(inc_scale, pressure_rate_scale, phase_scale)
        self._X = DeviceArray(np.zeros((self.N, 8), np.float32))
        self._Y = DeviceArray(np.zeros((self.N, 3), np.float32))
        self.samples = []     # (X, T) device arrays of recorded steps

    # -- plumbing ---------------------------------------------------------------
    def _check(self, rc, where):
        if rc != _lib.SUCCESS:
            raise CmgError(rc, where)

    def sync(self):
        self._check(self.L.cmg_synchronize(self.ctx.h), "cmg_synchronize")

    def close(self):
        if getattr(self, "h", None):
            self.L.cmg_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_scales(self, inc_scale, pressure_rate_scale, phase_scale):
        self._check(self.L.cmg_mlp_set_scales(self.h, inc_scale, pressure_rate_scale, phase_scale),
                    "cmg_mlp_set_scales")

    def weights(self):
        w = np.zeros(self.n_weights, np.float32)
        self._check(self.L.cmg_mlp_get_weights(self.h, w.ctypes.data), "cmg_mlp_get_weights")
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._check(self.L.cmg_mlp_set_weights(self.h, w.ctypes.data), "cmg_mlp_set_weights")

    # -- kernels on device arrays -------------------------------------------------
    def forward(self, X: DeviceArray, Y: DeviceArray, rows, bf16=False):
        self._check(self.L.cmg_mlp_forward(self.h, X.ptr, rows, int(bf16), Y.ptr), "cmg_mlp_forward")

    def loss_grad(self, X: DeviceArray, T: DeviceArray, rows, want_grad=False):
        loss = ctypes.c_double()
        g = np.zeros(self.n_weights, np.float32) if want_grad else None
        self._check(self.L.cmg_mlp_loss_grad(self.h, X.ptr, T.ptr, rows, ctypes.byref(loss),
                                             g.ctypes.data if want_grad else None), "cmg_mlp_loss_grad")
        return loss.value, g

    def adam_step(self, lr):
        self._check(self.L.cmg_mlp_adam_step(self.h, ctypes.c_float(lr)), "cmg_mlp_adam_step")

    # -- use inside a Newmark run -----------------------------------------------------
    def features(self, time, pressure_rate, out: DeviceArray | None = None):
        X = out if out is not None else self._X
        self._check(self.L.cmg_mlp_features(self.h, time, pressure_rate, X.ptr), "cmg_mlp_features")
        return X

    def record_before_step(self, time, pressure_rate):
        """features of the step about to run, kept on the device for training"""
        X = DeviceArray(np.zeros((self.N, 8), np.float32))
        self.features(time, pressure_rate, X)
        self.sync()
        self._pending = X

    def record_after_step(self):
        T = DeviceArray(np.zeros((self.N, 3), np.float32))
        self._check(self.L.cmg_mlp_targets(self.h, T.ptr), "cmg_mlp_targets")
        self.sync()
        self.samples.append((self._pending, T))
        self._pending = None

    def propose(self, time, pressure_rate, bf16=False):
        """network guess for the next Newmark step of the context (checked by the step)"""
        X = self.features(time, pressure_rate)
        self.forward(X, self._Y, self.N, bf16)
        self._check(self.L.cmg_mlp_propose_guess(self.h, self._Y.ptr), "cmg_mlp_propose_guess")

    def train(self, epochs=30, lr=1e-3, seed=0, log=None):
        """Adam over the recorded steps (one step's nodes per update), shuffled per epoch."""
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            tot = 0.0
            for k in rng.permutation(len(self.samples)):
                X, T = self.samples[k]
                loss, _ = self.loss_grad(X, T, self.N)
                self.adam_step(lr)
                tot += loss
            hist.append(tot / max(1, len(self.samples)))
            if log:
                log(f"[warm start] epoch {ep}: mean loss {hist[-1]:.4e}")
        self.sync()
        return hist
