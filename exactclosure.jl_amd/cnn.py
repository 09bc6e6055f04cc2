"""Learned closure: a small 3D CNN trained on (filtered velocity, exact
closure) pairs generated by the DNS-aided run itself.

The reference has no learned closure. The network is specified by the
workload: 4 layers of 3x3x3 convolutions with periodic padding, 3 -> 32 ->
32 -> 32 -> 3 channels, GELU between layers, BF16 weights and activations with
FP32 accumulation, trained with Adam. Forward, input-gradient and
weight-gradient kernels are hand-written MFMA implicit GEMMs
(`csrc/ec_cnn.hip`). The exact closure stays the arbiter: `closure_error`
compares a prediction with it, and `LearnedLES` falls back to it when the
prediction misses by more than a tolerance.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _abi
from ._abi import P3, ptrs


def _ptr_array(fields):
    arr = (ctypes.c_void_p * len(fields))()
    for i, t in enumerate(fields):
        arr[i] = t.data_ptr()
    return arr


class CNNClosure:
    """3 -> 32 -> 32 -> 32 -> 3 channel 3D CNN on an LES grid (n a multiple of 16).

    Inputs and targets are lists of samples, each sample three FP64 (n, n, n)
    CUDA tensors (the velocity or closure components, column-major like every
    field here). `in_scale` and `out_scale` map them to order-one network
    units; they are set from training data by `fit` or by hand."""

    def __init__(self, seed=0, ctx=None):
        self.ctx = ctx or _abi.context()
        h = ctypes.c_void_p()
        self.ctx.call("ec_cnn_create", ctypes.c_uint64(seed), ctypes.byref(h))
        self.h = h
        self.np = self.ctx.lib.ec_cnn_num_params(h)
        self.in_scale = 1.0
        self.out_scale = 1.0

    def close(self):
        if self.h:
            self.ctx.lib.ec_cnn_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- parameters ----------------------------------------------------------
    def _io(self, which, get, arr=None):
        buf = np.empty(self.np, dtype=np.float32) if get else np.ascontiguousarray(arr, dtype=np.float32)
        assert buf.size == self.np
        self.ctx.call("ec_cnn_params", self.h, which, 1 if get else 0,
                      buf.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), self.np)
        return buf

    def get_params(self):
        return self._io(0, True)

    def set_params(self, arr):
        self._io(0, False, arr)

    def get_grads(self):
        return self._io(1, True)

    # -- forward / backward / update -----------------------------------------
    def forward(self, inputs, out=None):
        """Run the network on a batch; `out` (optional) receives the predictions
        in physical units (3 FP64 fields per sample).  Inputs stored in another
        precision (BF16 or FP32 pairs) are widened to FP64 per batch."""
        inputs = [_f64(s) for s in inputs]
        self._keep = inputs  # the widened batch must outlive the asynchronous launch
        n = inputs[0][0].shape[0]
        flat = [t for s in inputs for t in s]
        for t in flat:
            assert t.dtype == torch.float64 and t.is_contiguous() and t.shape == (n, n, n)
        outs = None
        if out is not None:
            of = [t for s in out for t in s]
            outs = _ptr_array(of)
        self.ctx.call("ec_cnn_forward", self.h, n, len(inputs), _ptr_array(flat), float(self.in_scale),
                      outs, float(self.out_scale))

    def backward(self, targets):
        """Gradients of the mean squared error (network units) of the last forward; returns the loss."""
        targets = [_f64(s) for s in targets]
        flat = [t for s in targets for t in s]
        loss = ctypes.c_double()
        self.ctx.call("ec_cnn_backward", self.h, _ptr_array(flat), float(self.out_scale), ctypes.byref(loss))
        return loss.value

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.ctx.call("ec_cnn_adam", self.h, lr, beta1, beta2, eps)

    def predict(self, u):
        """Closure forcing predicted for one LES velocity (three FP64 fields)."""
        n = u[0].shape[0]
        m = [torch.empty((n, n, n), dtype=torch.float64, device=u[0].device) for _ in range(3)]
        self.forward([list(u)], out=[m])
        return m

    # -- training ------------------------------------------------------------
    def fit(self, inputs, targets, epochs=20, batch=8, lr=1e-3, seed=0, log=None):
        """Adam on the mean squared closure error; returns the mean loss per epoch."""
        if self.in_scale == 1.0 and self.out_scale == 1.0:
            self.in_scale = 1.0 / _rms(inputs)
            self.out_scale = 1.0 / _rms(targets)
        rng = np.random.default_rng(seed)
        hist = []
        for ep in range(epochs):
            order = rng.permutation(len(inputs))
            tot, cnt = 0.0, 0
            for b0 in range(0, len(order), batch):
                idx = order[b0:b0 + batch]
                self.forward([inputs[i] for i in idx])
                tot += self.backward([targets[i] for i in idx]) * len(idx)
                cnt += len(idx)
                self.adam(lr)
            hist.append(tot / cnt)
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4e}")
        return hist


def _f64(sample):
    return [t if t.dtype == torch.float64 else t.double().contiguous() for t in sample]


def _rms(samples):
    s, c = 0.0, 0
    for smp in samples:
        for t in smp:
            s += float(t.double().square().sum())
            c += t.numel()
    return (s / c) ** 0.5


def closure_error(pred, exact):
    """Relative L2 error of a predicted closure against the exact one (FP64 sums)."""
    num = sum(float((p.double() - e.double()).square().sum()) for p, e in zip(pred, exact))
    den = sum(float(e.double().square().sum()) for e in exact)
    return (num / den) ** 0.5


def generate_pairs(drv, set_index, dt, nsteps, every=1, log=None, dtype=None):
    """Run the DNS-aided driver (forcing form) and collect (P_H V u, exact
    closure forcing m) of LES set `set_index` after every `every` steps,
    stored in `dtype` (default: the LES precision, FP64; BF16 quarters the
    footprint of a large training set, 25 MB per pair at 128^3)."""
    X, Y = [], []
    L = drv.les[set_index]
    for s in range(nsteps):
        drv.step(dt)
        if (s + 1) % every == 0:
            X.append([t.to(dtype) if dtype else t.clone() for t in L.ubar])
            Y.append([t.to(dtype) if dtype else t.clone() for t in L.m])
        if log and (s + 1) % 64 == 0:
            torch.cuda.synchronize()
            log(f"[pairs] step {s + 1}/{nsteps}: {len(X)} pairs")
    return X, Y


class LearnedLES:
    """A posteriori LES of one set of a DNS-aided driver closed by the network,
    advanced through the same RK4 tableau as the driver's LES:
    du = F_H(u) + m_pred(u).  With `fallback_tol` set, a stage whose prediction
    misses the driver's exact closure by more than that relative L2 error uses
    the exact closure instead (the exact closure is the arbiter and fallback).
    `model=None` runs the no-model LES (m = 0)."""

    def __init__(self, drv, set_index, model=None, fallback_tol=None):
        from . import navierstokes as NS

        self.drv, self.si, self.model, self.tol = drv, set_index, model, fallback_tol
        L = drv.les[set_index]
        self.L = L
        self.g = L.g
        n = L.g.n
        dt = drv.dtype
        self.u0 = [t.clone() for t in L.u0]
        self.us = [torch.empty_like(t) for t in self.u0]
        self.acc = [torch.empty_like(t) for t in self.u0]
        self.du = [torch.empty_like(t) for t in self.u0]
        self.ps = NS.PoissonSolver(L.g, dt)
        self.fallbacks = 0
        self.stages = 0
        self.dt_code = _abi.EC_F64 if dt == torch.float64 else _abi.EC_F32

    def stage(self, st, dt):
        """Advance stage st; call after the driver evaluated its stage st."""
        c = self.drv.ctx
        gl = self.g
        u = self.u0 if st == 0 else self.us
        if self.model is not None:
            u64 = [t.double().contiguous() for t in u]
            m = self.model.predict(u64)
            self.stages += 1
            if self.tol is not None and closure_error(m, self.L.m) > self.tol:
                m = self.L.m
                self.fallbacks += 1
            # du = F_H(u) + m in one pass
            c.call("ec_ns_rhs_add", self.dt_code, gl.n, gl.l, self.drv.visc, ptrs(u, P3), ptrs(self.du, P3), 1.0,
                   ptrs(m, P3))
        else:
            c.call("ec_ns_rhs", self.dt_code, gl.n, gl.l, self.drv.visc, ptrs(u, P3), ptrs(self.du, P3), 0)
        c.call("ec_ns_pressure_solve", self.ps.h, ptrs(self.du, P3))
        c.call("ec_ns_rk4_update", self.dt_code, gl.n, gl.l, st, dt, ptrs(self.u0, P3), ptrs(self.us, P3),
               ptrs(self.du, P3), _abi.c_void_p(self.ps.p_ptr), ptrs(self.acc, P3))

    def error(self):
        """Relative L2 distance to the driver's exact LES (= the filtered DNS)."""
        ref = self.L.u0
        num = sum(float(((a.double() - b.double()) ** 2).sum()) for a, b in zip(self.u0, ref))
        den = sum(float((b.double() ** 2).sum()) for b in ref)
        return (num / den) ** 0.5
