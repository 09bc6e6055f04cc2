"""CPU oracle of the learned-closure CNN (test infrastructure only: imported by
tests/ and never by the product).

The reference has no learned closure (no ML dependency, no network code), so
this restates the network the workload specifies: 4 layers of 3x3x3
cross-correlations with periodic padding, 3 -> 32 -> 32 -> 32 -> 3 channels,
exact (erf) GELU after the first three layers, mean-squared-error loss, Adam.
Parameters are one flat array, layer by layer: W[27][cout][cin] with tap
t = 9 (dz + 1) + 3 (dy + 1) + (dx + 1) and y[p] = sum_t W[t] x[p + (dx, dy, dz)],
then b[cout].

Arrays are (sample, z, y, x, channel). Everything runs in FP64. With
`emulate=True` the values are rounded where the GPU rounds them: BF16 for
weights, inputs, activations, pre-activations and the gradients that feed the
matrix cores; FP32 for the accumulated layer outputs. Without it, the plain
FP64 network is evaluated, which finite differences pin.
"""
from __future__ import annotations

import numpy as np
from scipy.special import erf

CIN = (3, 32, 32, 32)
COUT = (32, 32, 32, 3)


def bf16(x):
    """Round to bfloat16 (nearest, ties to even), returned as FP64."""
    x32 = np.asarray(x, dtype=np.float32)
    b = x32.view(np.uint32).astype(np.uint64)
    b = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return (b & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def f32(x):
    return np.asarray(x, dtype=np.float32).astype(np.float64)


def num_params():
    return sum(27 * co * ci + co for ci, co in zip(CIN, COUT))


def unpack(params):
    """[(W[27, cout, cin], b[cout])] of the four layers."""
    out, o = [], 0
    for ci, co in zip(CIN, COUT):
        W = np.asarray(params[o:o + 27 * co * ci], dtype=np.float64).reshape(27, co, ci)
        o += 27 * co * ci
        b = np.asarray(params[o:o + co], dtype=np.float64)
        o += co
        out.append((W, b))
    return out


def pack(layers):
    return np.concatenate([np.concatenate([W.ravel(), b.ravel()]) for W, b in layers])


def taps():
    for t in range(27):
        yield t, t % 3 - 1, (t // 3) % 3 - 1, t // 9 - 1


def shift(X, dx, dy, dz):
    """Xs[p] = X[p + (dx, dy, dz)], periodic."""
    return np.roll(X, (-dz, -dy, -dx), axis=(1, 2, 3))


def conv(X, W):
    """y[p, co] = sum_t sum_ci X[p + o_t, ci] W[t, co, ci]."""
    Y = np.zeros(X.shape[:-1] + (W.shape[1],))
    for t, dx, dy, dz in taps():
        Y += shift(X, dx, dy, dz) @ W[t].T
    return Y


def conv_dgrad(dY, W):
    """dX[q, ci] = sum_t sum_co dY[q - o_t, co] W[t, co, ci]."""
    dX = np.zeros(dY.shape[:-1] + (W.shape[2],))
    for t, dx, dy, dz in taps():
        dX += np.roll(dY @ W[t], (dz, dy, dx), axis=(1, 2, 3))
    return dX


def conv_wgrad(X, dY):
    """dW[t, co, ci] = sum_p dY[p, co] X[p + o_t, ci]."""
    co, ci = dY.shape[-1], X.shape[-1]
    dW = np.zeros((27, co, ci))
    dYf = dY.reshape(-1, co)
    for t, dx, dy, dz in taps():
        dW[t] = dYf.T @ shift(X, dx, dy, dz).reshape(-1, ci)
    return dW


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / np.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / np.sqrt(2.0 * np.pi)


def fields_to_x(samples, scale=1.0):
    """Samples of three column-major (F-ordered [i, j, k]) fields -> (b, z, y, x, 3)."""
    return np.stack([np.stack([np.asarray(f).transpose(2, 1, 0) for f in s], axis=-1) for s in samples]) * scale


def x_to_fields(X, inv_scale=1.0):
    return [[np.asfortranarray(X[b, ..., c].transpose(2, 1, 0)) * inv_scale for c in range(X.shape[-1])]
            for b in range(X.shape[0])]


def forward(params, X, emulate=False):
    """Network output (FP64) and the cache the backward pass needs."""
    r = bf16 if emulate else (lambda a: a)
    r32 = f32 if emulate else (lambda a: a)
    layers = unpack(params)
    a = r(X)
    acts, pres = [a], []
    for l in range(3):
        W, b = layers[l]
        v = r32(conv(a, r(W)) + b)
        pres.append(r(v))
        a = r(gelu(v))
        acts.append(a)
    W, b = layers[3]
    out = r32(conv(a, r(W)) + b)
    return out, {"acts": acts, "pres": pres, "emulate": emulate}


def loss_and_grad(params, X, T, emulate=False):
    """Mean squared error over all points and channels, and its gradient with
    respect to the flat parameter array."""
    r = bf16 if emulate else (lambda a: a)
    r32 = f32 if emulate else (lambda a: a)
    layers = unpack(params)
    out, cache = forward(params, X, emulate)
    acts, pres = cache["acts"], cache["pres"]
    d = out - T
    loss = float(np.mean(d * d))
    dY = r(r32(2.0 * d / d.size))
    grads = [None] * 4
    for l in (3, 2, 1, 0):
        W, _ = layers[l]
        grads[l] = (conv_wgrad(acts[l], dY), dY.reshape(-1, dY.shape[-1]).sum(axis=0))
        if l > 0:
            dX = r32(conv_dgrad(dY, r(W)))
            dY = r(dX * gelu_grad(pres[l - 1]))
    return loss, pack(grads), out


def adam(params, grad, m1, m2, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected), FP64; returns new params, m1, m2."""
    m1 = b1 * m1 + (1 - b1) * grad
    m2 = b2 * m2 + (1 - b2) * grad * grad
    c1, c2 = 1 - b1 ** step, 1 - b2 ** step
    return params - lr * (m1 / c1) / (np.sqrt(m2 / c2) + eps), m1, m2


def init_params(seed):
    """The default initialisation's distribution (uniform within
    1/sqrt(27 cin)); not the GPU's exact stream."""
    rng = np.random.default_rng(seed)
    layers = []
    for ci, co in zip(CIN, COUT):
        bd = 1.0 / np.sqrt(27.0 * ci)
        layers.append((rng.uniform(-bd, bd, (27, co, ci)), rng.uniform(-bd, bd, co)))
    return pack(layers)
