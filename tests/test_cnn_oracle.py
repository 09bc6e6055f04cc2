"""CPU pins of the CNN oracle: brute-force convolution, adjoint identities,
finite differences of the FP64 loss gradient, BF16 rounding, GELU and Adam
against independent implementations (torch on the CPU)."""
import numpy as np
import pytest
import torch

import cnn_oracle as C


def test_conv_brute_force():
    rng = np.random.default_rng(0)
    n, ci, co = 3, 2, 2
    X = rng.standard_normal((1, n, n, n, ci))
    W = rng.standard_normal((27, co, ci))
    Y = C.conv(X, W)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                for o in range(co):
                    s = 0.0
                    for dz in (-1, 0, 1):
                        for dy in (-1, 0, 1):
                            for dx in (-1, 0, 1):
                                t = 9 * (dz + 1) + 3 * (dy + 1) + (dx + 1)
                                s += W[t, o] @ X[0, (z + dz) % n, (y + dy) % n, (x + dx) % n]
                    assert abs(Y[0, z, y, x, o] - s) < 1e-12


def test_conv_identity_and_shift():
    rng = np.random.default_rng(1)
    X = rng.standard_normal((2, 4, 5, 6, 3))
    W = np.zeros((27, 3, 3))
    W[13] = np.eye(3)                      # centre tap
    assert np.abs(C.conv(X, W) - X).max() == 0.0
    W = np.zeros((27, 3, 3))
    W[14] = np.eye(3)                      # dx = +1: y[x] = x[x + 1]
    assert np.abs(C.conv(X, W) - np.roll(X, -1, axis=3)).max() == 0.0


def test_adjoint_identities():
    rng = np.random.default_rng(2)
    X = rng.standard_normal((2, 4, 4, 4, 5))
    W = rng.standard_normal((27, 6, 5))
    dY = rng.standard_normal((2, 4, 4, 4, 6))
    lhs = np.sum(C.conv(X, W) * dY)
    assert abs(lhs - np.sum(X * C.conv_dgrad(dY, W))) < 1e-10 * abs(lhs)
    assert abs(lhs - np.sum(W * C.conv_wgrad(X, dY))) < 1e-10 * abs(lhs)


def test_loss_gradient_finite_differences():
    rng = np.random.default_rng(3)
    p = C.init_params(4)
    X = rng.standard_normal((1, 4, 4, 4, 3))
    T = rng.standard_normal((1, 4, 4, 4, 3))
    loss, g, _ = C.loss_and_grad(p, X, T)
    offs = np.cumsum([0] + [27 * co * ci + co for ci, co in zip(C.CIN, C.COUT)])
    # weights and biases of every layer
    idx = [int(rng.integers(offs[l], offs[l + 1])) for l in range(4) for _ in range(4)]
    idx += [int(offs[l + 1] - 1) for l in range(4)]
    for i in idx:
        h = 1e-5
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (C.loss_and_grad(pp, X, T)[0] - C.loss_and_grad(pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[i]) < 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_bf16_rounding_matches_torch():
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.standard_normal(5000) * 10.0 ** rng.integers(-6, 6, 5000),
                        [1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -9, -1 - 2 ** -8, 0.0, 65504.0]])
    ref = torch.tensor(x, dtype=torch.float32).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(C.bf16(x), ref)
    assert C.bf16(1 + 2 ** -8) == 1.0          # tie to even
    assert C.bf16(1 + 3 * 2 ** -9) == 1 + 2 ** -7


def test_gelu_matches_torch():
    x = torch.linspace(-6, 6, 1001, dtype=torch.float64, requires_grad=True)
    y = torch.nn.functional.gelu(x)
    y.sum().backward()
    assert np.abs(C.gelu(x.detach().numpy()) - y.detach().numpy()).max() < 1e-14
    assert np.abs(C.gelu_grad(x.detach().numpy()) - x.grad.numpy()).max() < 1e-14


def test_adam_matches_torch():
    rng = np.random.default_rng(6)
    p0 = rng.standard_normal(50)
    grads = [rng.standard_normal(50) for _ in range(3)]
    tp = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    p, m1, m2 = p0.copy(), np.zeros(50), np.zeros(50)
    for s, g in enumerate(grads, 1):
        tp.grad = torch.tensor(g)
        opt.step()
        p, m1, m2 = C.adam(p, g, m1, m2, s)
    assert np.abs(p - tp.detach().numpy()).max() < 1e-12


def test_emulated_close_to_fp64():
    rng = np.random.default_rng(7)
    p = C.init_params(8)
    X = rng.standard_normal((1, 4, 4, 4, 3))
    T = rng.standard_normal((1, 4, 4, 4, 3))
    l0, g0, o0 = C.loss_and_grad(p, X, T)
    l1, g1, o1 = C.loss_and_grad(p, X, T, emulate=True)
    assert np.linalg.norm(o1 - o0) < 3e-2 * np.linalg.norm(o0)
    assert np.linalg.norm(g1 - g0) < 5e-2 * np.linalg.norm(g0)
