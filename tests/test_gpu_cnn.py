"""GPU parity of the learned-closure CNN (MFMA forward, input-gradient and
weight-gradient kernels, Adam) against the CPU oracle, and what the trained
network gains over the LES without a model."""
import numpy as np
import pytest
import torch

import cnn_oracle as C
from test_gpu_ns import dev, host

pytestmark = pytest.mark.gpu


def rl2(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def _setup(gpu, n=16, batch=2, seed=0):
    rng = np.random.default_rng(seed)
    net = gpu.cnn.CNNClosure(seed=1)
    p = C.init_params(seed + 11).astype(np.float32)
    net.set_params(p)
    u = [[np.asfortranarray(rng.standard_normal((n, n, n))) for _ in range(3)] for _ in range(batch)]
    t = [[np.asfortranarray(0.3 * rng.standard_normal((n, n, n))) for _ in range(3)] for _ in range(batch)]
    return net, p.astype(np.float64), u, t


@pytest.mark.parametrize("n,batch", [(16, 2), (32, 1)])
def test_cnn_forward(gpu, n, batch):
    net, p, u, _ = _setup(gpu, n, batch)
    ud = [[dev(x) for x in s] for s in u]
    out = [[torch.empty_like(ud[0][0]) for _ in range(3)] for _ in range(batch)]
    net.forward(ud, out=out)
    ref, _ = C.forward(p, C.fields_to_x(u), emulate=True)
    got = C.fields_to_x([[host(x) for x in s] for s in out])
    assert rl2(got, ref) < 1e-3
    ref64, _ = C.forward(p, C.fields_to_x(u))
    assert rl2(got, ref64) < 3e-2      # BF16 against the plain FP64 network


def test_cnn_gradients(gpu):
    """Loss and all parameter gradients against the oracle with the GPU's
    rounding points (BF16 operands, FP32 accumulation), and against FP64."""
    net, p, u, t = _setup(gpu, 16, 2, seed=3)
    ud = [[dev(x) for x in s] for s in u]
    td = [[dev(x) for x in s] for s in t]
    net.forward(ud)
    loss = net.backward(td)
    g = net.get_grads().astype(np.float64)
    X, T = C.fields_to_x(u), C.fields_to_x(t)
    lref, gref, _ = C.loss_and_grad(p, X, T, emulate=True)
    assert abs(loss - lref) < 1e-4 * lref
    offs = np.cumsum([0] + [27 * co * ci + co for ci, co in zip(C.CIN, C.COUT)])
    for l in range(4):
        sl = slice(offs[l], offs[l + 1])
        assert rl2(g[sl], gref[sl]) < 2e-3, (l, rl2(g[sl], gref[sl]))
    l64, g64, _ = C.loss_and_grad(p, X, T)
    assert rl2(g, g64) < 5e-2


def test_cnn_adam_step(gpu):
    net, p, u, t = _setup(gpu, 16, 1, seed=5)
    net.forward([[dev(x) for x in s] for s in u])
    net.backward([[dev(x) for x in s] for s in t])
    g = net.get_grads().astype(np.float64)
    net.adam(1e-3)
    got = net.get_params().astype(np.float64)
    ref, _, _ = C.adam(p, g, np.zeros_like(p), np.zeros_like(p), 1)
    assert np.abs(got - ref).max() < 1e-6


def test_cnn_rejects_bad_shapes(gpu):
    net, _, _, _ = _setup(gpu, 16, 1)
    x = [[torch.zeros((12, 12, 12), dtype=torch.float64, device="cuda") for _ in range(3)]]
    with pytest.raises(RuntimeError):
        net.forward(x)


def test_cnn_learns_the_closure(gpu):
    """Pairs (P_H V u, exact closure forcing) from a 64^3 DNS-aided run with the
    16^3 volume-average LES; the trained network must predict held-out closures
    much better than the LES without a model (relative error 1), and its
    a posteriori LES must track the filtered DNS better than the no-model LES."""
    NS, D, CN = gpu.navierstokes, gpu.dnsaid, gpu.cnn
    n, visc = 64, 1.0 / 1000
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float64)
    u = NS.randomfield(g, ps, seed=7)
    dt = NS.propose_timestep(u, visc, 0.4)
    drv = D.DNSAidedLES(g, visc, [(16, "va")], dtype=torch.float64)
    drv.set_state(u)
    X, Y = CN.generate_pairs(drv, 0, dt, 48)
    net = CN.CNNClosure(seed=2)
    hist = net.fit(X[:40], Y[:40], epochs=60, batch=8, lr=2e-3, seed=0)
    assert hist[-1] < 0.5 * hist[0]
    errs = [CN.closure_error(net.predict(x), y) for x, y in zip(X[40:], Y[40:])]
    assert max(errs) < 0.7, errs
    # a posteriori: no model vs learned closure, from the same filtered state
    plain = CN.LearnedLES(drv, 0, None)
    learned = CN.LearnedLES(drv, 0, net, fallback_tol=None)
    drv.stage_hooks = [plain, learned]
    drv.run(dt, 10)
    drv.stage_hooks = []
    assert learned.error() < plain.error(), (learned.error(), plain.error())


def test_cnn_exact_closure_fallback(gpu):
    """The exact-closure fallback of the a posteriori LES (the arbiter of the
    learned closure): with a tolerance below every prediction's error each
    stage falls back and the LES follows the driver's exact LES (the filtered
    DNS) to FP64 rounding; with an infinite tolerance no stage does and the LES
    equals the pure network LES bit for bit.  Pairs stored in BF16, as the
    large training set is, train the network."""
    NS, D, CN = gpu.navierstokes, gpu.dnsaid, gpu.cnn
    n, visc = 64, 1.0 / 1000
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float64)
    u = NS.randomfield(g, ps, seed=9)
    dt = NS.propose_timestep(u, visc, 0.4)
    drv = D.DNSAidedLES(g, visc, [(16, "va")], dtype=torch.float64)
    drv.set_state(u)
    X, Y = CN.generate_pairs(drv, 0, dt, 16, dtype=torch.bfloat16)
    assert X[0][0].dtype == torch.bfloat16 and Y[0][0].dtype == torch.bfloat16
    net = CN.CNNClosure(seed=3)
    hist = net.fit(X, Y, epochs=10, batch=4, lr=2e-3, seed=0)
    assert hist[-1] < hist[0]
    learned = CN.LearnedLES(drv, 0, net, fallback_tol=None)
    every = CN.LearnedLES(drv, 0, net, fallback_tol=0.0)
    never = CN.LearnedLES(drv, 0, net, fallback_tol=float("inf"))
    drv.stage_hooks = [learned, every, never]
    drv.run(dt, 5)
    drv.stage_hooks = []
    print(f"\n[fallback] learned {learned.error():.3e}, always {every.error():.3e}, never {never.error():.3e}")
    assert every.fallbacks == every.stages == 20
    assert every.error() < 1e-12
    assert never.fallbacks == 0 and never.stages == 20
    assert never.error() == learned.error() and learned.error() > 1e-6
