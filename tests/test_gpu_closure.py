"""GPU parity of the filters, the exact-closure fluxes and the RK4 DNS-aided
LES driver against the FP64 CPU oracle, plus the exactness invariant
(navierstokes.jl:33: the classic+flux+div LES equals the filtered DNS to
rounding) and the FP32-pipeline closure error against FP64 (gate 1e-5)."""
import numpy as np
import pytest
import torch

from test_gpu_ns import P, dev, host, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,c,kind", [(8, 2, "va"), (12, 3, "va"), (16, 4, "va"), (15, 5, "va"),
                                      (16, 4, "fa"), (12, 3, "fa")])
def test_closure_quantities(gpu, O, n, c, kind):
    NS = gpu.navierstokes
    g = NS.Grid(1.0, n)
    visc = 0.01
    u = O.randomfield(n, seed=n + c)
    p = np.asfortranarray(np.random.default_rng(c).standard_normal((n, n, n)))
    ref = O.closure_quantities(u, p, visc, n // c, c, kind)
    cs = NS.ClosureSet(g, n // c, kind)
    NS.closure_([cs], [NS.Field(g, dev(x)) for x in u], NS.Field(g, dev(p)), visc)
    for a in range(3):
        assert rel(host(cs.ubar[a]), ref[f"ubar{a + 1}"]) < 1e-12
    for m, ab in enumerate(O.NONSYM_ORDER):
        assert rel(host(cs.flux[m]), ref["rDHi" + ab]) < 1e-12


@pytest.mark.parametrize("n,sets", [(16, [(4, "va")]), (15, [(5, "va"), (3, "fa")]), (12, [(2, "va"), (6, "fa")]),
                                    (32, [(16, "va"), (8, "va"), (4, "va"), (4, "fa")]),
                                    (48, [(3, "va"), (3, "fa"), (16, "fa"), (2, "va")]),
                                    (30, [(5, "fa"), (5, "va")]),
                                    (64, [(16, "va"), (8, "va"), (4, "va"), (4, "fa")])])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_filter_vel(gpu, O, n, sets, prec):
    """V(u) and V(k), k = du - grad p, against the oracle's separable filters;
    and V(k) = -div_H^ns (S r) (the flux form) on the same data."""
    NS = gpu.navierstokes
    dt = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-12 if prec == "f64" else 2e-6
    rng = np.random.default_rng(n)
    v = [np.asfortranarray(rng.standard_normal((n, n, n))) for _ in range(3)]
    p = np.asfortranarray(rng.standard_normal((n, n, n)))
    k = O.pressuregradient([x.copy(order="F") for x in v], p)
    ctx = gpu._abi.context()
    vd = [dev(x, dt) for x in v]
    outs = []
    arr = (gpu._abi.VelSet * len(sets))()
    for i, (c, kind) in enumerate(sets):
        o = [torch.empty((n // c,) * 3, dtype=torch.float64, device="cuda") for _ in range(3)]
        outs.append(o)
        arr[i].nbar = n // c
        arr[i].kind = 0 if kind == "va" else 1
        for a in range(3):
            arr[i].out[a] = o[a].data_ptr()
    P3 = gpu._abi.P3
    code = 1 if prec == "f64" else 0
    # n = 64: z chunks of 16 planes, so windows meet chunk boundaries
    ctx.call("ec_filter_vel_set_tiles", 16, 16 if n == 64 else 64)
    try:
        for with_p in (False, True):
            ctx.call("ec_ns_filter_vel", code, n, 1.0, P3(*[x.data_ptr() for x in vd]),
                     dev(p, dt).data_ptr() if with_p else None, len(sets), arr)
            for (c, kind), o in zip(sets, outs):
                F = O.closure_filters(c, kind)
                for a in range(3):
                    ref = O.filter_coarse((k if with_p else v)[a], n // c, *F[f"ubar{a + 1}"][1])
                    assert rel(host(o[a]), ref) < tol, (c, kind, a, with_p)
    finally:
        ctx.call("ec_filter_vel_set_tiles", 16, 256)
    # commutation: V(k) for k = -div r equals -div^ns_H (S r) for the volume average
    if prec == "f64":
        u = O.randomfield(n, seed=1)
        for c, kind in sets:
            if kind != "va":
                continue
            q = O.closure_quantities(u, p, 0.01, n // c, c, "va")
            du = O.rhs(u, 0.01)
            kk = O.pressuregradient([x.copy(order="F") for x in du], p)
            lhs = [O.filter_coarse(kk[a], n // c, *O.closure_filters(c, "va")[f"ubar{a + 1}"][1]) for a in range(3)]
            rhs = O.tensordivergence_nonsym([q["rDHi" + ab] for ab in O.NONSYM_ORDER])
            for a in range(3):
                assert rel(lhs[a], rhs[a]) < 1e-11


def test_filter_coarse_gaussian(gpu, O):
    """lazycoarsegrain(lazyfilter(u, GDH)) with the reference's Gaussian-composed
    kernel at odd compression (get_kernels, src/N.jl:1008-1019)."""
    NS = gpu.navierstokes
    n, c = 15, 3
    g, gb = NS.Grid(1.0, n), NS.Grid(1.0, n // c)
    FD = NS.gaussian(g, 2 * NS.spacing(gb), 2)
    FDH = NS.composekernel(NS.tophat(g, c), FD)
    u = O.randomfield(n, seed=3)
    # the oracle builds its kernel with its own gaussian / composekernel / tophat
    OFDH = O.composekernel(O.tophat(c), O.gaussian(1.0 / n, 2.0 * c / n, 2))
    ref = O.coarsegrain(O.applyfilter(u[0], OFDH, OFDH, OFDH), n // c, (True, False, False))
    ub = NS.Field(gb)
    NS.applyfilter_coarsegrain(ub, NS.Field(g, dev(u[0])), NS.KernelProduct([FDH] * 3), (True, False, False))
    assert rel(ub.numpy(), ref) < 1e-12
    pt = NS.Field(gb)
    NS.coarsegrain_(pt, NS.Field(g, dev(u[1])), (False, True, False))
    assert rel(pt.numpy(), O.coarsegrain(u[1], n // c, (False, True, False))) < 1e-15


@pytest.mark.parametrize("closure", ["flux", "forcing"])
@pytest.mark.parametrize("n,sets", [(16, [(4, "va")]), (24, [(3, "va"), (8, "va")]), (16, [(4, "fa"), (2, "va")]),
                                    (32, [(16, "va"), (8, "va"), (4, "va"), (4, "fa")])])
def test_dnsaid_rk4_step_fp64(gpu, O, n, sets, closure):
    NS, D = gpu.navierstokes, gpu.dnsaid
    visc, dt = 1.0 / 500, 2e-3
    g = NS.Grid(1.0, n)
    u = O.randomfield(n, seed=5)
    drv = D.DNSAidedLES(g, visc, [(n // c, k) for c, k in sets], dtype=torch.float64, closure=closure)
    drv.set_state([dev(x) for x in u])
    drv.run(dt, 2)
    u_cpu, les_cpu, tau_cpu = O.dnsaid_rk4([x.copy(order="F") for x in u], sets, visc, dt, 2, closure=closure)
    for a in range(3):
        assert rel(host(drv.u0[a]), u_cpu[a]) < 1e-11
    for si, (L, lc) in enumerate(zip(drv.les, les_cpu)):
        for a in range(3):
            assert rel(host(L.u0[a]), lc[a]) < 1e-10
        # closure of the last stage evaluated: 9 flux components or 3 forcing components
        terms = L.cset.tau if closure == "flux" else L.m
        for m in range(len(terms)):
            assert rel(host(terms[m]), tau_cpu[si][m]) < 1e-9
    errs = drv.les_errors()
    for (c, k), e in zip(sets, errs):
        if k == "va" or closure == "forcing":
            assert e < 1e-12, (c, e)  # exactness invariant (the forcing form is exact for any filter)


def test_exactness_many_steps_fp64(gpu, O):
    NS, D = gpu.navierstokes, gpu.dnsaid
    n, visc = 32, 1.0 / 2000
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float64)
    u = NS.randomfield(g, ps, seed=123)
    dt = NS.propose_timestep(u, visc, 0.4)
    drv = D.DNSAidedLES(g, visc, [(8, "va"), (4, "va"), (16, "va")], dtype=torch.float64, closure="flux")
    drv.set_state(u)
    drv.run(dt, 10)
    assert max(drv.les_errors()) < 1e-12
    assert drv.divergence_norm() < 1e-13


@pytest.mark.parametrize("closure", ["flux", "forcing"])
def test_fp32_pipeline_closure_error(gpu, O, closure):
    """Closure from the same state: FP32 fields + FP64 commutator vs all-FP64.
    Gate: relative L2 <= 1e-5 over all components and LES points."""
    NS, D = gpu.navierstokes, gpu.dnsaid
    n, visc = 64, 1.0 / 4000
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float64)
    u64 = NS.randomfield(g, ps, seed=123)
    u32 = [x.data.float() for x in u64]
    res = {}
    for dt, u in ((torch.float32, u32), (torch.float64, [x.double() for x in u32])):
        drv = D.DNSAidedLES(g, visc, [(16, "va"), (8, "va"), (4, "va"), (4, "fa")], dtype=dt, closure=closure)
        drv.set_state(u)
        res[dt] = drv.closure_terms()
    for t32, t64 in zip(res[torch.float32], res[torch.float64]):
        num = sum(float(((a - b) ** 2).sum()) for a, b in zip(t32, t64))
        den = sum(float((b ** 2).sum()) for b in t64)
        assert (num / den) ** 0.5 < 1e-5


@pytest.mark.parametrize("n,prec,sets", [(32, "f64", [(16, 0), (8, 0), (4, 0), (4, 1)]),
                                         (64, "f32", [(16, 0), (8, 0), (4, 0), (4, 1)]),
                                         (64, "f64", [(16, 0), (8, 0), (4, 0), (4, 1)]),
                                         (48, "f32", [(3, 0), (3, 1), (16, 1), (2, 0)])])
@pytest.mark.parametrize("stage", [0, 1, 2, 3])
def test_filter_vel_rk4_fused(gpu, n, prec, sets, stage):
    """ec_ns_filter_vel_rk4 equals ec_ns_filter_vel followed by ec_ns_rk4_update
    on the same fields: the filter outputs bit for bit, the stage update to
    rounding (fused for the volume-16/8/4 + face-4 list, with 16-plane chunks at
    n = 64; the other list takes the two calls)."""
    dt = torch.float64 if prec == "f64" else torch.float32
    code = 1 if prec == "f64" else 0
    gen = torch.Generator(device="cuda").manual_seed(10 * n + stage)
    f = lambda: torch.randn((n, n, n), device="cuda", dtype=dt, generator=gen)  # noqa: E731
    du, u0, us, acc = [f() for _ in range(3)], [f() for _ in range(3)], [f() for _ in range(3)], [f() for _ in range(3)]
    p = f()
    copies = lambda ts: [t.clone() for t in ts]  # noqa: E731
    state_a = [copies(u0), copies(us), copies(acc)]
    state_b = [copies(u0), copies(us), copies(acc)]

    def outsets():
        outs = []
        arr = (gpu._abi.VelSet * len(sets))()
        for i, (c, kind) in enumerate(sets):
            o = [torch.empty((n // c,) * 3, dtype=torch.float64, device="cuda") for _ in range(3)]
            outs.append(o)
            arr[i].nbar = n // c
            arr[i].kind = kind
            for a in range(3):
                arr[i].out[a] = o[a].data_ptr()
        return outs, arr

    oa, arra = outsets()
    ob, arrb = outsets()
    ctx = gpu._abi.context()
    step = 1e-3
    ctx.call("ec_filter_vel_set_tiles", 16, 16 if n == 64 else 64)
    try:
        ctx.call("ec_ns_filter_vel", code, n, 1.0, P(du, 3), p.data_ptr(), len(sets), arra)
        ctx.call("ec_ns_rk4_update", code, n, 1.0, stage, step, P(state_a[0], 3), P(state_a[1], 3), P(du, 3),
                 p.data_ptr(), P(state_a[2], 3))
        ctx.call("ec_ns_filter_vel_rk4", code, n, 1.0, P(du, 3), gpu._abi.c_void_p(p.data_ptr()), len(sets), arrb,
                 stage, step, P(state_b[0], 3), P(state_b[1], 3), P(state_b[2], 3))
    finally:
        ctx.call("ec_filter_vel_set_tiles", 16, 256)
    for x, y in zip(oa, ob):
        for a in range(3):
            assert torch.equal(x[a], y[a])
    tol = 1e-14 if prec == "f64" else 1e-6
    for xs, ys in zip(state_a, state_b):
        for x, y in zip(xs, ys):
            assert float((x - y).abs().max()) <= tol * float(x.abs().max())



@pytest.mark.parametrize("n", [512, 1024])
@pytest.mark.parametrize("stage", [1, 3])
def test_filter_vel_rk4_fused_large(gpu, n, stage):
    """Full-size FP32 grids with the benchmark's set list and its default
    256-plane chunks (every tile edge, chunk edge and emit of the launch): the
    fused filter + RK4 stage against ec_ns_filter_vel followed by
    ec_ns_rk4_update on the same fields, the filter outputs bit for bit and
    the stage state to FP32 rounding (stage 1 reads and writes acc, stage 3
    writes u0)."""
    sets = [(16, 0), (8, 0), (4, 0), (4, 1)]
    gen = torch.Generator(device="cuda").manual_seed(n + stage)
    f = lambda: torch.randn((n, n, n), device="cuda", generator=gen)  # noqa: E731
    du = [f() for _ in range(3)]
    p = f()
    state_a = [[f() for _ in range(3)] for _ in range(3)]  # u0, us, acc
    state_b = [[t.clone() for t in xs] for xs in state_a]

    def outsets():
        outs = []
        arr = (gpu._abi.VelSet * len(sets))()
        for i, (c, kind) in enumerate(sets):
            o = [torch.empty((n // c,) * 3, dtype=torch.float64, device="cuda") for _ in range(3)]
            outs.append(o)
            arr[i].nbar = n // c
            arr[i].kind = kind
            for a in range(3):
                arr[i].out[a] = o[a].data_ptr()
        return outs, arr

    oa, arra = outsets()
    ob, arrb = outsets()
    ctx = gpu._abi.context()
    step = 1e-4
    ctx.call("ec_ns_filter_vel", 0, n, 1.0, P(du, 3), p.data_ptr(), len(sets), arra)
    ctx.call("ec_ns_rk4_update", 0, n, 1.0, stage, step, P(state_a[0], 3), P(state_a[1], 3), P(du, 3),
             p.data_ptr(), P(state_a[2], 3))
    ctx.call("ec_ns_filter_vel_rk4", 0, n, 1.0, P(du, 3), gpu._abi.c_void_p(p.data_ptr()), len(sets), arrb,
             stage, step, P(state_b[0], 3), P(state_b[1], 3), P(state_b[2], 3))
    torch.cuda.synchronize()
    for x, y in zip(oa, ob):
        for a in range(3):
            assert torch.equal(x[a], y[a])
    for xs, ys in zip(state_a, state_b):
        for x, y in zip(xs, ys):
            assert float((x - y).abs().max()) <= 1e-6 * float(x.abs().max())
    del du, p, state_a, state_b, oa, ob
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n,prec,sets", [(64, "f64", [(16, 0), (8, 0), (4, 0), (4, 1)]),
                                         (48, "f32", [(3, 0), (3, 1), (16, 1), (2, 0)]),
                                         (512, "f32", [(16, 0), (8, 0), (4, 0), (4, 1)]),
                                         (1024, "f32", [(16, 0), (8, 0), (4, 0), (4, 1)])])
def test_filter_vel_rk4_states_fused(gpu, n, prec, sets):
    """ec_ns_filter_vel_rk4_states over the four stages of a step equals
    ec_ns_filter_vel followed by ec_ns_rk4_update_states on the same fields:
    the filter outputs bit for bit, the stage states u1, u2, u3 and the new
    u0 to rounding (FMA contraction may differ between the kernels; fused
    for the volume-16/8/4 + face-4 list, at the benchmark's 256-plane chunks
    for n >= 512; the other list takes the two calls).  Each stage gets its
    own du and p."""
    dt = torch.float64 if prec == "f64" else torch.float32
    code = 1 if prec == "f64" else 0
    gen = torch.Generator(device="cuda").manual_seed(3 * n + 1)
    f = lambda: torch.randn((n, n, n), device="cuda", dtype=dt, generator=gen)  # noqa: E731
    u0 = [f() for _ in range(3)]
    ua, ub = [x.clone() for x in u0], [x.clone() for x in u0]
    del u0
    sta = [[torch.empty_like(ua[0]) for _ in range(3)] for _ in range(3)]
    stb = [[torch.empty_like(ua[0]) for _ in range(3)] for _ in range(3)]
    pa, pb = P([x for s_ in sta for x in s_], 9), P([x for s_ in stb for x in s_], 9)

    def outsets():
        outs = []
        arr = (gpu._abi.VelSet * len(sets))()
        for i, (c, kind) in enumerate(sets):
            o = [torch.empty((n // c,) * 3, dtype=torch.float64, device="cuda") for _ in range(3)]
            outs.append(o)
            arr[i].nbar = n // c
            arr[i].kind = kind
            for a in range(3):
                arr[i].out[a] = o[a].data_ptr()
        return outs, arr

    ctx = gpu._abi.context()
    step = 1e-3 if n <= 64 else 1e-4
    small = n <= 64
    tol = 1e-14 if prec == "f64" else 1e-6
    if small:
        ctx.call("ec_filter_vel_set_tiles", 16, 16)
    try:
        for st in range(4):
            du = [f() for _ in range(3)]
            p = f()
            oa, arra = outsets()
            ob, arrb = outsets()
            ctx.call("ec_ns_filter_vel", code, n, 1.0, P(du, 3), p.data_ptr(), len(sets), arra)
            ctx.call("ec_ns_rk4_update_states", code, n, 1.0, st, step, P(ua, 3), pa, P(du, 3), p.data_ptr())
            ctx.call("ec_ns_filter_vel_rk4_states", code, n, 1.0, P(du, 3), gpu._abi.c_void_p(p.data_ptr()),
                     len(sets), arrb, st, step, P(ub, 3), pb)
            torch.cuda.synchronize()
            for x, y in zip(oa, ob):
                for a in range(3):
                    assert torch.equal(x[a], y[a]), (st, a)
            if st < 3:
                for a in range(3):
                    x, y = sta[st][a], stb[st][a]
                    assert float((x - y).abs().max()) <= tol * float(x.abs().max()), (st, a)
            del du, p, oa, ob
    finally:
        if small:
            ctx.call("ec_filter_vel_set_tiles", 16, 256)
    for x, y in zip(ua, ub):
        assert float((x - y).abs().max()) <= tol * float(x.abs().max())
    del ua, ub, sta, stb
    torch.cuda.empty_cache()


def test_stream_after_orders_and_refuses(gpu):
    """ec_stream_after: work queued on one context after the call runs after
    what the other context had queued (a copy on the second stream reads the
    result of a long kernel sequence on the first), and the call refuses a
    context paired with itself or with a context on the same stream."""
    abi = gpu._abi
    a = abi.context()
    s2 = torch.cuda.Stream()
    b = abi.Context(0, s2.cuda_stream)
    c = abi.Context(0, a.stream)
    x = torch.zeros(1 << 24, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    for _ in range(50):  # many dependent passes on the first stream
        a.call("ec_axpby", abi.EC_F64, x.numel(), 0.0, abi.c_void_p(x.data_ptr()), 1.0, abi.c_void_p(x.data_ptr()))
    a.call("ec_axpby", abi.EC_F64, x.numel(), 0.0, abi.c_void_p(x.data_ptr()), 1.0, abi.c_void_p(x.data_ptr()))
    ones = torch.ones_like(x)
    torch.cuda.synchronize()
    a.call("ec_axpby", abi.EC_F64, x.numel(), 1.0, abi.c_void_p(ones.data_ptr()), 1.0, abi.c_void_p(x.data_ptr()))
    b.call("ec_stream_after", a.h)
    b.call("ec_d2d", abi.c_void_p(y.data_ptr()), abi.c_void_p(x.data_ptr()), x.numel() * 8)
    b.call("ec_sync")
    assert bool((y == 1.0).all())
    for bad in (a, c):
        with pytest.raises(abi.ECError):
            bad.call("ec_stream_after", a.h)
    del x, y, ones


@pytest.mark.parametrize("n,sets,prec", [(256, [(16, "va"), (32, "va"), (64, "va"), (64, "fa")], "f32"),
                                         (32, [(8, "va"), (4, "fa")], "f64")])
def test_les_stream_same_bits(gpu, O, n, sets, prec):
    """The LES stage work on a second stream (ec_stream_after orders it after
    its stage's filter and the next filter after it) gives the one-stream
    step's values bit for bit, eagerly and as a captured graph whose LES work
    is a parallel branch (replayed with eager, synchronising work between the
    replays), for the DNS state and every LES state."""
    NS, D = gpu.navierstokes, gpu.dnsaid
    dt_t = torch.float32 if prec == "f32" else torch.float64
    g = NS.Grid(1.0, n)
    u = O.randomfield(n, seed=n + 5) if n <= 64 else None
    if u is None:
        ps = NS.PoissonSolver(g, torch.float64)
        u = [x.data for x in NS.randomfield(g, ps, "peak", 5.0, 0.5, seed=11)]
    else:
        u = [dev(x) for x in u]
    one = D.DNSAidedLES(g, 1.0 / 2000, sets, dtype=dt_t, les_stream=False)
    two = D.DNSAidedLES(g, 1.0 / 2000, sets, dtype=dt_t, les_stream=True)
    assert two.les_ctx is not None and one.les_ctx is None
    one.set_state([x.to(dt_t) for x in u])
    two.set_state([x.to(dt_t) for x in u])
    dt = 2e-3 if n <= 64 else 5e-4

    def same():
        torch.cuda.synchronize()
        for a in range(3):
            assert torch.equal(one.u0[a], two.u0[a])
            for L1, L2 in zip(one.les, two.les):
                assert torch.equal(L1.u0[a], L2.u0[a])
                assert torch.equal(L1.v0[a], L2.v0[a])

    for _ in range(2):
        one.step(dt)
        two.step(dt)
        same()
    two.capture(dt)
    for _ in range(3):
        one.step(dt)
        two.step(dt)
        two.divergence_norm()  # eager, synchronising work between the replays
        same()
    two.release_graphs()


@pytest.mark.parametrize("n,sets,prec", [(256, [(16, "va"), (32, "va"), (64, "va"), (64, "fa")], "f32"),
                                         (32, [(8, "va"), (4, "fa")], "f64")])
def test_captured_step_replays_the_eager_step(gpu, O, n, sets, prec):
    """A whole DNS-aided RK4 step (DNS kernels, every LES grid's forcing,
    projection and update) captured into one HIP graph and replayed four
    times, with eager work and synchronisation between the replays (a
    reduction, an eager step of another driver), equals the eager step bit
    for bit after every replay, for the DNS and every LES state."""
    NS, D = gpu.navierstokes, gpu.dnsaid
    dt_t = torch.float32 if prec == "f32" else torch.float64
    g = NS.Grid(1.0, n)
    u = O.randomfield(n, seed=n + 3) if n <= 64 else None
    if u is None:
        ps = NS.PoissonSolver(g, torch.float64)
        u = [x.data for x in NS.randomfield(g, ps, "peak", 5.0, 0.5, seed=9)]
    else:
        u = [dev(x) for x in u]
    eager = D.DNSAidedLES(g, 1.0 / 2000, sets, dtype=dt_t)
    graph = D.DNSAidedLES(g, 1.0 / 2000, sets, dtype=dt_t)
    eager.set_state([x.to(dt_t) for x in u])
    graph.set_state([x.to(dt_t) for x in u])
    dt = 2e-3 if n <= 64 else 5e-4
    graph.capture(dt)
    for _ in range(4):
        eager.step(dt)
        graph.step(dt)
        graph.divergence_norm()  # eager, synchronising work between the replays
        torch.cuda.synchronize()
        for a in range(3):
            assert torch.equal(eager.u0[a], graph.u0[a])
            for Le, Lg in zip(eager.les, graph.les):
                assert torch.equal(Le.u0[a], Lg.u0[a])
    graph.release_graphs()
