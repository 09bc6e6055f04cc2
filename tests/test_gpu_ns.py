"""GPU parity of the Navier-Stokes operators and the FFT projection against the
FP64 CPU oracle (same seeded inputs).  FP64 kernels must agree to ~1e-12
relative; FP32 kernels to the FP32 tolerances written at each test."""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SIZES = [8, 12, 15, 16, 32]


def dev(a, dtype=torch.float64):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0))).to("cuda", dtype).contiguous()


def host(t):
    return np.asfortranarray(t.double().cpu().numpy().transpose(2, 1, 0))


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def rand_u(O, n, seed):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(rng.standard_normal((n, n, n))) for _ in range(3)]


def P(ts, k):
    from ctypes import c_void_p
    return (c_void_p * k)(*[t.data_ptr() for t in ts])


@pytest.mark.parametrize("n", SIZES)
def test_rhs_fp64(gpu, O, n):
    u = rand_u(O, n, n)
    visc, l = 0.013, 1.0
    ref = O.rhs(u, visc, l)
    ud = [dev(x) for x in u]
    du = [torch.empty_like(x) for x in ud]
    ctx = gpu._abi.context()
    ctx.call("ec_ns_rhs", 1, n, l, visc, P(ud, 3), P(du, 3), 0)
    torch.cuda.synchronize()
    for a in range(3):
        assert rel(host(du[a]), ref[a]) < 1e-12
    # doadd: du = du - div sigma
    ctx.call("ec_ns_rhs", 1, n, l, visc, P(ud, 3), P(du, 3), 1)
    for a in range(3):
        assert rel(host(du[a]), 2 * ref[a]) < 1e-12


@pytest.mark.parametrize("n", [64, 256, 320])
def test_rhs_marching_tiles_fp64(gpu, O, n):
    """Grids that take the z-marching kernel (n >= 256: 64 x 8 column tiles,
    chunks of 16 to 64 planes) and one that does not: every tile and chunk
    boundary, with and without doadd."""
    test_rhs_fp64(gpu, O, n)


@pytest.mark.parametrize("n", [16, 256, 320])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_rhs_div(gpu, O, n, prec):
    """Fused du = -div sigma(u), d = div_h du (marching kernel for n % 64 == 0,
    the two plain kernels otherwise) against the oracle's rhs and divergence;
    then the FFT solve on d gives the pressure of ec_ns_pressure_solve."""
    dt = torch.float64 if prec == "f64" else torch.float32
    tol = 1e-12 if prec == "f64" else 2e-5
    u = rand_u(O, n, 3 * n + 1)
    visc, l = 0.011, 1.0
    ref = O.rhs(u, visc, l)
    dref = O.divergence(ref, l)
    ud = [dev(x, dt) for x in u]
    du = [torch.empty_like(x) for x in ud]
    d = torch.empty_like(ud[0])
    ctx = gpu._abi.context()
    code = 1 if prec == "f64" else 0
    ctx.call("ec_ns_rhs_div", code, n, l, visc, P(ud, 3), P(du, 3), gpu._abi.c_void_p(d.data_ptr()), 0.0)
    for a in range(3):
        assert rel(host(du[a]), ref[a]) < tol
    # the divergence cancels: measure it against the size of its terms
    scale = max(np.abs(r).max() for r in ref) * n
    assert np.abs(host(d) - dref).max() / scale < (1e-13 if prec == "f64" else 1e-6)
    if n & (n - 1):
        return  # the pressure comparison runs on the power-of-two grids
    NS = gpu.navierstokes
    g = NS.Grid(l, n)
    ps1, ps2 = NS.PoissonSolver(g, dt), NS.PoissonSolver(g, dt)
    ctx.call("ec_ns_pressure_solve", ps1.h, P(du, 3))
    ctx.call("ec_ns_rhs_div", code, n, l, visc, P(ud, 3), P(du, 3), gpu._abi.c_void_p(ps2.p_ptr), 0.0)
    ctx.call("ec_poisson_solve", ps2.h)
    torch.cuda.synchronize()
    p1, p2 = ps1.p.data, ps2.p.data
    assert float((p1 - p2).abs().max()) <= (1e-12 if prec == "f64" else 1e-5) * float(p1.abs().max())


@pytest.mark.parametrize("n", [16, 256])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_rhs_div_state_divergence(gpu, O, n, prec):
    """d = div_h du + r div_h u: the divergence of the stage state folded into
    the Poisson right-hand side (r = 6/dt at the last RK4 stage), against the
    oracle's divergence of du and of u (an unprojected random u, so div_h u
    is of the size of its terms)."""
    dt = torch.float64 if prec == "f64" else torch.float32
    u = rand_u(O, n, 5 * n + 3)
    visc, l, r = 0.011, 1.0, 6.0 / 1e-3
    ref = O.rhs(u, visc, l)
    dref = O.divergence(ref, l) + r * O.divergence(u, l)
    ud = [dev(x, dt) for x in u]
    du = [torch.empty_like(x) for x in ud]
    d = torch.empty_like(ud[0])
    ctx = gpu._abi.context()
    ctx.call("ec_ns_rhs_div", 1 if prec == "f64" else 0, n, l, visc, P(ud, 3), P(du, 3),
             gpu._abi.c_void_p(d.data_ptr()), r)
    scale = (max(np.abs(x).max() for x in ref) + r * max(np.abs(x).max() for x in u)) * n
    assert np.abs(host(d) - dref).max() / scale < (1e-13 if prec == "f64" else 1e-6)


@pytest.mark.parametrize("n", [512, 1024])
def test_rhs_div_large_fp32(gpu, n):
    """Full-size FP32 grids: the fused march (64 x 16 column tiles) returns the
    du of ec_ns_rhs (its own 64 x 8 march of the same stress cell) and, as d,
    the divergence of ec_ns_divergence applied to its own du, on the same
    seeded field; every tile and chunk edge of the launch is in play."""
    l, visc = 1.0, 1e-3
    gen = torch.Generator(device="cuda").manual_seed(n)
    u = [torch.randn((n, n, n), device="cuda", generator=gen) for _ in range(3)]
    du = [torch.empty_like(x) for x in u]
    d = torch.empty_like(u[0])
    ctx = gpu._abi.context()
    ctx.call("ec_ns_rhs_div", 0, n, l, visc, P(u, 3), P(du, 3), gpu._abi.c_void_p(d.data_ptr()), 0.0)
    ref = [torch.empty_like(x) for x in u]
    ctx.call("ec_ns_rhs", 0, n, l, visc, P(u, 3), P(ref, 3), 0)
    torch.cuda.synchronize()
    for a in range(3):
        scale = float(ref[a].abs().max())
        assert float((du[a] - ref[a]).abs().max()) <= 2e-6 * scale, a
    del u
    dref = torch.empty_like(d)
    ctx.call("ec_ns_divergence", 0, n, l, P(du, 3), gpu._abi.c_void_p(dref.data_ptr()))
    torch.cuda.synchronize()
    # the divergence cancels: measure it against the size of its terms
    scale = max(float(r.abs().max()) for r in du) * n / l
    assert float((d - dref).abs().max()) <= 1e-6 * scale
    del du, d, ref, dref
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n", [16, 256])
@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("s", [1.0, -1.0])
def test_rhs_add(gpu, O, n, prec, s):
    """du = s F(u) + add (FP64 addend, one rounding) on the plain kernel (16)
    and the marching kernel (256): against the oracle's s rhs + add, and
    bitwise against the two-pass form it replaces (the RHS written, then an
    FP64 axpy of the addend: ec_axpby for FP64 fields, ec_axpy2 for FP32)."""
    dt = torch.float64 if prec == "f64" else torch.float32
    code = 1 if prec == "f64" else 0
    u = rand_u(O, n, 5 * n + 3)
    add = rand_u(O, n, 5 * n + 4)
    visc, l = 0.012, 1.0
    ref = O.rhs(u, visc, l)
    ud = [dev(x, dt) for x in u]
    ad = [dev(x) for x in add]
    du = [torch.empty_like(x) for x in ud]
    ctx = gpu._abi.context()
    ctx.call("ec_ns_rhs_add", code, n, l, visc, P(ud, 3), P(du, 3), s, P(ad, 3))
    tol = 1e-12 if prec == "f64" else 2e-5
    for a in range(3):
        want = s * ref[a] + add[a]
        assert rel(host(du[a]), want) < tol, a
    two = [torch.empty_like(x) for x in ud]
    ctx.call("ec_ns_rhs", code, n, l, visc, P(ud, 3), P(two, 3), 0)
    for a in range(3):
        if prec == "f64":
            ctx.call("ec_axpby", 1, two[a].numel(), 1.0, ad[a].data_ptr(), s, two[a].data_ptr())
        else:
            if s != 1.0:
                two[a].neg_()  # exact
            ctx.call("ec_axpy2", 0, two[a].numel(), 1.0, 1, ad[a].data_ptr(), two[a].data_ptr())
    torch.cuda.synchronize()
    for a in range(3):
        assert torch.equal(du[a], two[a]), a


@pytest.mark.parametrize("n", [16, 32, 256])
def test_rhs_fp32(gpu, O, n):
    u = rand_u(O, n, 100 + n)
    visc, l = 0.01, 1.0
    ref = O.rhs(u, visc, l)
    ud = [dev(x, torch.float32) for x in u]
    du = [torch.empty_like(x) for x in ud]
    gpu._abi.context().call("ec_ns_rhs", 0, n, l, visc, P(ud, 3), P(du, 3), 0)
    for a in range(3):
        # FP32 rounding of inputs and a ~40-term stencil: 1e-5 relative to the field max
        assert rel(host(du[a]), ref[a]) < 2e-5


@pytest.mark.parametrize("n", [8, 12])
def test_stress_and_tensordivergence(gpu, O, n):
    u = rand_u(O, n, 7 * n)
    visc = 0.02
    s_ref = O.stresstensor(u, visc)
    ud = [dev(x) for x in u]
    s = [torch.empty_like(ud[0]) for _ in range(6)]
    ctx = gpu._abi.context()
    ctx.call("ec_ns_stresstensor", 1, n, 1.0, visc, P(ud, 3), P(s, 6))
    for a in range(6):
        assert rel(host(s[a]), s_ref[a]) < 1e-13
    du = [torch.empty_like(ud[0]) for _ in range(3)]
    ctx.call("ec_ns_tensordivergence", 1, n, 1.0, P(s, 6), P(du, 3), 0)
    ref = O.tensordivergence(s_ref)
    for a in range(3):
        assert rel(host(du[a]), ref[a]) < 1e-12
    # non-symmetric with FP64 sigma into FP32 and FP64 du
    rng = np.random.default_rng(n)
    s9 = [np.asfortranarray(rng.standard_normal((n, n, n))) for _ in range(9)]
    ref9 = O.tensordivergence_nonsym(s9, du=[x.copy(order="F") for x in ref], doadd=True)
    s9d = [dev(x) for x in s9]
    ctx.call("ec_ns_tensordivergence_nonsym", 1, n, 1.0, 1, P(s9d, 9), P(du, 3), 1)
    for a in range(3):
        assert rel(host(du[a]), ref9[a]) < 1e-12
    du32 = [dev(x, torch.float32) for x in ref]
    ctx.call("ec_ns_tensordivergence_nonsym", 0, n, 1.0, 1, P(s9d, 9), P(du32, 3), 1)
    for a in range(3):
        assert rel(host(du32[a]), ref9[a]) < 1e-5


@pytest.mark.parametrize("n", SIZES)
def test_divergence_gradient(gpu, O, n):
    u = rand_u(O, n, 3 * n)
    ud = [dev(x) for x in u]
    d = torch.empty_like(ud[0])
    ctx = gpu._abi.context()
    ctx.call("ec_ns_divergence", 1, n, 1.0, P(ud, 3), d.data_ptr())
    assert rel(host(d), O.divergence(u)) < 1e-13
    p = np.asfortranarray(np.random.default_rng(n).standard_normal((n, n, n)))
    ref = O.pressuregradient([x.copy(order="F") for x in u], p)
    ctx.call("ec_ns_pressuregradient", 1, n, 1.0, P(ud, 3), dev(p).data_ptr())
    for a in range(3):
        assert rel(host(ud[a]), ref[a]) < 1e-13


@pytest.mark.parametrize("n", SIZES + [48, 64, 128, 256])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_fft_roundtrip_and_spectrum(gpu, O, n, prec):
    NS = gpu.navierstokes
    dt = torch.float64 if prec == "f64" else torch.float32
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, dt)
    x = np.asfortranarray(np.random.default_rng(n).standard_normal((n, n, n)))
    X = NS.rfft(dev(x, dt), ps)
    ref = np.fft.rfftn(x, axes=(2, 1, 0))  # [kx, ky, kz]
    got = X.cpu().numpy().transpose(2, 1, 0)
    tol = 1e-12 if prec == "f64" else 3e-6
    assert np.abs(got - ref).max() / np.abs(ref).max() < tol
    back = NS.irfft(X, ps)
    assert rel(host(back), x) < tol


@pytest.mark.parametrize("n,prec", [(512, "f64"), (512, "f32"), (1024, "f32")])
def test_fft_plane_waves_large(gpu, n, prec):
    """Full-size grids: a sum of plane waves cos(2 pi (a i + b j + c k) / n + phi)
    has the half spectrum n^3/2 e^{i phi} at (a, b, c) (0 < a < n/2) and zero
    elsewhere; the inverse returns the field."""
    NS = gpu.navierstokes
    dt = torch.float64 if prec == "f64" else torch.float32
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, dt)
    modes = [(1, 0, 0, 0.3, 1.0), (5, n - 3, 7, -1.1, 0.5), (n // 2 - 1, n // 2, n - 1, 2.0, 0.25)]
    ar = torch.arange(n, device="cuda", dtype=torch.float64)
    x = torch.zeros((n, n, n), device="cuda", dtype=dt)
    for a, b, c, phi, amp in modes:  # x[k, j, i], i fastest (column-major (i, j, k))
        ph = (2 * np.pi / n) * ((a * ar).view(1, 1, n) + (b * ar).view(1, n, 1) + (c * ar).view(n, 1, 1))
        x += (amp * torch.cos(ph + phi)).to(dt)
        del ph
    X = NS.rfft(x, ps)
    half = n ** 3 / 2
    tol = 1e-12 if prec == "f64" else 3e-6
    for a, b, c, phi, amp in modes:
        v = complex(X[c, b, a].item())
        ref = amp * half * complex(np.cos(phi), np.sin(phi))
        assert abs(v - ref) < tol * half, (a, b, c, v, ref)
        X[c, b, a] = 0
    assert float(X.abs().max()) < tol * half
    for a, b, c, phi, amp in modes:
        X[c, b, a] = complex(amp * half * np.cos(phi), amp * half * np.sin(phi))
    back = NS.irfft(X, ps)
    assert float(back.sub(x).abs().max()) < 10 * tol * 1.75  # max |x| <= 1.75
    del X, back, x
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n", [512, 1024])
def test_poisson_plane_waves_large(gpu, n):
    """FP32 Poisson solve at full size: the discrete Laplacian of the staggered
    grid maps cos(2 pi (a i + b j + c k) / n + phi) to lambda times itself,
    lambda = -(4 / h^2) (sin^2(pi a / n) + sin^2(pi b / n) + sin^2(pi c / n)),
    so the right-hand side cos(...) has the solution cos(...) / lambda.

    One mode per solve.  FP32 rounding noise that the transforms leave in the
    low-|lambda| bins is amplified by |lambda| / lambda_min there, so the bound
    on |lambda p - f| is 2e-6 + 5e-9 |lambda| / lambda_min: tight (~2e-6) for
    modes near lambda_min, where a wrong eigenvalue index or scale shows at
    once.  Modes (n/2 - 1, 0, 0) are left out: their leakage is amplified by
    ~n^2/pi^2 and an FP32 library FFT misses them by ~1e-3 as well."""
    NS = gpu.navierstokes
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float32)
    h = 1.0 / n
    lmin = 4 / h ** 2 * np.sin(np.pi / n) ** 2
    ctx = gpu._abi.context()
    ar = torch.arange(n, device="cuda", dtype=torch.float64)
    for a, b, c, phi in [(1, 0, 0, 0.3), (0, 0, 1, 0.5), (0, 1, 0, -0.5), (2, n - 1, 1, 1.0), (5, n - 3, 7, -1.1),
                         (n // 2 - 1, n // 2, n - 1, 2.0), (0, 2, n // 2, 0.7), (n // 2, 1, 0, 0.4)]:
        lam = -(4 / h ** 2) * sum(np.sin(np.pi * m / n) ** 2 for m in (a, b, c))
        w = torch.cos((2 * np.pi / n) * ((a * ar).view(1, 1, n) + (b * ar).view(1, n, 1)
                                         + (c * ar).view(n, 1, 1)) + phi)  # [k, j, i], i fastest
        f = w.float()
        ctx.call("ec_d2d", gpu._abi.c_void_p(ps.p_ptr), gpu._abi.c_void_p(f.data_ptr()), f.numel() * 4)
        ctx.call("ec_poisson_solve", ps.h)
        got = ps.p.data
        err = float((got.double() * lam - w).abs().max())
        assert err < 2e-6 + 5e-9 * abs(lam) / lmin, (a, b, c, err)
        del w, f, got
    torch.cuda.empty_cache()



@pytest.mark.parametrize("n", [512, 1024])
def test_z_poisson_tridiagonal(gpu, n):
    """The FP32 Poisson solve with its z pass as per-column cyclic tridiagonal
    solves (two first-order recursions in FP64) against the transform z pass,
    both against the FP64 solve of the same right-hand side: the direct form
    is at least as close to it.  A single mode on the column (kx, ky) = (0, 1),
    the slowest-decaying recursion (r = 1 - O(1/n)), comes back as d / lambda
    with lambda the discrete Laplacian's eigenvalue.  Times both forms."""
    NS, abi = gpu.navierstokes, gpu._abi
    ctx = abi.context()
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float32)
    ps64 = NS.PoissonSolver(g, torch.float64)
    gen = torch.Generator(device="cuda").manual_seed(n + 7)
    d = torch.randn((n, n, n), dtype=torch.float32, device="cuda", generator=gen)
    d -= d.mean()
    h = 1.0 / n

    def solve(src, solver, nbytes):
        ctx.call("ec_d2d", abi.c_void_p(solver.p_ptr), abi.c_void_p(src.data_ptr()), nbytes)
        ctx.call("ec_poisson_solve", solver.h)
        return solver.p.data.clone()

    ref = solve(d.double(), ps64, d.numel() * 8)
    ar = torch.arange(n, device="cuda", dtype=torch.float64)
    # index order (x, y, z) of the fields: mode 1 along y and z
    mode = torch.cos(2 * math.pi * (ar.view(1, n, 1) + ar.view(1, 1, n)) / n).expand(n, n, n).contiguous()
    lam = -4.0 / h**2 * 2 * math.sin(math.pi / n) ** 2
    out, ms, err, merr = {}, {}, {}, {}
    try:
        for on in (1, 0):
            ctx.call("ec_poisson_set_z_tridiagonal", on)
            out[on] = solve(d, ps, d.numel() * 4)
            err[on] = float((out[on].double() - ref).abs().max() / ref.abs().max())
            pm = solve(mode.float(), ps, d.numel() * 4)
            merr[on] = float((pm.double() - mode / lam).abs().max() / (1.0 / abs(lam)))
            del pm
            torch.cuda.synchronize()
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(3):
                ctx.call("ec_poisson_solve", ps.h)
            t1.record()
            t1.synchronize()
            ms[on] = t0.elapsed_time(t1) / 3
    finally:
        ctx.call("ec_poisson_set_z_tridiagonal", 1)
    print(f"\n[poisson n={n} fp32] vs FP64: tridiagonal z {err[1]:.2e} ({ms[1]:.3f} ms), transform z {err[0]:.2e} "
          f"({ms[0]:.3f} ms); mode (0,1,1): {merr[1]:.2e} / {merr[0]:.2e}")
    assert err[1] < 2e-6 and err[1] <= 1.5 * err[0] + 1e-7
    assert merr[1] < 1e-5 and merr[0] < 1e-5
    del out, ref, d, mode
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n", [512, 1024])
def test_fft_cols_persistent_same_bits(gpu, n):
    """The persistent column passes (blocks that prefetch their next tile) and
    the one-block-per-tile passes run the same stages on the same values:
    forward and inverse transforms and the fused Poisson solve agree bit for
    bit.  Also times the Poisson solve both ways."""
    NS = gpu.navierstokes
    ctx = gpu._abi.context()
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float32)
    gen = torch.Generator(device="cuda").manual_seed(n + 1)
    x = torch.randn((n, n, n), dtype=torch.float32, device="cuda", generator=gen)
    out, ms = {}, {}
    try:
        for on in (1, 0):
            ctx.call("ec_fft_set_cols_persistent", on)
            X = NS.rfft(x, ps)
            back = NS.irfft(X, ps)
            ctx.call("ec_d2d", gpu._abi.c_void_p(ps.p_ptr), gpu._abi.c_void_p(x.data_ptr()), x.numel() * 4)
            ctx.call("ec_poisson_solve", ps.h)
            p = ps.p.data
            torch.cuda.synchronize()
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            t0.record()
            for _ in range(3):
                ctx.call("ec_poisson_solve", ps.h)
            t1.record()
            t1.synchronize()
            ms[on] = t0.elapsed_time(t1) / 3
            out[on] = (X, back, p)
            del X, back, p
    finally:
        ctx.call("ec_fft_set_cols_persistent", 1)
    print(f"\n[poisson solve n={n} fp32] persistent column passes {ms[1]:.3f} ms, one block per tile {ms[0]:.3f} ms")
    for a, b in zip(out[1], out[0]):
        assert torch.equal(a, b)
    assert float((out[1][1] - x).abs().max() / x.abs().max()) < 3e-6
    del out, x
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n", SIZES + [64])
def test_project_fp64(gpu, O, n):
    NS = gpu.navierstokes
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float64)
    u = rand_u(O, n, 11 * n)
    uref = [x.copy(order="F") for x in u]
    pref = O.project(uref)
    ud = [NS.Field(g, dev(x)) for x in u]
    NS.project_(ud, ps)
    for a in range(3):
        assert rel(ud[a].numpy(), uref[a]) < 1e-12
    assert rel(ps.p.numpy(), pref) < 1e-11
    d = O.divergence([x.numpy() for x in ud])
    assert np.abs(d).max() < 1e-10 * n


@pytest.mark.parametrize("n", [32, 64])
def test_project_fp32_divergence(gpu, O, n):
    NS = gpu.navierstokes
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float32)
    u = rand_u(O, n, 5 * n)
    ud = [NS.Field(g, dev(x, torch.float32)) for x in u]
    NS.project_(ud, ps)
    uh = [x.numpy() for x in ud]
    d = O.divergence(uh)
    h = 1.0 / n
    norm = h * np.sqrt((d ** 2).sum()) / np.sqrt(sum((x ** 2).sum() for x in uh))
    assert norm < 1e-6  # the FP32 divergence gate of the workload
    uref = [x.copy(order="F") for x in u]
    O.project(uref)
    for a in range(3):
        assert rel(uh[a], uref[a]) < 1e-5


@pytest.mark.parametrize("n", [8, 16, 32])
def test_randomfield(gpu, O, n):
    NS = gpu.navierstokes
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float64)
    u = NS.randomfield(g, ps, "peak", 5.0, 0.5, seed=123)
    ref = O.randomfield(n, seed=123)
    for a in range(3):
        assert rel(u[a].numpy(), ref[a]) < 1e-10
    e = sum(NS.sumsq(x) for x in u) / 2 / n ** 3
    assert abs(e - 0.5) < 1e-6


def test_reductions(gpu):
    NS = gpu.navierstokes
    x = torch.randn(1000003, dtype=torch.float64, device="cuda")
    y = torch.randn(1000003, dtype=torch.float64, device="cuda")
    assert abs(NS.maxabs(x) - x.abs().max().item()) == 0
    assert abs(NS.sumsq(x) - (x * x).sum().item()) < 1e-9 * (x * x).sum().item()
    assert abs(NS.diffsq(x, y) - ((x - y) ** 2).sum().item()) < 1e-9 * ((x - y) ** 2).sum().item()
    assert NS.sumsq(x[:0]) == 0.0
    xf = x.float()
    assert abs(NS.maxabs(xf) - xf.abs().max().item()) == 0


@pytest.mark.parametrize("stage", [0, 1, 2, 3])
def test_rk4_update(gpu, O, stage):
    n, dt = 12, 0.01
    rng = np.random.default_rng(stage)
    f = lambda: np.asfortranarray(rng.standard_normal((n, n, n)))
    u0, us, du, acc = [f() for _ in range(3)], [f() for _ in range(3)], [f() for _ in range(3)], [f() for _ in range(3)]
    p = f()
    k = O.pressuregradient([x.copy(order="F") for x in du], p)
    b = (1, 2, 2, 1)[stage]
    a = (0.5, 0.5, 1.0)[stage] if stage < 3 else None
    T = [[dev(x) for x in v] for v in (u0, us, du, acc)]
    gpu._abi.context().call("ec_ns_rk4_update", 1, n, 1.0, stage, dt, P(T[0], 3), P(T[1], 3), P(T[2], 3),
                            dev(p).data_ptr(), P(T[3], 3))
    for c in range(3):
        acc_new = k[c] if stage == 0 else acc[c] + b * k[c]
        if stage < 3:
            assert rel(host(T[3][c]), acc_new) < 1e-13
            assert rel(host(T[1][c]), u0[c] + a * dt * k[c]) < 1e-13
        else:
            assert rel(host(T[0][c]), u0[c] + dt / 6 * acc_new) < 1e-13


@pytest.mark.parametrize("stage", [0, 1, 2, 3])
def test_rk4_update_large_fp32(gpu, stage):
    """The 16-byte vector FP32 update at n = 512 (periodic x/y/z neighbours of p
    across vector, row and plane ends) against an FP64 torch evaluation of the
    same stage on the same FP32 inputs: acc += b k, us = u0 + a dt k, k = du - grad p."""
    n, dt, l = 512, 0.01, 1.0
    ih = n / l
    g = torch.Generator(device="cuda").manual_seed(100 + stage)
    f = lambda: torch.randn((n, n, n), device="cuda", generator=g)  # noqa: E731
    u0, us, du, acc = [[f() for _ in range(3)] for _ in range(4)]
    p = f()
    ref_u0 = [x.double() for x in u0]
    ref_acc = [x.double() for x in acc]
    pd = p.double()
    gp = [(torch.roll(pd, -1, dims=2 - a) - pd) * ih for a in range(3)]  # [k, j, i]: x is dim 2
    gpu._abi.context().call("ec_ns_rk4_update", 0, n, l, stage, dt, P(u0, 3), P(us, 3), P(du, 3), p.data_ptr(),
                            P(acc, 3))
    b = (1, 2, 2, 1)[stage]
    for c in range(3):
        k = du[c].double() - gp[c]
        acc_new = k if stage == 0 else ref_acc[c] + b * k
        if stage < 3:
            a = (0.5, 0.5, 1.0)[stage]
            assert float((acc[c].double() - acc_new).abs().max()) < 1e-6 * float(acc_new.abs().max())
            want = ref_u0[c] + a * dt * k
            assert float((us[c].double() - want).abs().max()) < 1e-6 * float(want.abs().max())
        else:
            want = ref_u0[c] + dt / 6 * acc_new
            assert float((u0[c].double() - want).abs().max()) < 1e-6 * float(want.abs().max())
        del k, acc_new
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n,prec", [(12, "f64"), (13, "f64"), (64, "f32"), (512, "f32")])
def test_rk4_update_states_matches_accumulator_form(gpu, O, n, prec):
    """Four stages of the stored-state RK4 (ec_ns_rk4_update_states) against
    the accumulator form (ec_ns_rk4_update) fed the same stage tendencies
    (du, p per stage): the stage states u1, u2, u3 equal the accumulator
    form's us of that stage bit for bit (same arithmetic), and the new u0
    agrees to rounding of the stored states (FP64 1e-14; FP32 2 ulp of |u|).
    n = 13 takes the one-cell-per-thread path (not 16-byte aligned rows), the
    others the 16-byte one.  At n = 12 the new u0 is also checked against the
    oracle's pressure gradient and the textbook tableau in FP64."""
    dt_t = torch.float64 if prec == "f64" else torch.float32
    code = 1 if prec == "f64" else 0
    step, l = 0.01, 1.0
    gen = torch.Generator(device="cuda").manual_seed(7 * n)
    f = lambda: torch.randn((n, n, n), device="cuda", dtype=dt_t, generator=gen)  # noqa: E731
    u0 = [f() for _ in range(3)]
    dus = [[f() for _ in range(3)] for _ in range(4)]
    ps_ = [f() for _ in range(4)]
    ctx = gpu._abi.context()
    # accumulator form
    ua, us, acc = [x.clone() for x in u0], [torch.empty_like(x) for x in u0], [torch.empty_like(x) for x in u0]
    us_seen = []
    # stored-state form
    ub = [x.clone() for x in u0]
    ust = [[torch.empty_like(x) for x in u0] for _ in range(3)]
    ust_p = P([x for s_ in ust for x in s_], 9)
    for st in range(4):
        ctx.call("ec_ns_rk4_update", code, n, l, st, step, P(ua, 3), P(us, 3), P(dus[st], 3), ps_[st].data_ptr(),
                 P(acc, 3))
        if st < 3:
            us_seen.append([x.clone() for x in us])
        ctx.call("ec_ns_rk4_update_states", code, n, l, st, step, P(ub, 3), ust_p, P(dus[st], 3),
                 ps_[st].data_ptr())
    torch.cuda.synchronize()
    for st in range(3):
        for a in range(3):
            assert torch.equal(ust[st][a], us_seen[st][a]), (st, a)
    tol = 1e-14 if prec == "f64" else 2.5e-7
    for a in range(3):
        assert float((ub[a] - ua[a]).abs().max()) <= tol * float(ua[a].abs().max()), a
    if n == 12:
        # the textbook tableau on the oracle's k = du - grad p of every stage
        ks = [O.pressuregradient([host(x) for x in dus[st]], host(ps_[st])) for st in range(4)]
        for a in range(3):
            want = host(u0[a]) + step / 6 * (ks[0][a] + 2 * ks[1][a] + 2 * ks[2][a] + ks[3][a])
            assert rel(host(ub[a]), want) < 1e-13
    del u0, dus, ps_, ua, us, acc, ub, ust, us_seen
    torch.cuda.empty_cache()


@pytest.mark.parametrize("n", [12, 13, 64])
def test_rk4_update_states_without_pressure(gpu, n):
    """p = NULL: k = du.  Four stages against the accumulator form given a
    zero pressure field (the form the LES grids' filtered velocity used): the
    stage states bit for bit, the new state to FP64 rounding."""
    step = 0.01
    gen = torch.Generator(device="cuda").manual_seed(11 * n)
    f = lambda: torch.randn((n, n, n), device="cuda", dtype=torch.float64, generator=gen)  # noqa: E731
    u0 = [f() for _ in range(3)]
    dus = [[f() for _ in range(3)] for _ in range(4)]
    zero = torch.zeros((n, n, n), device="cuda", dtype=torch.float64)
    ctx = gpu._abi.context()
    ua, us, acc = [x.clone() for x in u0], [torch.empty_like(x) for x in u0], [torch.empty_like(x) for x in u0]
    ub = [x.clone() for x in u0]
    ust = [[torch.empty_like(x) for x in u0] for _ in range(3)]
    ust_p = P([x for s_ in ust for x in s_], 9)
    for st in range(4):
        ctx.call("ec_ns_rk4_update", 1, n, 1.0, st, step, P(ua, 3), P(us, 3), P(dus[st], 3), zero.data_ptr(),
                 P(acc, 3))
        ctx.call("ec_ns_rk4_update_states", 1, n, 1.0, st, step, P(ub, 3), ust_p, P(dus[st], 3), None)
        torch.cuda.synchronize()
        if st < 3:
            for a in range(3):
                assert torch.equal(ust[st][a], us[a]), (st, a)
    for a in range(3):
        want = u0[a] + step / 6 * (dus[0][a] + 2 * dus[1][a] + 2 * dus[2][a] + dus[3][a])
        assert float((ub[a] - want).abs().max()) <= 1e-14 * float(want.abs().max())
        assert float((ub[a] - ua[a]).abs().max()) <= 1e-14 * float(ua[a].abs().max())


@pytest.mark.parametrize("n,prec", [(16, "f64"), (64, "f64"), (256, "f32")])
def test_project_into(gpu, n, prec):
    """dst = P src leaves src unchanged and equals copy + in-place projection
    bit for bit (same divergence, solve and gradient arithmetic)."""
    NS = gpu.navierstokes
    dt = torch.float64 if prec == "f64" else torch.float32
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, dt)
    gen = torch.Generator(device="cuda").manual_seed(n)
    src = [torch.randn((n, n, n), device="cuda", dtype=dt, generator=gen) for _ in range(3)]
    keep = [x.clone() for x in src]
    dst = [torch.empty_like(x) for x in src]
    ref = [x.clone() for x in src]
    ctx = gpu._abi.context()
    ctx.call("ec_ns_project", ps.h, P(ref, 3))
    ctx.call("ec_ns_project_into", ps.h, P(src, 3), P(dst, 3))
    for a in range(3):
        assert torch.equal(src[a], keep[a])
        assert torch.equal(dst[a], ref[a])


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_axpby(gpu, prec):
    """y = a x + b y in place (the closure forcing m = V k - F_H(P_H v))."""
    dt = torch.float64 if prec == "f64" else torch.float32
    g = torch.Generator(device="cuda").manual_seed(7)
    x = torch.randn(100003, device="cuda", dtype=dt, generator=g)
    y = torch.randn(100003, device="cuda", dtype=dt, generator=g)
    want = 1.5 * x.double() - 0.25 * y.double()
    gpu._abi.context().call("ec_axpby", 1 if prec == "f64" else 0, x.numel(), 1.5, x.data_ptr(), -0.25,
                            y.data_ptr())
    assert float((y.double() - want).abs().max()) <= (1e-15 if prec == "f64" else 1e-6) * float(want.abs().max())


@pytest.mark.parametrize("n,prec", [(16, "f64"), (32, "f64"), (24, "f64"), (64, "f32")])
def test_spectrum(gpu, O, n, prec):
    """spectrum(u, spectral_stuff(grid), poisson) (src/N.jl:547-590) against the
    oracle's shell sums on the same field: the shells 1..n/3 and their
    wavenumbers 2 pi / l * shell; all shells through the C entry point."""
    NS = gpu.navierstokes
    dt = torch.float64 if prec == "f64" else torch.float32
    l = 2.0
    g = NS.Grid(l, n)
    ps = NS.PoissonSolver(g, dt)
    u = rand_u(O, n, 9 * n)
    uf = tuple(NS.Field(g, dev(x, dt)) for x in u)
    stuff = NS.spectral_stuff(g)
    assert stuff["shells"] == list(range(1, (2 * (n // 2)) // 3 + 1))
    sp = NS.spectrum(uf, stuff, ps)
    nshell = int(np.floor(np.sqrt(3) * (n // 2))) + 1
    ref = O.spectrum_shells(u, nshell)
    tol = 1e-12 if prec == "f64" else 1e-5
    for k, e, sh in zip(sp["k"], sp["s"], stuff["shells"]):
        assert abs(k - 2 * np.pi / l * sh) < 1e-12
        assert abs(e - ref[sh]) <= tol * ref.max(), (sh, e, ref[sh])
    import ctypes
    out = (ctypes.c_double * nshell)()
    gpu._abi.context().call("ec_ns_spectrum", ps.h, P([x.data for x in uf], 3), nshell, out)
    got = np.array(out[:])
    assert np.abs(got - ref).max() <= tol * ref.max()
    logs = NS.spectral_stuff(g, npoint=5)["shells"]
    assert logs[0] == 1 and logs[-1] == (2 * (n // 2)) // 3 and logs == sorted(set(logs))


@pytest.mark.parametrize("n", [512, 1024])
def test_spectrum_large_fp32(gpu, n):
    """Full-size FP32: all shells together hold the kinetic energy mean(|u|^2)/2
    (Parseval, FP64 sums on both sides, 1e-5); a single mode cos(2 pi (a i +
    b j + c k)/n) of amplitude A put in one component lands whole in shell
    floor(|k|) with energy A^2/4."""
    NS = gpu.navierstokes
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float32)
    gen = torch.Generator(device="cuda").manual_seed(n + 5)
    u = tuple(NS.Field(g, torch.randn((n, n, n), device="cuda", generator=gen)) for _ in range(3))
    import ctypes
    nshell = int(np.floor(np.sqrt(3) * (n // 2))) + 1
    out = (ctypes.c_double * nshell)()
    ctx = gpu._abi.context()
    ctx.call("ec_ns_spectrum", ps.h, P([x.data for x in u], 3), nshell, out)
    energy = sum(float((x.data.double() ** 2).sum()) for x in u) / 2 / n ** 3
    assert abs(sum(out[:]) - energy) <= 1e-5 * energy
    ar = torch.arange(n, device="cuda", dtype=torch.float64)
    a, b, c, amp = 7, n - 5, 3, 1.5  # signed wavenumbers (7, -5, 3)
    w = amp * torch.cos((2 * np.pi / n) * ((a * ar).view(1, 1, n) + (b * ar).view(1, n, 1) + (c * ar).view(n, 1, 1))
                        + 0.4)  # [k, j, i], i fastest
    z = torch.zeros((n, n, n), device="cuda")
    one = (NS.Field(g, z), NS.Field(g, w.float()), NS.Field(g, z))
    del w
    ctx.call("ec_ns_spectrum", ps.h, P([x.data for x in one], 3), nshell, out)
    sh = int(np.floor(np.sqrt(7 ** 2 + 5 ** 2 + 3 ** 2)))
    got = np.array(out[:])
    assert abs(got[sh] - amp ** 2 / 4) <= 1e-5 * amp ** 2 / 4
    assert np.abs(np.delete(got, sh)).max() <= 1e-9
    del u, one, z
    torch.cuda.empty_cache()


def test_invalid_arguments_raise(gpu):
    NS = gpu.navierstokes
    ctx = gpu._abi.context()
    with pytest.raises(gpu._abi.ECError):
        NS.PoissonSolver(NS.Grid(1.0, 7 * 11), torch.float64)  # factor 7 unsupported
    with pytest.raises(gpu._abi.ECError):
        P3 = gpu._abi.P3
        ctx.call("ec_ns_rk4_update", 1, 8, 1.0, 5, 0.1, P3(), P3(), P3(), None, P3())
    with pytest.raises(gpu._abi.ECError):
        import ctypes
        g8 = NS.Grid(1.0, 8)
        ps8 = NS.PoissonSolver(g8, torch.float64)
        u8 = NS.vectorfield(g8)
        ctx.call("ec_ns_spectrum", ps8.h, P([x.data for x in u8], 3), 0, (ctypes.c_double * 1)())


@pytest.mark.parametrize("n", [16, 32])
def test_steppers_fe_wray3_vs_oracle(gpu, O, n):
    """step_forwardeuler! and step_wray3! (src/N.jl:309-346) through the host
    mirror (the fused RHS, axpy and the FFT projection on the GPU, FP64)
    against the oracle's steppers over 5 steps each, to 1e-12."""
    NS = gpu.navierstokes
    g = NS.Grid(1.0, n)
    visc, dt = 0.01, 2e-3
    u0 = O.randomfield(n, seed=n + 1)
    ps = NS.PoissonSolver(g, torch.float64)
    for name in ("fe", "wray3"):
        u = tuple(NS.Field.from_numpy(g, x) for x in u0)
        du, us = NS.vectorfield(g), NS.vectorfield(g)
        ref = [x.copy(order="F") for x in u0]
        for _ in range(5):
            if name == "fe":
                NS.step_forwardeuler_(u, du, ps, visc, dt)
                O.step_forwardeuler(ref, visc, dt)
            else:
                NS.step_wray3_(u, du, us, ps, visc, dt)
                O.step_wray3(ref, visc, dt)
        for a in range(3):
            assert rel(u[a].numpy(), ref[a]) < 1e-12, (name, a)


def test_steppers_convergence_order_on_gpu(gpu, O):
    """Observed order in dt of the GPU FP64 steppers over [0, 0.08] at n = 16
    (the pins of the oracle's steppers, run on the product path): forward
    Euler 1, Wray RK3 3, the RK4 DNS step of DNSAidedLES (no LES grids) 4."""
    NS, D = gpu.navierstokes, gpu.dnsaid
    n, visc, T = 16, 0.01, 0.08
    g = NS.Grid(1.0, n)
    u0 = O.randomfield(n, seed=5)
    ps = NS.PoissonSolver(g, torch.float64)

    def run(kind, m):
        dt = T / m
        if kind == "rk4":
            drv = D.DNSAidedLES(g, visc, [], dtype=torch.float64)
            drv.set_state([dev(x) for x in u0])
            drv.run(dt, m)
            return [host(x) for x in drv.u0]
        u = tuple(NS.Field.from_numpy(g, x) for x in u0)
        du, us = NS.vectorfield(g), NS.vectorfield(g)
        for _ in range(m):
            if kind == "fe":
                NS.step_forwardeuler_(u, du, ps, visc, dt)
            else:
                NS.step_wray3_(u, du, us, ps, visc, dt)
        return [x.numpy() for x in u]

    for kind, ms, lo, hi in (("fe", (8, 16, 32, 64), 0.9, 1.2), ("wray3", (4, 8, 16, 32), 2.8, 3.4),
                             ("rk4", (4, 8, 16, 32), 3.8, 4.4)):
        us = [run(kind, m) for m in ms]
        d = [max(np.abs(a - b).max() for a, b in zip(us[i], us[i + 1])) for i in range(len(ms) - 1)]
        p = np.log2(d[-2] / d[-1])
        assert lo < p < hi, (kind, p, d)
