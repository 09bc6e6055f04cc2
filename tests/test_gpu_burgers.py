"""GPU parity of the batched Burgers DNS-aided LES persistent kernel against the
FP64 CPU oracle (per instance), graph replay against eager launches, and the
exactness of the swap-filter LES (src/Burgers.jl:1087-1089 identity)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _small_batch():
    n = [64, 128, 96, 256, 48, 512, 60]
    c = [4, 8, 3, 16, 6, 64, 5]
    return n, c


def test_burgers_rhs(gpu, O):
    n, L, visc = 200, 2 * np.pi, 5e-4
    u = O.burgers_ic(n, 3)
    ud = torch.from_numpy(u).cuda()
    du = torch.empty_like(ud)
    gpu._abi.context().call("ec_burgers_rhs", 1, n, L, visc, ud.data_ptr(), du.data_ptr())
    assert rel(du.cpu().numpy(), O.burgers_rhs(u, visc, L)) < 1e-12


@pytest.mark.parametrize("threads", [128, 256, 1024])
def test_batch_fp64_matches_oracle(gpu, O, threads):
    B = gpu.burgers
    n, c = _small_batch()
    visc, L, seed, steps = 5e-3, 2 * np.pi, 11, 40
    b = B.BurgersBatch(n, c, visc=visc, L=L, cfl=0.3, seed=seed, dtype="f64")
    b.set_threads(threads)
    b.run(steps, use_graph=False)
    for i in range(len(n)):
        u0 = O.burgers_ic(n[i], seed + i)
        dt = b.dt(i)
        assert abs(dt - 0.3 * O.burgers_cfl(u0, visc, L)) < 1e-15 * dt
        u, us, um = O.burgers_dnsaid_rk4(u0, c[i], visc, dt, steps, L)
        assert rel(b.get(i, "dns"), u) < 1e-12
        assert rel(b.get(i, "swapfil"), us) < 1e-11
        assert rel(b.get(i, "nomodel"), um) < 1e-11
        # exactness: the swap-filter LES is the filtered DNS
        _, w = O.tophat_weights(c[i]) if c[i] % 2 else O.trapezoid_weights(c[i])
        vu = O.burgers_filter_coarse(u, n[i] // c[i], w, -(c[i] // 2))
        assert rel(b.get(i, "swapfil"), vu) < 1e-12


@pytest.mark.parametrize("prec,threads,n,c", [
    # 8 points per thread; LES arrays of PT/4 = 2 points (largest LES grid 512 <= 2 x 256)
    ("f64", 256, [2048, 1024, 2048], [8, 4, 16]),
    # 8 points per thread; LES arrays of PT/2 = 4 points (LES grid 1024 > 2 x 256)
    ("f64", 256, [2048, 1024], [2, 4]),
    # the benchmark's build: 1024 threads, 16 points per thread, LES arrays of 4
    ("f32", 1024, [16384, 8192, 2048], [4, 8, 64]),
    ("f64", 1024, [16384, 4096], [16, 4]),
])
def test_batch_points_per_thread_builds(gpu, O, prec, threads, n, c):
    """Builds with several points per thread (the register arrays of the DNS
    state and of the LES states in their PT/2 and PT/4 widths) against the
    oracle; FP64 to 1e-12, FP32 to 2e-5 after 10 steps."""
    B = gpu.burgers
    visc, L, seed, steps = 5e-4, 2 * np.pi, 21, 10
    b = B.BurgersBatch(n, c, visc=visc, L=L, cfl=0.3, seed=seed, dtype=prec)
    b.set_threads(threads)
    b.run(steps, use_graph=False)
    tol = 1e-12 if prec == "f64" else 2e-5
    for i in range(len(n)):
        u0 = O.burgers_ic(n[i], seed + i)
        u, us, um = O.burgers_dnsaid_rk4(u0, c[i], visc, b.dt(i), steps, L)
        assert rel(b.get(i, "dns"), u) < tol, i
        assert rel(b.get(i, "swapfil"), us) < 10 * tol, i
        assert rel(b.get(i, "nomodel"), um) < 10 * tol, i
    b.close()


def test_batch_fp32_and_graph_replay(gpu, O):
    B = gpu.burgers
    n, c = _small_batch()
    visc, L, seed = 5e-4, 2 * np.pi, 3
    eager = B.BurgersBatch(n, c, visc=visc, L=L, seed=seed, dtype="f32")
    graph = B.BurgersBatch(n, c, visc=visc, L=L, seed=seed, dtype="f32")
    for _ in range(3):
        eager.run(25, use_graph=False)
        graph.run(25, use_graph=True)  # captured once, replayed
        torch.cuda.synchronize()
        for i in range(len(n)):
            for w in ("dns", "swapfil", "nomodel"):
                assert np.array_equal(eager.get(i, w), graph.get(i, w))
    for i in range(len(n)):
        u0 = O.burgers_ic(n[i], seed + i)
        u, us, um = O.burgers_dnsaid_rk4(u0, c[i], visc, eager.dt(i), 75, L)
        assert rel(eager.get(i, "dns"), u) < 1e-4
        assert rel(eager.get(i, "swapfil"), us) < 1e-4


def test_baseline_batch_shape(gpu):
    B = gpu.burgers
    n, c = B.baseline_batch(4096)
    assert len(n) == 4096 and int(n.sum()) == 1638 * 2048 + 1229 * 4096 + 819 * 8192 + 410 * 16384
    b = B.BurgersBatch(n[:64], c[:64], dtype="f32")
    b.run(2)
    torch.cuda.synchronize()
    assert np.isfinite(b.get(0, "dns")).all()


def test_invalid_batch_rejected(gpu):
    B = gpu.burgers
    with pytest.raises(gpu._abi.ECError):
        B.BurgersBatch([100], [3])  # 3 does not divide 100
    with pytest.raises(gpu._abi.ECError):
        B.BurgersBatch([32768], [4], dtype="f64")  # exceeds LDS
    # 128 threads reach 64 points each: 8192 points, not 16384
    b = B.BurgersBatch([16384, 2048], [4, 8], dtype="f32")
    with pytest.raises(gpu._abi.ECError):
        b.set_threads(128)
    b.set_threads(256)
    b.close()


@pytest.mark.parametrize("n,nH,kind,dr,ns", [(225, 15, "tophat", 1.0, 4), (225, 15, "gaussian", 2.0, 4),
                                             (135, 45, "tophat", 3.0, 3), (64, 64, "tophat", 3.0, 2),
                                             (13500, 300, "gaussian", 2.0, 6), (9000, 10, "tophat", 8.0, 2)])
def test_burgers_fractions(gpu, O, n, nH, kind, dr, ns):
    """compute_fractions (src/Burgers.jl:1022-1099) on the GPU against the
    oracle on the same snapshots, FP64: the mean fractions to 1e-12 (the last
    case is the reference's own setup, nh = 13500, nH = 300, Delta = 2H,
    Gaussian, on a few samples); the GPU path builds its kernels with the host
    mirror of tophat/gaussian_weights and build_doublekernel, the oracle with
    its own restatement of them.  The fine pass stages each
    block's stretch of the row in LDS; the 9000 / 10 case (double kernel of
    8101 taps) is wider than that stage and takes the direct pass."""
    import math
    B = gpu.burgers
    L, visc = 2 * math.pi, 5e-4
    rng = np.random.default_rng(n + nH)
    x = np.arange(n) * L / n
    U = np.array([sum(rng.standard_normal() / k * np.cos(k * x + rng.uniform(0, 2 * np.pi)) for k in range(1, 11))
                  for _ in range(ns)])
    got = B.compute_fractions(torch.from_numpy(U).cuda(), nH, L, visc, dr, kind)
    les, dbl = O.burgers_les_kernels(n, nH, L, dr, kind)
    want = O.burgers_fractions(U, nH, L, visc, les, dbl)
    for key, w in zip(("classic", "flux", "div"), want):
        assert abs(got[key] - w) < 1e-12, (key, got[key], w)
    assert abs(sum(got.values()) - 1) < 1e-9


@pytest.mark.parametrize("n,nH,kind,dr,ns", [(225, 15, "tophat", 1.0, 4), (225, 45, "gaussian", 2.0, 3),
                                             (64, 64, "tophat", 3.0, 2), (13500, 300, "gaussian", 2.0, 5),
                                             (13500, 2700, "tophat", 4.0, 3)])
def test_burgers_dissipation(gpu, O, n, nH, kind, dr, ns):
    """compute_dissipation (src/Burgers.jl:447-516) on the GPU against the
    oracle on the same snapshots, FP64, every coefficient of every model to
    1e-10 of the model's largest (the parts are differences of O(1) stresses
    filtered by sums of up to 2K + 1 terms that the GPU contracts into FMAs, so
    they carry about 1e-15 absolute rounding against values near 1e-3); the
    reference's setup sizes are included; the layout is the reference's
    vec(d[:, m, :]), coarse point fastest."""
    import math
    B = gpu.burgers
    L, visc = 2 * math.pi, 5e-4
    U = _smag_samples(n, ns, n + 3 * nH, L)
    got = B.compute_dissipation(torch.from_numpy(U).cuda(), nH, L, visc, dr, kind)
    les, dbl = O.burgers_les_kernels(n, nH, L, dr, kind)
    want = O.burgers_dissipation(U, nH, L, visc, les, dbl)
    assert got["nomodel"].cpu().tolist() == [0.0]
    for key, w in zip(("class_m", "class_p", "swapfil"), want):
        g = got[key].cpu().numpy()
        assert g.shape == (ns * nH,)
        scale = np.abs(w).max()
        assert np.abs(g - w.ravel()).max() <= 1e-10 * scale, (key, np.abs(g - w.ravel()).max(), scale)


def _smag_samples(n, ns, seed, L):
    rng = np.random.default_rng(seed)
    x = np.arange(n) * L / n
    return np.array([sum(rng.standard_normal() / k * np.cos(k * x + rng.uniform(0, 2 * np.pi)) for k in range(1, 11))
                     for _ in range(ns)])


@pytest.mark.parametrize("n,nH,kind,dr", [(225, 15, "gaussian", 2.0), (225, 45, "tophat", 1.0),
                                          (13500, 300, "gaussian", 2.0), (13500, 900, "gaussian", 0.0),
                                          (4000, 8, "tophat", 6.0), (9000, 10, "tophat", 8.0)])
def test_smagorinsky_fit(gpu, O, n, nH, kind, dr):
    """The four sums of the Smagorinsky fit (smagorinsky.jl:24-36) against the
    oracle on the same snapshots (1e-10 relative; FP64 sums in another
    order); theta from them; the reference's setup sizes included.  The last
    two cases have kernels wider than the LDS-staged passes hold (K = 1750:
    direct fit pass; K = 4050: direct fine and fit passes)."""
    import math
    B = gpu.burgers
    L, visc = 2 * math.pi, 5e-4
    U = _smag_samples(n, 3, n + nH, L)
    les, dbl = B.les_kernels(n, nH, L, dr, kind)
    ol, od = O.burgers_les_kernels(n, nH, L, dr, kind)
    want = O.smagorinsky_sums(U, nH, L, visc, dr * L / nH, ol, od)
    Ud = torch.from_numpy(U).cuda()
    import ctypes
    out = (ctypes.c_double * 4)()
    dp = ctypes.POINTER(ctypes.c_double)
    wl, wd = np.ascontiguousarray(les[1]), np.ascontiguousarray(dbl[1])
    gpu._abi.context().call("ec_burgers_smag_fit", n, nH, L, visc, dr * L / nH, 3, gpu._abi.c_void_p(Ud.data_ptr()),
                            les[0], wl.ctypes.data_as(dp), dbl[0], wd.ctypes.data_as(dp), out)
    got = np.array(out[:])
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max(), (got, want)
    if want[0] < 0 and want[2] < 0:
        th = B.smagorinsky_coefficients(Ud, nH, L, visc, dr, kind)
        assert abs(th["classic"] - math.sqrt(-want[0] / want[1])) < 1e-9
        assert abs(th["discret"] - math.sqrt(-want[2] / want[3])) < 1e-9


@pytest.mark.parametrize("n,nH,kind,dr,C", [(225, 15, "gaussian", 2.0, 2e-3), (192, 64, "tophat", 1.0, 0.0),
                                            (13500, 300, "gaussian", 2.0, 1e-3), (13500, 2700, "gaussian", 2.0, 4e-8)])
def test_smagorinsky_solve(gpu, O, n, nH, kind, dr, C):
    """solve_smagorinsky (src/Burgers.jl:682-772) on the GPU, every sample in
    one launch, against the oracle's per-sample loop (same CFL steps, the
    reference's accumulating RK4 stages), FP64 to 1e-11; nH = 2700 (eleven
    points per thread) with C near the fitted theta^2 (Delta^2 + H^2) of that
    grid (the CFL step ignores the eddy viscosity, so a C far above it blows
    the explicit scheme up in both codes)."""
    import math
    B = gpu.burgers
    L, visc, tstop = 2 * math.pi, 5e-4, 0.1
    U = _smag_samples(n, 4, 3 * n + nH, L)
    H = L / nH
    theta = math.sqrt(C / ((dr * H) ** 2 + H ** 2))
    got = B.solve_smagorinsky(torch.from_numpy(U).cuda(), nH, theta, L, visc, dr, kind, tstop).cpu().numpy()
    _, dbl = O.burgers_les_kernels(n, nH, L, dr, kind)
    want = O.smagorinsky_solve(U, nH, L, visc, C, tstop, 0.3, dbl)
    assert np.isfinite(want).all()
    assert np.abs(got - want).max() <= 1e-11 * np.abs(want).max()


def test_study_entry_points_reject_bad_arguments(gpu):
    """The closure-study entry points refuse inconsistent sizes before any
    launch: nH not dividing n, kernels wider than the grid, an LES state
    beyond the 4096 points one block holds."""
    import ctypes
    ctx = gpu._abi.context()
    dp = ctypes.POINTER(ctypes.c_double)
    w = np.ones(3) / 3
    u = torch.zeros((2, 90), dtype=torch.float64, device="cuda")
    out = (ctypes.c_double * 4)()
    with pytest.raises(gpu._abi.ECError):
        ctx.call("ec_burgers_fractions", 90, 7, 1.0, 1e-3, 2, gpu._abi.c_void_p(u.data_ptr()), 1,
                 w.ctypes.data_as(dp), 1, w.ctypes.data_as(dp), out)
    d = torch.empty((3, 2 * 7), dtype=torch.float64, device="cuda")
    with pytest.raises(gpu._abi.ECError):
        ctx.call("ec_burgers_dissipation", 90, 7, 1.0, 1e-3, 2, gpu._abi.c_void_p(u.data_ptr()), 1,
                 w.ctypes.data_as(dp), 1, w.ctypes.data_as(dp), gpu._abi.c_void_p(d.data_ptr()))
    big = np.ones(201) / 201
    with pytest.raises(gpu._abi.ECError):
        ctx.call("ec_burgers_dissipation", 90, 9, 1.0, 1e-3, 2, gpu._abi.c_void_p(u.data_ptr()), 100,
                 big.ctypes.data_as(dp), 1, w.ctypes.data_as(dp), gpu._abi.c_void_p(d.data_ptr()))
    with pytest.raises(gpu._abi.ECError):
        ctx.call("ec_burgers_smag_fit", 90, 9, 1.0, 1e-3, 0.1, 2, gpu._abi.c_void_p(u.data_ptr()), 100,
                 big.ctypes.data_as(dp), 1, w.ctypes.data_as(dp), out)
    ub = torch.empty((2, 9000), dtype=torch.float64, device="cuda")
    u2 = torch.zeros((2, 18000), dtype=torch.float64, device="cuda")
    with pytest.raises(gpu._abi.ECError):
        ctx.call("ec_burgers_smag_solve", 18000, 9000, 1.0, 1e-3, 0.0, 0.1, 0.3, 2, gpu._abi.c_void_p(u2.data_ptr()),
                 1, w.ctypes.data_as(dp), gpu._abi.c_void_p(ub.data_ptr()))


def test_burgers_randomfield_vs_oracle(gpu, O):
    """randomfield with bump_profile (src/Burgers.jl:192-202) for several
    samples in one launch against the oracle's numpy irfft of the whole
    spectrum (the GPU sums modes up to 6 kpeak): 1e-13 of the field's maximum,
    energy to 1e-12."""
    B = gpu.burgers
    n, kp, E, L = 1350, 10, 2.0, 2 * np.pi
    u = B.randomfield(n, kp, E, nsamples=3, seed=7, L=L).cpu().numpy()
    for q in range(3):
        want = O.burgers_randomfield(n, kp, E, 7 + q, L)
        assert rel(u[q], want) < 1e-13, q
        assert abs(np.sum(u[q] ** 2) / 2 * (L / n) - E) < 1e-12 * E


@pytest.mark.parametrize("nh,nH,kind,dr", [(1350, 30, "gaussian", 2.0), (1350, 90, "tophat", 1.0),
                                           (1350, 270, "gaussian", 2.0), (2025, 135, "gaussian", 0.0)])
def test_burgers_dns_aided_les_vs_oracle(gpu, O, nh, nH, kind, dr):
    """dns_aided_les (src/Burgers.jl:350-445), every sample in one persistent
    launch, against the oracle's per-sample loop: same number of CFL steps
    and all seven fields (DNS, filtered DNS, the four LES) to 1e-11 of their
    maximum after t = 0.01; the swap-filter LES equals the filtered DNS to
    rounding (the exact closure)."""
    B = gpu.burgers
    L, visc, tstop = 2 * np.pi, 5e-4, 0.01
    ustart = B.randomfield(nh, 10, 2.0, nsamples=2, seed=3, L=L)
    got = B.dns_aided_les(ustart, nH, visc, dr * L / nH, tstop, 0.3, kind, L)
    les, dbl = O.burgers_les_kernels(nh, nH, L, dr, kind)
    u0 = ustart.cpu().numpy()
    for q in range(2):
        want, steps = O.burgers_dns_aided_les(u0[q], nH, L, visc, les, dbl, tstop)
        assert got["steps"][q] == steps
        for key in B.FIELDS:
            assert rel(got[key][q].cpu().numpy(), want[key]) < 1e-11, (key, q)
        f = got["dns_fil"][q]
        assert float((got["swapfil"][q] - f).norm() / f.norm()) < 1e-14
