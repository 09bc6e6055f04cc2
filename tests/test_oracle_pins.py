"""Pins of the FP64 CPU oracle (oracle/) by what the mathematics fixes.

The reference holds no golden vectors and cannot run here (Julia is absent),
so every oracle function is pinned by checks chosen so a plausible mistake
(dropped term, wrong sign or index, transposed operand) fails one of them:

  FFT          numpy.fft (rfftn / fft) and a brute-force DFT
  stress/RHS   viscous part against the spectral Laplacian eigenvalues
               (-4/h^2 sum sin^2(pi k/n)); inviscid energy conservation of the
               divergence-form convection for div-free u; momentum
               conservation; second-order convergence to the continuous
               Taylor-Green right-hand side
  projection   div_h(P u) = 0; P idempotent; P(grad phi) = 0; eigenvalue
               closed form for one Fourier mode
  filters      constants preserved; box transfer function on a Fourier mode;
               the commutation identity V(-div_h r) = -div_H^ns (S r) for a
               random tensor r (which fixes every offset of S)
  closure      exactness of the classic+flux+div LES (navierstokes.jl:33:
               5.7e-15) under forward Euler and RK4, odd and even c, with and
               without the Gaussian; error ordering of the five variants
               (navierstokes.jl:30-34)
  randomfield  divergence-free, total energy, shell spectrum follows the
               profile (src/N.jl:441-449)
  Burgers      RK4 4th-order convergence; swap-filter exactness; the
               identity classic + flux + div = swap (src/Burgers.jl:1087-1089);
               closure fractions: mirror symmetry of every part under the
               staggered reflection, the comp = 1 and constant-field limits;
               dissipation coefficients: the pinned parts times a closed-form
               coarse velocity gradient, mirror symmetry
  spectrum     full-FFT shells, Parseval, single modes
  steppers     observed order under dt-halving: forward Euler 1, Wray RK3 3,
               the NS RK4 tableau 4 (mutated tableaus fail)
  SFS          volume swap-filter stress = exact closure; surface line flux
               exact across the face; every filter window from its region
               definition; all stresses vanish at n_les = n_dns; dissipation by
               summation by parts and locality
"""
import math

import numpy as np
import pytest


def rnd(n, seed, k=3):
    rng = np.random.default_rng(seed)
    return [np.asfortranarray(rng.standard_normal((n, n, n))) for _ in range(k)]


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [4, 6, 8, 9, 10, 12, 15, 16, 20, 24])
def test_fft_matches_numpy(O, n):
    x = rnd(n, n, 1)[0]
    X = O.rfft3(x)
    ref = np.fft.rfftn(x, axes=(2, 1, 0))
    assert np.abs(X - ref).max() < 1e-12 * np.abs(ref).max()
    assert np.abs(O.irfft3(X) - x).max() < 1e-13
    z = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    assert np.abs(O.fft1d(z) - np.fft.fft(z)).max() < 1e-12 * n
    assert np.abs(O.fft1d(z, +1) - n * np.fft.ifft(z)).max() < 1e-12 * n


def test_fft_brute_force_dft(O):
    n = 5
    x = rnd(n, 1, 1)[0]
    X = O.rfft3(x)
    idx = np.arange(n)
    for kx in range(n // 2 + 1):
        for ky in range(n):
            for kz in (0, 3):
                ph = np.exp(-2j * np.pi * (kx * idx[:, None, None] + ky * idx[None, :, None] + kz * idx[None, None, :]) / n)
                assert abs(X[kx, ky, kz] - (x * ph).sum()) < 1e-12


# ---------------------------------------------------------------- RHS
def spectral_laplacian(f):
    n = f.shape[0]
    F = np.fft.fftn(f)
    s = np.sin(np.pi * np.fft.fftfreq(n) * n / n) ** 2
    lam = -4 * n ** 2 * (s[:, None, None] + s[None, :, None] + s[None, None, :])
    return np.real(np.fft.ifftn(F * lam))


def test_rhs_viscous_part_is_laplacian(O):
    n, visc, eps = 12, 0.7, 1e-7
    u = rnd(n, 2)
    O.project(u)
    du = O.rhs([eps * x for x in u], visc)
    for a in range(3):
        lap = spectral_laplacian(u[a])
        # du = visc * lap(u) + O(eps^2) convection for a div-free field
        assert np.abs(du[a] / eps - visc * lap).max() < 1e-5 * np.abs(visc * lap).max()


def test_rhs_inviscid_energy_conservation(O):
    n = 16
    u = O.randomfield(n, seed=9)
    du = O.rhs(u, 0.0)
    e = sum(float((a * b).sum()) for a, b in zip(u, du))
    scale = sum(float(np.abs(a * b).sum()) for a, b in zip(u, du))
    assert abs(e) < 1e-13 * scale
    # and momentum: sum of each component vanishes (periodic flux form)
    for a in range(3):
        assert abs(du[a].sum()) < 1e-12 * np.abs(du[a]).sum()
    # a wrong stencil loses the invariant: a shifted stress breaks it
    s = O.stresstensor(u, 0.0)
    s_bad = list(s)
    s_bad[3] = np.roll(s[3], 1, axis=0)
    du_bad = O.tensordivergence(s_bad)
    e_bad = sum(float((a * b).sum()) for a, b in zip(u, du_bad))
    assert abs(e_bad) > 1e-6 * scale


def tg_rhs_error(O, n, visc):
    h = 1.0 / n
    x = np.arange(n) * h
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    c = h / 2
    # Taylor-Green velocity sampled at the face centres
    u1 = np.sin(2 * np.pi * (X + h)) * np.cos(2 * np.pi * (Y + c)) * np.cos(2 * np.pi * (Z + c))
    u2 = -np.cos(2 * np.pi * (X + c)) * np.sin(2 * np.pi * (Y + h)) * np.cos(2 * np.pi * (Z + c))
    u3 = 0 * u1
    du = O.rhs([np.asfortranarray(u1), np.asfortranarray(u2), np.asfortranarray(u3)], visc)
    def exact(px, py, pz):
        s, co = np.sin, np.cos
        k = 2 * np.pi
        U = s(k * px) * co(k * py) * co(k * pz)
        V = -co(k * px) * s(k * py) * co(k * pz)
        Ux = k * co(k * px) * co(k * py) * co(k * pz); Uy = -k * s(k * px) * s(k * py) * co(k * pz)
        Vx = k * s(k * px) * s(k * py) * co(k * pz); Vy = -k * co(k * px) * co(k * py) * co(k * pz)
        lapU, lapV = -3 * k * k * U, -3 * k * k * V
        # -div(u u) + visc lap u for the div-free TG field: -(u.grad)u + visc lap u
        return -(U * Ux + V * Uy) + visc * lapU, -(U * Vx + V * Vy) + visc * lapV
    e1, _ = exact(X + h, Y + c, Z + c)
    _, e2 = exact(X + c, Y + h, Z + c)
    return max(np.abs(du[0] - e1).max(), np.abs(du[1] - e2).max())


def test_rhs_second_order_taylor_green(O):
    e = [tg_rhs_error(O, n, 0.01) for n in (16, 32, 64)]
    r1, r2 = e[0] / e[1], e[1] / e[2]
    assert 3.5 < r1 < 4.5 and 3.7 < r2 < 4.3, e


# ---------------------------------------------------------------- projection
@pytest.mark.parametrize("n", [8, 9, 16])
def test_projection_properties(O, n):
    u = rnd(n, 3 * n)
    O.project(u)
    assert np.abs(O.divergence(u)).max() < 1e-11 * n
    v = [x.copy(order="F") for x in u]
    O.project(v)
    assert max(np.abs(a - b).max() for a, b in zip(u, v)) < 1e-13
    phi = rnd(n, 7, 1)[0]
    g = [np.asfortranarray(np.zeros_like(phi)) for _ in range(3)]
    O.pressuregradient(g, -phi)  # g = +grad phi (forward differences)
    O.project(g)
    assert max(np.abs(x).max() for x in g) < 1e-11 * n


def test_projection_single_mode_pressure(O):
    n = 16
    h = 1.0 / n
    i = np.arange(n)
    k = 3
    phi = np.asfortranarray(np.cos(2 * np.pi * k * i / n)[:, None, None] * np.ones((n, n, n)))
    u = [np.asfortranarray(np.zeros((n, n, n))) for _ in range(3)]
    O.pressuregradient(u, -phi)  # u = grad phi
    p = O.project(u)
    # L p = div u = L phi  =>  p = phi (zero-mean part)
    assert np.abs(p - phi).max() < 1e-12


# ---------------------------------------------------------------- filters
@pytest.mark.parametrize("c", [2, 3, 4, 5])
def test_filters_constants_and_transfer(O, c):
    n = 6 * c
    for name, (src, fs) in O.closure_filters(c, "va").items():
        assert all(abs(w.sum() - 1) < 1e-15 for w, _ in fs), name
    # box of cells on a Fourier mode along y: transfer sin(pi k c/n)/(c sin(pi k/n))
    k = 2
    j = np.arange(n)
    f = np.asfortranarray(np.ones((n, n, n)) * np.cos(2 * np.pi * k * (j + 0.5) / n)[None, :, None])
    out = O.filter_coarse(f, n // c, O._point(), O._box(c), O._point())
    J = np.arange(n // c)
    centre = (c * J + c / 2) / n
    expect = np.sin(np.pi * k * c / n) / (c * np.sin(np.pi * k / n)) * np.cos(2 * np.pi * k * centre)
    assert np.abs(out[0, :, 0] - expect).max() < 1e-13


@pytest.mark.parametrize("c", [2, 3, 4, 5])
def test_commutation_identity(O, c):
    """V(-div_h r) = -div_H^ns (S r) to rounding for a random symmetric tensor r."""
    n = 5 * c if c % 2 else 4 * c
    nb = n // c
    rng = np.random.default_rng(c)
    s = [np.asfortranarray(rng.standard_normal((n, n, n))) for _ in range(6)]  # 11,22,33,12,23,31
    p = np.asfortranarray(rng.standard_normal((n, n, n)))
    rfull = [s[0] + p, s[1] + p, s[2] + p, s[3], s[4], s[5]]
    du = O.tensordivergence(rfull)
    F = O.closure_filters(c, "va")
    V = [O.filter_coarse(du[a], nb, *F[f"ubar{a + 1}"][1]) for a in range(3)]
    src = {"r11": rfull[0], "r22": rfull[1], "r33": rfull[2], "s12": s[3], "s23": s[4], "s31": s[5]}
    Sr = [O.filter_coarse(src[F["rDHi" + ab][0]], nb, *F["rDHi" + ab][1]) for ab in O.NONSYM_ORDER]
    rhs = O.tensordivergence_nonsym(Sr)
    for a in range(3):
        assert np.abs(V[a] - rhs[a]).max() < 1e-12 * np.abs(V[a]).max()


# ---------------------------------------------------------------- closure exactness
@pytest.mark.parametrize("n,nl,dr", [(15, 5, 0.0), (15, 5, 2.0), (16, 4, 0.0), (24, 8, 2.0)])
def test_exactness_forward_euler(O, n, nl, dr):
    u = O.randomfield(n, seed=2, profile="power", param=-5 / 3, totalenergy=1.0)
    les = O.dnsaid_project_fe(u, nl, 2e-3, 4, 1e-3, delta_ratio=dr)
    assert O.compute_errors(les["u_cfd"], les["u_les"]) < 1e-13


def test_error_ordering_odd_c(O):
    """Same ordering as the reference's pins (navierstokes.jl:30-34):
    no-model > classic > classic+flux > symmetrised > exact ~ 1e-15."""
    n, nl = 30, 6
    u = O.randomfield(n, seed=0, profile="power", param=-5 / 3, totalenergy=1.0)
    les = O.dnsaid_project_fe(u, nl, 2e-4, 8, 2e-3, delta_ratio=2.0)
    e = {k: O.compute_errors(les[k], les["u_les"]) for k in ("u_nomo", "u_c", "u_cf", "u_cfd_symm", "u_cfd")}
    assert e["u_nomo"] > e["u_c"] > e["u_cf"] > e["u_cfd_symm"] > 1e3 * e["u_cfd"]
    assert e["u_cfd"] < 1e-13


@pytest.mark.parametrize("sets", [[(4, "va")], [(3, "va")], [(2, "va"), (4, "va")]])
def test_exactness_rk4(O, sets):
    n = 12
    u = O.randomfield(n, seed=4)
    u2, les, _ = O.dnsaid_rk4([x.copy(order="F") for x in u], sets, 1e-3, 2e-3, 3)
    for (c, kind), l in zip(sets, les):
        ref = [O.filter_coarse(u2[a], n // c, *O.closure_filters(c, kind)[f"ubar{a + 1}"][1]) for a in range(3)]
        O.project(ref)
        assert O.compute_errors(l, ref) < 1e-13


# ---------------------------------------------------------------- randomfield
def test_randomfield_spectrum(O):
    n = 32
    u = O.randomfield(n, seed=123, profile="peak", param=5.0, totalenergy=0.5)
    assert np.abs(O.divergence(u)).max() < 1e-10
    E = sum((x ** 2).sum() for x in u) / 2 / n ** 3
    assert abs(E - 0.5) < 1e-3
    # shell energies follow the profile k^4 exp(-2 (k/5)^2)
    spec = np.zeros(n)
    kk = np.fft.fftfreq(n) * n
    K = np.sqrt(kk[:, None, None] ** 2 + kk[None, :, None] ** 2 + kk[None, None, :] ** 2)
    for x in u:
        X = np.fft.fftn(x) / n ** 3
        np.add.at(spec, np.floor(K).astype(int).ravel() % n, (np.abs(X) ** 2 / 2).ravel())
    kd = int(np.floor(np.sqrt(3) * n / 2))
    prof = np.array([k ** 4 * np.exp(-2 * (k / 5) ** 2) for k in range(kd + 1)])
    target = 0.5 * prof / prof.sum()
    for k in range(1, 12):
        assert abs(spec[k] - target[k]) < 0.02 * target[k] + 1e-6, (k, spec[k], target[k])


# ---------------------------------------------------------------- spectrum
def _isqrt(kk):
    s = np.floor(np.sqrt(kk.astype(np.float64))).astype(np.int64)
    s[(s + 1) ** 2 <= kk] += 1
    s[s ** 2 > kk] -= 1
    return s


@pytest.mark.parametrize("n", [4, 8, 12])
def test_spectrum_full_fft_and_parseval(O, n):
    """Shells of the half spectrum with the conjugate bins counted twice
    (getshells, src/N.jl:478-545) equal the shells of the full complex FFT
    (numpy.fft.fftn over all signed wavenumbers), and they sum to the kinetic
    energy mean(|u|^2)/2 (Parseval) -- the identity the getshells comment
    states for even n (src/N.jl:516-523).  A missing or misplaced conjugate
    weight, a wrong normalisation or a shell off by one fails here."""
    u = rnd(n, 40 + n)
    nshell = int(np.floor(np.sqrt(3) * (n // 2))) + 1
    s = O.spectrum_shells(u, nshell)
    kw = np.round(np.fft.fftfreq(n) * n).astype(np.int64)
    sh = _isqrt(kw[:, None, None] ** 2 + kw[None, :, None] ** 2 + kw[None, None, :] ** 2)
    full = np.zeros(nshell)
    for x in u:
        F = np.fft.fftn(x) / n ** 3
        np.add.at(full, sh.ravel(), (np.abs(F) ** 2 / 2).ravel())
    assert np.abs(s - full).max() < 1e-14 * full.max()
    energy = sum((x ** 2).sum() for x in u) / 2 / n ** 3
    assert abs(s.sum() - energy) < 1e-13 * energy


def test_spectrum_single_modes(O):
    """amp cos(2 pi (a i + b j + c k)/n + phi) in one component has energy
    amp^2/4, all of it in shell floor(sqrt(a^2 + b^2 + c^2)); modes with
    kx = 0 (stored twice in the half spectrum, weight 1) and kx > 0 (stored
    once, weight 2) must both give it."""
    n = 16
    nshell = 14
    i = np.arange(n)
    for a, b, c, amp in [(3, 0, 0, 1.0), (0, 2, 5, 0.5), (2, -3, 1, 2.0), (0, 0, 4, 1.5), (5, 5, 5, 1.0)]:
        ph = 2 * np.pi * (a * i[:, None, None] + b * i[None, :, None] + c * i[None, None, :]) / n + 0.3
        x = np.asfortranarray(amp * np.cos(ph))
        z = np.zeros_like(x, order="F")
        for comp in range(3):
            u = [z, z, z]
            u[comp] = x
            s = O.spectrum_shells(u, nshell)
            k = int(math.isqrt(a * a + b * b + c * c))
            assert abs(s[k] - amp ** 2 / 4) < 1e-13, (a, b, c, comp, s[k])
            rest = np.delete(s, k)
            assert np.abs(rest).max() < 1e-13


# ---------------------------------------------------------------- Burgers closure fractions
def _bkernels(n, nH, L, dr, kind):
    from conftest import ROOT  # noqa: F401
    h, H = L / n, L / nH
    comp = n // nH
    fvm = (comp // 2, np.full(comp, 1.0 / comp))
    if kind == "tophat":
        R = int(np.rint(H * dr / h)) // 2
        les = (R, np.full(2 * R + 1, 1.0 / (2 * R + 1)))
    elif H * dr < 1e-15:
        les = (0, np.ones(1))  # gaussian_weights at Delta = 0 (src/Burgers.jl:146-156)
    else:
        s2 = (H * dr / h) ** 2 / 12
        R = int(np.rint(3 * math.sqrt(s2)))
        w = np.exp(-np.arange(-R, R + 1) ** 2 / (2 * s2))
        les = (R, w / w.sum())
    return les, (fvm[0] + les[0], np.convolve(fvm[1], les[1]))


def _bsamples(n, ns, seed):
    rng = np.random.default_rng(seed)
    x = np.arange(n) * 2 * np.pi / n
    return np.array([sum(rng.standard_normal() / k * np.cos(k * x + rng.uniform(0, 2 * np.pi))
                         for k in range(1, 11)) for _ in range(ns)])


@pytest.mark.parametrize("kind,dr", [("tophat", 1.0), ("tophat", 2.0), ("gaussian", 2.0)])
def test_burgers_fractions_reflection(O, kind, dr):
    """The staggered reflection u~[i] = -u[n - i] (faces i = 1..n, u[0] = u[n])
    maps the Burgers stress to itself at the mirrored centre, so with odd comp
    and symmetric kernels every closure part at coarse centre J equals the
    part at nH + 1 - J of the original field; a centre or face sampled one
    cell off, or a filter applied in the wrong direction, breaks this."""
    n, nH, L, visc = 225, 15, 2 * math.pi, 5e-3
    les, dbl = _bkernels(n, nH, L, dr, kind)
    U = _bsamples(n, 3, 7)
    idx = (n - np.arange(n) - 2) % n  # 0-based form of u~[i] = -u[n - i]
    Ur = -U[:, idx]
    a = O.burgers_fractions(U, nH, L, visc, les, dbl, parts=True)
    b = O.burgers_fractions(Ur, nH, L, visc, les, dbl, parts=True)
    for x, y in zip(a, b):
        assert np.abs(y - x[:, ::-1]).max() < 1e-12 * np.abs(x).max()
    fa = O.burgers_fractions(U, nH, L, visc, les, dbl)
    fb = O.burgers_fractions(Ur, nH, L, visc, les, dbl)
    assert np.abs(fa - fb).max() < 1e-13
    assert abs(fa.sum() - 1.0) < 1e-12  # fractions of each point sum to one


def test_burgers_fractions_degenerate_cases(O):
    """comp = 1 (the LES grid is the DNS grid): the double kernel is the LES
    kernel and the two stresses of the flux part coincide, so flux and div
    vanish exactly and the classic part carries everything; a constant field
    has no closure at all (all parts zero to rounding)."""
    n, L, visc = 64, 2 * math.pi, 1e-2
    les, dbl = _bkernels(n, n, L, 3.0, "tophat")
    assert les[0] == 1 and dbl[0] == 1
    U = _bsamples(n, 2, 3)
    cl, fl, dv = O.burgers_fractions(U, n, L, visc, les, dbl, parts=True)
    assert np.abs(fl).max() == 0 and np.abs(dv).max() == 0 and np.abs(cl).max() > 1e-3
    f = O.burgers_fractions(U, n, L, visc, les, dbl)
    assert abs(f[0] - 1) < 1e-9 and f[1] == 0 and f[2] == 0
    les, dbl = _bkernels(135, 9, L, 2.0, "gaussian")
    C = np.full((2, 135), 0.7)
    # (the fractions of rounding-level parts are not meaningful; eps only keeps them finite)
    assert all(np.abs(x).max() < 1e-15 for x in O.burgers_fractions(C, 9, L, visc, les, dbl, parts=True))


@pytest.mark.parametrize("n,nH,kind,dr", [(225, 15, "tophat", 1.0), (225, 45, "gaussian", 2.0),
                                          (135, 9, "tophat", 2.0)])
def test_burgers_dissipation_closed_form_gradient(O, n, nH, kind, dr):
    """compute_dissipation's coefficients are the closure parts times the
    coarse velocity gradient: class_m = classic g, class_p = (classic + flux) g,
    swapfil = (classic + flux + div) g, with the parts of the pinned fractions
    restatement and g_J = (vu_J - vu_{J-1}) / H written from the field's
    Fourier modes: u[i] = sum_k c_k cos(k t_i + p_k), t_i = 2 pi i / n, and a
    symmetric double kernel give vu_J = sum_k c_k T_k cos(k t_f + p_k) at the
    coarse face f = (J + 1) comp - 1 with T_k = sum_j w_j cos(2 pi k j / n).  A
    gradient of the wrong sign, one cell off or at the wrong face fails."""
    L, visc = 2 * math.pi, 5e-3
    les, dbl = _bkernels(n, nH, L, dr, kind)
    rng = np.random.default_rng(n + nH)
    t = 2 * np.pi * np.arange(n) / n
    modes = [[(rng.standard_normal() / k, k, rng.uniform(0, 2 * np.pi)) for k in range(1, 9)] for _ in range(3)]
    U = np.array([sum(c * np.cos(k * t + p) for c, k, p in m) for m in modes])
    K, wd = dbl
    comp = n // nH
    tf = 2 * np.pi * ((np.arange(nH) + 1) * comp - 1) / n
    g = []
    for m in modes:
        vu = sum(c * sum(wd[j + K] * math.cos(2 * math.pi * k * j / n) for j in range(-K, K + 1)) * np.cos(k * tf + p)
                 for c, k, p in m)
        g.append((vu - np.roll(vu, 1)) / (L / nH))
    g = np.array(g)
    cl, fl, dv = O.burgers_fractions(U, nH, L, visc, les, dbl, parts=True)
    d = O.burgers_dissipation(U, nH, L, visc, les, dbl)
    for got, part in zip(d, (cl, cl + fl, cl + fl + dv)):
        want = part * g
        assert np.abs(got - want).max() < 1e-11 * np.abs(want).max()


def test_burgers_dissipation_reflection(O):
    """Under the staggered reflection u~[i] = -u[n - i] both the closure parts
    and the coarse velocity gradient are even, so every dissipation
    coefficient at coarse centre J equals the one at nH + 1 - J of the
    original field (a forward difference instead of the backward one breaks
    this)."""
    n, nH, L, visc = 225, 15, 2 * math.pi, 5e-3
    les, dbl = _bkernels(n, nH, L, 2.0, "gaussian")
    U = _bsamples(n, 3, 11)
    Ur = -U[:, (n - np.arange(n) - 2) % n]
    for x, y in zip(O.burgers_dissipation(U, nH, L, visc, les, dbl),
                    O.burgers_dissipation(Ur, nH, L, visc, les, dbl)):
        assert np.abs(y - x[:, ::-1]).max() < 1e-12 * np.abs(x).max()


@pytest.mark.parametrize("kind,dr", [("tophat", 2.0), ("gaussian", 2.0), ("gaussian", 0.0)])
def test_smagorinsky_sums_reflection(O, kind, dr):
    """<S,T> and <S,S> of both fits are unchanged by the staggered reflection
    u~[i] = -u[n - i]: S and T are centre quantities, even under it, and the
    coarse stencil U[i - a - 1], U[i + a] is centred on i; a stencil one cell
    off or a filter in the wrong direction breaks the invariance."""
    n, nH, L, visc = 225, 15, 2 * math.pi, 5e-3
    les, dbl = _bkernels(n, nH, L, dr, kind)
    U = _bsamples(n, 3, 11)
    Ur = -U[:, (n - np.arange(n) - 2) % n]
    a = O.smagorinsky_sums(U, nH, L, visc, dr * L / nH, les, dbl)
    b = O.smagorinsky_sums(Ur, nH, L, visc, dr * L / nH, les, dbl)
    assert np.abs(a - b).max() < 1e-12 * np.abs(a).max()
    assert a[1] > 0 and a[3] > 0


def test_smagorinsky_sums_identity_limit(O):
    """comp = 1 and Delta = 0: every filter is the identity, both sub-filter
    stresses vanish and the two Smagorinsky stresses coincide."""
    n, L, visc = 64, 2 * math.pi, 1e-2
    les, dbl = _bkernels(n, n, L, 0.0, "gaussian")
    assert les[0] == 0 and len(les[1]) == 1 and dbl[0] == 0
    U = _bsamples(n, 2, 5)
    st = O.smagorinsky_sums(U, n, L, visc, 0.0, les, dbl)
    assert st[0] == 0 and st[2] == 0 and st[1] > 0 and abs(st[1] - st[3]) < 1e-14 * st[1]


def test_smagorinsky_solve_first_order_and_reflection(O):
    """The reference's RK4 loop lands each axpy! on the running state
    (src/Burgers.jl:748-755), so its stages sit at u0 + dt/2 k1, u0 + dt/2 (k1
    + k2), ...: a consistent but first-order scheme.  With C = 0 the solve is
    the coarse Burgers flow: against the oracle's pinned classical RK4 at a
    tiny step its error halves with the CFL factor (order 1; the textbook
    tableau would drop it 16-fold).  With C > 0 the solve commutes with the
    staggered reflection (the Smagorinsky stencil and the coarse sampling)."""
    n, nH, L, visc, tstop = 192, 64, 2 * math.pi, 1e-2, 0.05
    les, dbl = _bkernels(n, nH, L, 1.0, "tophat")
    U = _bsamples(n, 1, 13)
    ub0 = O.smagorinsky_solve(U, nH, L, visc, 0.0, 0.0, 0.3, dbl)[0]  # tstop 0: the filtered start
    ref = ub0.copy()
    O.lib().or_burgers_rk4(nH, L, visc, tstop / 4000, 4000, O._p(ref))
    e = [np.abs(O.smagorinsky_solve(U, nH, L, visc, 0.0, tstop, c, dbl)[0] - ref).max() for c in (0.1, 0.05, 0.025)]
    assert 1.7 < e[0] / e[1] < 2.3 and 1.7 < e[1] / e[2] < 2.3, e
    Ur = -U[:, (n - np.arange(n) - 2) % n]
    C = 1e-3  # theta^2 (Delta^2 + H^2) for theta ~ 0.2, Delta = H ~ 0.1
    a = O.smagorinsky_solve(U, nH, L, visc, C, tstop, 0.3, dbl)[0]
    b = O.smagorinsky_solve(Ur, nH, L, visc, C, tstop, 0.3, dbl)[0]
    assert np.abs(b - (-a[(nH - np.arange(nH) - 2) % nH])).max() < 1e-12 * np.abs(a).max()
    assert np.abs(a - O.smagorinsky_solve(U, nH, L, visc, 0.0, tstop, 0.3, dbl)[0]).max() > 1e-6


def test_randn_moments(O):
    x = O.randn(400000, 5)
    assert abs(x.mean()) < 5e-3 and abs(x.std() - 1) < 5e-3
    assert np.array_equal(O.randn(10, 5, 7), x[7:17])


# ---------------------------------------------------------------- Burgers
def test_burgers_rk4_fourth_order(O):
    n, visc, L = 256, 0.05, 2 * np.pi
    u0 = O.burgers_ic(n, 1, kmax=3)
    T = 0.2
    ref = O.burgers_rk4(u0, visc, T / 640, 640, L)
    errs = [np.abs(O.burgers_rk4(u0, visc, T / m, m, L) - ref).max() for m in (20, 40, 80)]
    assert 14 < errs[0] / errs[1] < 18 and 14 < errs[1] / errs[2] < 18, errs


@pytest.mark.parametrize("c", [3, 4, 5, 8, 64])
def test_burgers_swap_exactness_and_identity(O, c):
    n, visc, L = 64 * c if c < 64 else 4096, 5e-4, 2 * np.pi
    u0 = O.burgers_ic(n, 2)
    dt = 0.3 * O.burgers_cfl(u0, visc, L)
    u, us, um = O.burgers_dnsaid_rk4(u0, c, visc, dt, 20, L)
    w = np.full(c + (c % 2 == 0), 1.0 / c)
    if c % 2 == 0:
        w[0] = w[-1] = 0.5 / c
    vu = O.burgers_filter_coarse(u, n // c, w, -(c // 2))
    assert np.abs(us - vu).max() < 1e-13 * np.abs(vu).max()
    assert np.abs(um - vu).max() > 1e-6


@pytest.mark.parametrize("c", [2, 3, 4, 7, 8])
def test_burgers_commutation(O, c):
    """V(delta_coll s) = delta_coll_H(Pc s) for any s: the identity that makes the
    swap-filter closure exact (src/Burgers.jl:1087-1089 decomposition)."""
    n, L = 40 * c, 2 * np.pi
    h, H = L / n, L * c / n
    s = np.random.default_rng(c).standard_normal(n)
    ds = (np.roll(s, -1) - s) / h
    if c % 2:
        wv, wp = np.full(c, 1.0 / c), np.ones(1)
    else:
        wv = np.full(c + 1, 1.0 / c)
        wv[0] = wv[-1] = 0.5 / c
        wp = np.array([0.5, 0.5])
    lhs = O.burgers_filter_coarse(ds, n // c, wv, -(c // 2))
    ps = O.burgers_filter_coarse(s, n // c, wp, -(c // 2))
    rhs = (np.roll(ps, -1) - ps) / H
    assert np.abs(lhs - rhs).max() < 1e-12 * np.abs(lhs).max()


# ---------------------------------------------------------------- NS time steppers
def _ns_order(O, step, ms, n=16, visc=0.01, T=0.08, seed=5):
    """Observed order log2(|u_m - u_2m| / |u_2m - u_4m|) of a stepper over
    [0, T] with m, 2m, 4m ... steps, from a divergence-free random field."""
    u0 = O.randomfield(n, seed=seed)

    def run(m):
        u = [x.copy(order="F") for x in u0]
        for _ in range(m):
            step(u, T / m)
        return u

    us = [run(m) for m in ms]
    d = [max(np.abs(a - b).max() for a, b in zip(us[i], us[i + 1])) for i in range(len(ms) - 1)]
    return [math.log2(d[i] / d[i + 1]) for i in range(len(d) - 1)]


def test_ns_forward_euler_first_order(O):
    """step_forwardeuler! (src/N.jl:309-318): first order in dt."""
    p = _ns_order(O, lambda u, dt: O.step_forwardeuler(u, 0.01, dt), (8, 16, 32, 64))
    assert 0.9 < p[-1] < 1.2, p


def test_ns_wray3_third_order(O):
    """step_wray3! (src/N.jl:320-346), a = (8/15, 5/12, 3/4), b = (1/4, 0):
    third order in dt (the projection at every stage keeps the stages on the
    divergence-free subspace, where the scheme is the plain low-storage RK3)."""
    p = _ns_order(O, lambda u, dt: O.step_wray3(u, 0.01, dt), (4, 8, 16, 32))
    assert 2.8 < p[-1] < 3.4, p


def test_ns_rk4_fourth_order_and_mutants(O, monkeypatch):
    """The RK4 DNS step of dnsaid_rk4 (tableau of rk4_timestep!,
    src/Burgers.jl:120-142, projected stage tendencies): fourth order in dt;
    a wrong stage time (a3 = 1/2) or a wrong weight order (b = 1, 2, 1, 2)
    drops it to first order."""
    step = lambda u, dt: O.dnsaid_rk4(u, [], 0.01, dt, 1)
    p = _ns_order(O, step, (4, 8, 16, 32))
    assert 3.8 < p[-1] < 4.4, p
    monkeypatch.setattr(O, "RK4_A", (0.5, 0.5, 0.5))
    assert _ns_order(O, step, (4, 8, 16, 32))[-1] < 1.5
    monkeypatch.setattr(O, "RK4_A", (0.5, 0.5, 1.0))
    monkeypatch.setattr(O, "RK4_B", (1.0, 2.0, 1.0, 2.0))
    assert _ns_order(O, step, (4, 8, 16, 32))[-1] < 1.5


@pytest.mark.parametrize("c", [2, 3, 4])
def test_face_average_keeps_divergence_free(O, c):
    """The face average of a discretely divergence-free field is coarse
    divergence-free (its coarse fluxes are face sums, so div_H V u is the cell
    average of div_h u); the volume average is not (the projection of
    apply_DHpi, src/N.jl:1228-1237, does real work there)."""
    n = 8 * c
    u = O.randomfield(n, seed=c)
    div = {}
    for kind in ("fa", "va"):
        F = O.closure_filters(c, kind)
        v = [O.filter_coarse(O._f(u[a]), n // c, *F[f"ubar{a + 1}"][1]) for a in range(3)]
        div[kind] = np.abs(O.divergence(v)).max() / (max(np.abs(x).max() for x in v) * n // c)
    assert div["fa"] < 1e-14 and div["va"] > 1e-3, div


def test_burgers_bump_randomfield(O):
    """randomfield with bump_profile (src/Burgers.jl:192-202): energy sum u^2/2 h
    equals the requested one, the mean is zero, and the Fourier amplitudes
    follow (k/kp)^4 exp(-2 (k/kp)^2) up to one common factor (a wrong profile
    exponent, a lost factor 2 of the Hermitian half or a shifted k fails)."""
    n, kp, E, L = 2048, 10, 2.0, 2 * np.pi
    u = O.burgers_randomfield(n, kp, E, 5, L)
    assert abs(np.sum(u ** 2) / 2 * (L / n) - E) < 1e-12 * E
    assert abs(u.mean()) < 1e-13
    c = np.abs(np.fft.rfft(u)) / n
    k = np.arange(1, 31)  # amplitudes above 1e-5 of the peak (the rest sink into rounding)
    prof = (k / kp) ** 4 * np.exp(-2 * (k / kp) ** 2)
    ratio = c[1:31] / prof
    assert np.abs(ratio / ratio[9] - 1).max() < 1e-8


@pytest.mark.parametrize("nH,kind", [(30, "gaussian"), (90, "tophat"), (270, "gaussian")])
def test_burgers_dns_aided_les_swap_exactness(O, nH, kind):
    """dns_aided_les (src/Burgers.jl:350-445) with odd compressions: the
    swap-filter LES equals the filtered DNS (dns_fil) to rounding after every
    forward-Euler step (the exact closure), while no model, classic and
    classic+flux stay O(1e-3) away; a swapped fsu/vsu or a centre one cell
    off breaks the first."""
    nh, L, visc = 1350, 2 * np.pi, 5e-4
    u = O.burgers_randomfield(nh, 10, 2.0, 3, L)
    les, dbl = O.burgers_les_kernels(nh, nH, L, 2.0, kind)
    f, steps = O.burgers_dns_aided_les(u, nH, L, visc, les, dbl, 0.01)
    err = {k: np.linalg.norm(f[k] - f["dns_fil"]) / np.linalg.norm(f["dns_fil"])
           for k in ("nomodel", "class_m", "class_p", "swapfil")}
    assert steps >= 10 and err["swapfil"] < 1e-14
    assert min(err["nomodel"], err["class_m"], err["class_p"]) > 1e-4


# ---------------------------------------------------------------- SFS tensors and dissipation
NONSYM = ["11", "21", "31", "12", "22", "32", "13", "23", "33"]


def test_tensordissipation_summation_by_parts(O):
    """sum_I d[I] = sum_I u . tensordivergence_nonsym(sigma): the staggered
    sigma : grad u is minus the work of the staggered divergence, so a product
    at the wrong edge, a derivative of the wrong sign or one cell off fails."""
    n, l = 7, 2.0
    rng = np.random.default_rng(3)
    sig = [np.asfortranarray(rng.standard_normal((n, n, n))) for _ in range(9)]
    u = rnd(n, 4)
    d = O.tensordissipation(sig, u, l)
    du = O.tensordivergence_nonsym(sig, l=l)
    want = sum(float(np.sum(a * b)) for a, b in zip(u, du))
    assert abs(d.sum() - want) < 1e-12 * np.abs(d).sum()


@pytest.mark.parametrize("ab", NONSYM)
def test_tensordissipation_locality(O, ab):
    """A single nonzero sigma_ab at index J0 feeds exactly the centres its
    position touches: the centre J0 for a diagonal entry (with delta_a u_a
    there); for an edge x_J0 + (e_a + e_b) h/2 the four centres J0, J0 + e_a,
    J0 + e_b, J0 + e_a + e_b in equal parts (with delta_b u_a on the edge)."""
    n = 6
    h = 1.0 / n
    a, b = int(ab[0]) - 1, int(ab[1]) - 1
    J0 = np.array([2, 3, 1])
    sig = [np.asfortranarray(np.zeros((n, n, n))) for _ in range(9)]
    sig[NONSYM.index(ab)][tuple(J0)] = 1.0
    u = rnd(n, 5)
    d = O.tensordissipation(sig, u)
    ea, eb = np.eye(3, dtype=int)[a], np.eye(3, dtype=int)[b]
    if a == b:
        expect = {tuple(J0)}
        val = (u[a][tuple(J0)] - u[a][tuple(J0 - ea)]) / h
    else:
        expect = {tuple(J0 + sa * ea + sb * eb) for sa in (0, 1) for sb in (0, 1)}
        val = (u[a][tuple(J0 + eb)] - u[a][tuple(J0)]) / h / 4
    got = {tuple(int(x) for x in I) for I in zip(*np.nonzero(np.abs(d) > 1e-14))}
    assert got == expect
    for I in expect:
        assert abs(d[I] - val) < 1e-12 * abs(val)


@pytest.mark.parametrize("experiment", ["volavg", "project_volavg", "surfavg"])
def test_sfs_tensors_vanish_without_compression(O, experiment):
    """With n_les = n_dns every filter is the identity and the LES pressure
    is the DNS one: all three sub-filter stresses vanish (a window one cell
    off, a missing pressure or a wrong source does not)."""
    n = 6
    u = O.randomfield(n, seed=2, totalenergy=1.0)
    t = O.sfs_tensors(u, n, 0.01, experiment)
    scale = max(np.abs(x).max() for x in O.stresstensor(u, 0.01))
    for k in ("classic", "swapfil", "swapfil_symm"):
        assert max(np.abs(x).max() for x in t[k]) < 1e-12 * scale, k


@pytest.mark.parametrize("experiment,c", [("volavg", 2), ("volavg", 3), ("project_volavg", 2),
                                          ("project_volavg", 3)])
def test_sfs_swapfil_is_the_exact_closure(O, experiment, c):
    """-div^ns sigma_swapfil = V(P du) - P_H du_les: the volume swap-filter
    stress is the exact closure of the volume-filtered LES (P du the projected
    DNS right-hand side, du_les the LES right-hand side of fu), which fixes
    every window and position of the swap filter, its pressure and the sign
    of the LES stress."""
    n = 4 * c if c % 2 == 0 else 5 * c
    nb = n // c
    visc = 0.01
    u = O.randomfield(n, seed=7, totalenergy=1.0)
    t = O.sfs_tensors(u, nb, visc, experiment)
    du = O.rhs(u, visc)
    O.project(du)
    F = O.closure_filters(c, "va")
    fpdu = [O.filter_coarse(du[a], nb, *F[f"ubar{a + 1}"][1]) for a in range(3)]
    dul = O.rhs(t["fu"], visc)
    O.project(dul)
    lhs = O.tensordivergence_nonsym(t["swapfil"])
    for a in range(3):
        assert np.abs(lhs[a] - (fpdu[a] - dul[a])).max() < 1e-11 * np.abs(fpdu[a]).max()
    # the classic stress is not exact there
    cl = O.tensordivergence_nonsym(t["classic"])
    assert max(np.abs(cl[a] - (fpdu[a] - dul[a])).max() / np.abs(fpdu[a]).max() for a in range(3)) > 1e-3


@pytest.mark.parametrize("c", [2, 3])
def test_sfs_volume_filters_tile_the_domain(O, c):
    """The volume filters (of u and of the tensor r) are partitions of unity
    over the DNS grid: their LES means equal the DNS means, and the classic
    filter of the symmetric r is symmetric."""
    n = 4 * c
    nb = n // c
    rng = np.random.default_rng(c)
    src = {k: np.asfortranarray(rng.standard_normal((n, n, n)))
           for k in ("u1", "u2", "u3", "r11", "r22", "r33", "s12", "s23", "s31")}
    q = {}
    for name, (s, fs) in O.sfs_filters(c, "volume").items():
        if name.startswith(("fu", "fr")):
            q[name] = O.filter_coarse(src[s], nb, *fs)
            assert abs(q[name].mean() - src[s].mean()) < 1e-13, name
    for ab in NONSYM:
        assert np.array_equal(q["fr" + ab], q["fr" + ab[::-1]]), ab


@pytest.mark.parametrize("c", [2, 3])
def test_sfs_surface_line_flux_is_exact_across_the_face(O, c):
    """Face average: the transverse part of the swap flux (the line averages
    off the diagonal) is exact, F(-sum_{b != a} delta_b r_ab) =
    -sum_{b != a} delta_b^H (L r)_ab (the face sum of a transverse derivative
    telescopes to the face's edges).  The normal part is not (the face
    average of delta_a r_aa is a fine-scale difference), which is why the
    surface swap-filter LES has an error at all."""
    n = 4 * c if c % 2 == 0 else 5 * c
    nb = n // c
    rng = np.random.default_rng(11 + c)
    z = np.asfortranarray(np.zeros((n, n, n)))
    off = [np.asfortranarray(rng.standard_normal((n, n, n))) for _ in range(3)]  # r12, r23, r31
    du = O.tensordivergence([z, z, z] + off)                                    # -div r, diagonal zero
    F = O.closure_filters(c, "fa")
    Fdu = [O.filter_coarse(du[a], nb, *F[f"ubar{a + 1}"][1]) for a in range(3)]
    src = {"s12": off[0], "s23": off[1], "s31": off[2]}
    q = O.sfs_filters(c, "surface")
    zc = np.asfortranarray(np.zeros((nb, nb, nb)))
    Lr = [zc if ab[0] == ab[1] else O.filter_coarse(src[q["sr" + ab][0]], nb, *q["sr" + ab][1]) for ab in NONSYM]
    lhs = O.tensordivergence_nonsym(Lr)
    for a in range(3):
        assert np.abs(lhs[a] - Fdu[a]).max() < 1e-12 * np.abs(Fdu[a]).max()


def _stagger(source):
    """Fine position offset (in cells) of a source along each axis: 1 at a
    face, 1/2 at a centre (u_a at face a; sigma_ab, a != b, at faces a and b;
    the diagonal r_aa at the centre)."""
    if source.startswith("u"):
        a = int(source[1]) - 1
        return [1.0 if d == a else 0.5 for d in range(3)]
    ab = source[1:]
    a, b = int(ab[0]) - 1, int(ab[1]) - 1
    if a == b:
        return [0.5] * 3
    return [1.0 if d in (a, b) else 0.5 for d in range(3)]


def _region(name, kind):
    """(LES position offset in coarse cells, region width in coarse cells) per
    axis of one sfs_tensors_* quantity, as the filter's definition states it:
    volume filters average the coarse-cell-sized box around the LES position;
    a surface filter is one plane normal to its direction (width 0); the
    swap filters sample the plane normal to the derivative direction b
    (volume) or the face normal to a and, off the diagonal, the line along
    the third direction (surface)."""
    if name.startswith("fu"):
        a = int(name[2]) - 1
        pos = [1.0 if d == a else 0.5 for d in range(3)]
        width = [1.0] * 3 if kind == "volume" else [0.0 if d == a else 1.0 for d in range(3)]
        return pos, width
    ab = name[2:]
    a, b = int(ab[0]) - 1, int(ab[1]) - 1
    pos = [0.5] * 3 if a == b else [1.0 if d in (a, b) else 0.5 for d in range(3)]
    if kind == "volume":
        width = [1.0] * 3 if name.startswith("fr") else [0.0 if d == b else 1.0 for d in range(3)]
    elif name.startswith("fr") or a == b:
        width = [0.0 if d == a else 1.0 for d in range(3)]
    else:
        width = [0.0 if d in (a, b) else 1.0 for d in range(3)]
    return pos, width


def _region_weights(n, c, J, pos, width, s):
    """Trapezoid weights of the fine samples x_i = i + s (cells) over the
    region [X - w c/2, X + w c/2], X = c (J + pos): 1/c inside, 1/(2c) on the
    ends; a plane (w = 0) takes the sample at X or the mean of the two at
    X -+ 1/2."""
    X = c * (J + pos)
    w = {}
    for i in range(-2 * n, 3 * n):
        dx = abs(i + s - X)
        if width > 0:
            half = width * c / 2
            wt = 1.0 / c if dx < half - 1e-9 else (0.5 / c if abs(dx - half) < 1e-9 else 0.0)
        else:
            wt = 1.0 if dx < 1e-9 else (0.5 if abs(dx - 0.5) < 1e-9 else 0.0)
        if wt:
            w[i] = w.get(i, 0.0) + wt
    return w


@pytest.mark.parametrize("kind", ["volume", "surface"])
@pytest.mark.parametrize("c", [2, 3, 4])
def test_sfs_filters_follow_their_regions(O, kind, c):
    """Every quantity of sfs_filters on a field that varies along one axis
    only equals the trapezoid average of the fine samples over the region
    its definition names (position and width per axis), for every axis and
    a random profile: this fixes each window, offset and end weight."""
    n = 3 * c
    nb = n // c
    rng = np.random.default_rng(c)
    for name, (source, fs) in O.sfs_filters(c, kind).items():
        pos, width = _region(name, kind)
        s = _stagger(source if not source.startswith("r") else "r" + name[2:])
        for d in range(3):
            prof = rng.standard_normal(n)
            shape = [1, 1, 1]
            shape[d] = n
            f = np.asfortranarray(np.broadcast_to(prof.reshape(shape), (n, n, n)))
            got = O.filter_coarse(f, nb, *fs)
            idx = [0, 0, 0]
            for J in range(nb):
                idx[d] = J
                wts = _region_weights(n, c, J, pos[d], width[d], s[d])
                assert abs(sum(wts.values()) - 1) < 1e-12, (name, d)
                want = sum(wt * prof[i % n] for i, wt in wts.items())
                assert abs(got[tuple(idx)] - want) < 1e-12, (name, d, J)
