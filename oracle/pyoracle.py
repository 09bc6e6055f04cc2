"""FP64 CPU oracle for the DNS-aided LES hot path (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only to check the product against it.  The product
(``exactclosure.jl_amd``) never imports it.

Fields are Fortran-ordered float64 arrays of shape (n, n, n) indexed [i, j, k]
with x fastest -- the memory layout of a Julia ``Array{Float64,3}`` and of the
GPU path.  The heavy loops live in ``liboracle.so`` (plain C + OpenMP, built by
``oracle/Makefile``); this file restates the reference's composition logic:
filter kernels (src/N.jl:592-637, 1008-1020), closure quantities
(src/N.jl:1022-1135), the DNS-aided LES drivers (src/N.jl:1137-1332) and
``compute_errors`` (src/N.jl:1334-1349).

Parity status of each part is recorded with its tests; the pieces pinned by
reference-independent mathematics are listed in the module-level docstring of
``tests/test_oracle_pins.py``.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        _lib = ctypes.CDLL(_LIB_PATH)
        d = ctypes.c_double
        vp = ctypes.c_void_p
        i = ctypes.c_int
        _lib.or_maxabs.restype = d
        _lib.or_sumsq.restype = d
        _lib.or_burgers_cfl.restype = d
        _lib.or_stresstensor.argtypes = [i, d, d] + [vp] * 9
        _lib.or_tensordivergence.argtypes = [i, d] + [vp] * 9 + [i]
        _lib.or_tensordivergence_nonsym.argtypes = [i, d, vp, vp, vp, vp, i]
        _lib.or_divergence.argtypes = [i, d, vp, vp, vp, vp]
        _lib.or_pressuregradient.argtypes = [i, d, vp, vp, vp, vp]
        _lib.or_project.argtypes = [i, d, vp, vp, vp, vp]
        _lib.or_rfft3.argtypes = [i, vp, vp]
        _lib.or_irfft3.argtypes = [i, vp, vp]
        _lib.or_fft1d.argtypes = [i, i, vp]
        _lib.or_randn.argtypes = [ctypes.c_size_t, ctypes.c_ulonglong, ctypes.c_ulonglong, vp]
        _lib.or_randomfield.argtypes = [i, d, i, d, d, ctypes.c_ulonglong, vp, vp, vp]
        _lib.or_applyfilter.argtypes = [i, vp, i, vp, i, vp, i, vp, vp]
        _lib.or_coarsegrain.argtypes = [i, i, vp, vp, vp]
        _lib.or_filter_coarse.argtypes = [i, i, vp, i, i, vp, i, i, vp, i, i, vp, vp]
        _lib.or_maxabs.argtypes = [ctypes.c_size_t, vp]
        _lib.or_sumsq.argtypes = [ctypes.c_size_t, vp]
        _lib.or_burgers_stress.argtypes = [i, d, d, vp, vp]
        _lib.or_burgers_rhs.argtypes = [i, d, d, vp, vp]
        _lib.or_burgers_rk4.argtypes = [i, d, d, d, i, vp]
        _lib.or_burgers_cfl.argtypes = [i, d, d, vp]
        _lib.or_burgers_filter_coarse.argtypes = [i, i, vp, i, i, vp, vp]
        _lib.or_burgers_dnsaid_rk4.argtypes = [i, i, d, d, d, i, vp, vp, vp]
        _lib.or_burgers_ic.argtypes = [i, i, ctypes.c_ulonglong, vp]
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _f(a):
    """Fortran-ordered float64 copy-or-view."""
    return np.asfortranarray(a, dtype=np.float64)


def field(n):
    return np.zeros((n, n, n), dtype=np.float64, order="F")


# --------------------------------------------------------------------------
# NS operators (src/N.jl:94-307)
# --------------------------------------------------------------------------

def stresstensor(u, visc, l=1.0):
    """stresstensor(u, visc) materialised; order 11, 22, 33, 12, 23, 31 (src/N.jl:94-118)."""
    u = [_f(x) for x in u]
    n = u[0].shape[0]
    s = [field(n) for _ in range(6)]
    lib().or_stresstensor(n, l, visc, *[_p(x) for x in u], *[_p(x) for x in s])
    return s


def tensordivergence(sigma, du=None, doadd=False, l=1.0):
    """tensordivergence!(du, sigma, doadd) (src/N.jl:120-160)."""
    sigma = [_f(x) for x in sigma]
    n = sigma[0].shape[0]
    if du is None:
        du = [field(n) for _ in range(3)]
    lib().or_tensordivergence(n, l, *[_p(x) for x in sigma], *[_p(x) for x in du], int(doadd))
    return du


def tensordivergence_nonsym(sigma9, du=None, doadd=False, l=1.0):
    """tensordivergence_nonsym! (src/N.jl:162-200); order 11,21,31,12,22,32,13,23,33."""
    sigma9 = [_f(x) for x in sigma9]
    n = sigma9[0].shape[0]
    if du is None:
        du = [field(n) for _ in range(3)]
    arr = (ctypes.c_void_p * 9)(*[x.ctypes.data for x in sigma9])
    lib().or_tensordivergence_nonsym(n, l, ctypes.cast(arr, ctypes.c_void_p),
                                     *[_p(x) for x in du], int(doadd))
    return du


def divergence(u, l=1.0):
    u = [_f(x) for x in u]
    n = u[0].shape[0]
    d = field(n)
    lib().or_divergence(n, l, *[_p(x) for x in u], _p(d))
    return d


def pressuregradient(u, p, l=1.0):
    n = p.shape[0]
    lib().or_pressuregradient(n, l, *[_p(x) for x in u], _p(_f(p)))
    return u


def project(u, l=1.0):
    """project!(u, poisson) in place; returns the pressure (src/N.jl:256-307)."""
    for x in u:
        assert x.flags.f_contiguous and x.dtype == np.float64
    n = u[0].shape[0]
    p = field(n)
    lib().or_project(n, l, *[_p(x) for x in u], _p(p))
    return p


def rhs(u, visc, l=1.0):
    """-div(sigma(u)) with doadd = false: stresstensor + tensordivergence!."""
    return tensordivergence(stresstensor(u, visc, l), l=l)


def rfft3(x):
    x = _f(x)
    n = x.shape[0]
    out = np.zeros((n // 2 + 1, n, n), dtype=np.complex128, order="F")
    lib().or_rfft3(n, _p(x), _p(out))
    return out


def irfft3(X):
    X = np.asfortranarray(X, dtype=np.complex128)
    n = X.shape[1]
    out = field(n)
    lib().or_irfft3(n, _p(X), _p(out))
    return out


def spectrum_shells(u, nshell):
    """Energy per integer shell, FP64: getshells (src/N.jl:478-545) and
    spectrum (src/N.jl:576-590).  Shell s holds the half-spectrum bins with
    s^2 <= k1^2 + k2^2 + k3^2 < (s+1)^2 (k1 = x index 0..n/2, k2, k3 from
    wavenumber_int, src/N.jl:32); each bin contributes |u^ / n^3|^2 / 2 per
    component, and the bins 1 <= k1 <= n/2 - 1 count twice (the "conjinds" of
    getshells: their conjugates are not stored by the rfft).  Returns shells
    0..nshell-1."""
    n = u[0].shape[0]
    kmax = n // 2
    k1 = np.arange(kmax + 1)
    kw = np.array([i if i <= (n - 1) // 2 else i - n for i in range(n)])
    kk = k1[:, None, None] ** 2 + kw[None, :, None] ** 2 + kw[None, None, :] ** 2
    shell = np.floor(np.sqrt(kk.astype(np.float64))).astype(np.int64)
    # integer correction of the float square root
    shell[(shell + 1) ** 2 <= kk] += 1
    shell[shell ** 2 > kk] -= 1
    w = np.where((k1 >= 1) & (k1 <= kmax - 1), 2.0, 1.0)[:, None, None]
    e = np.zeros(kk.shape)
    for x in u:
        X = rfft3(x) / float(n) ** 3
        e += np.abs(X) ** 2 / 2
    e *= w
    keep = shell < nshell
    return np.bincount(shell[keep].ravel(), weights=e[keep].ravel(), minlength=nshell)[:nshell]


def burgers_fractions(U, nH, L, visc, les, dbl, parts=False):
    """compute_fractions (src/Burgers.jl:1022-1099) for the DNS samples U
    (nsamples, n), FP64: the LES kernel les = (R, w) and the double kernel
    dbl = (K, w).  Restated with periodic shifts (np.roll): lazyfilter
    y[i] = sum_r w[r + R] x[i + r] (src/Burgers.jl:164-172), stress at centre
    i from faces i - 1 and i (src/Burgers.jl:97-102), lazycoarsegrain at the
    1-based fine index J comp - !stag (comp / 2) (src/Burgers.jl:174-180).
    Returns the mean fractions (classic, flux, div), or with parts=True the
    per-point closure parts (classic, flux, div) of shape (nsamples, nH)."""
    U = np.asarray(U, dtype=np.float64)
    ns, n = U.shape
    comp = n // nH
    a = comp // 2
    h, H = L / n, L / nH
    Rl, wl = les
    K, wd = dbl

    def stress(u, sp):
        ua = np.roll(u, 1, axis=-1)  # u[i - 1]
        return -((ua + u) ** 2) / 8 + visc * (u - ua) / sp

    def filt(x, R, w):
        out = np.zeros_like(x)
        for r in range(-R, R + 1):
            out += w[r + R] * np.roll(x, -r, axis=-1)  # x[i + r]
        return out

    J1 = np.arange(1, nH + 1)
    centre = J1 * comp - a - 1  # 0-based forms of the 1-based coarse-grain indices
    face = J1 * comp - 1
    s = stress(U, h)
    VU = filt(U, K, wd)
    fsu = filt(s, Rl, wl)[:, centre]
    vsu = filt(s, K, wd)[:, centre]
    svu = stress(VU[:, face], H)
    sVU = stress(VU, h)[:, centre]
    cl, fl, dv = vsu - sVU, sVU - svu, fsu - vsu
    if parts:
        return cl, fl, dv
    tot = np.abs(cl) + np.abs(fl) + np.abs(dv) + np.finfo(np.float64).eps
    per = [np.sum(np.abs(x) / tot, axis=1) / nH for x in (cl, fl, dv)]
    return np.array([float(np.mean(x)) for x in per])


def burgers_dissipation(U, nH, L, visc, les, dbl):
    """compute_dissipation (src/Burgers.jl:447-516) for the DNS samples U
    (nsamples, n), FP64: d = sigma * delta_stag(vu) (dissipation,
    src/Burgers.jl:204-208) for sigma_class_m = vsu - sVU, sigma_class_p =
    vsu - svu and sigma_swapfil = fsu - svu, with the filtered quantities of
    compute_fractions (same restatement as burgers_fractions above) and
    delta_stag(vu)_J = (vu_J - vu_{J-1}) / H (src/Burgers.jl:86).  The
    three-argument stress / delta_stag / coarse-grain convolutions that
    compute_dissipation calls are not defined in src/Burgers.jl; they are read
    as these two-argument forms.  Returns (class_m, class_p, swapfil), each of
    shape (nsamples, nH)."""
    U = np.asarray(U, dtype=np.float64)
    ns, n = U.shape
    comp = n // nH
    a = comp // 2
    h, H = L / n, L / nH
    Rl, wl = les
    K, wd = dbl
    J1 = np.arange(1, nH + 1)
    centre = J1 * comp - a - 1
    face = J1 * comp - 1
    s = _bstress(np.roll(U, 1, axis=-1), U, h, visc)
    VU = _bfilt(U, K, wd)
    fsu = _bfilt(s, Rl, wl)[:, centre]
    vsu = _bfilt(s, K, wd)[:, centre]
    vu = VU[:, face]
    vum = np.roll(vu, 1, axis=-1)  # vu[J - 1]
    svu = _bstress(vum, vu, H, visc)
    sVU = _bstress(np.roll(VU, 1, axis=-1), VU, h, visc)[:, centre]
    dv = (vu - vum) / H
    return (vsu - sVU) * dv, (vsu - svu) * dv, (fsu - svu) * dv


def _bstress(ua, ub, sp, visc):
    return -((ua + ub) ** 2) / 8 + visc * (ub - ua) / sp


def _bfilt(x, R, w):
    out = np.zeros_like(x)
    for r in range(-R, R + 1):
        out += w[r + R] * np.roll(x, -r, axis=-1)  # x[i + r]
    return out


def smagorinsky_sums(U, nH, L, visc, delta, les, dbl):
    """The four sums behind theta_classic and theta_discret (smagorinsky.jl:
    24-36) over the fields of smagorinsky_fields (src/Burgers.jl:547-680), FP64:
    <S_h,T_h>, <S_h,S_h>, <S_H,T_H>, <S_H,S_H> over all points and samples."""
    U = np.asarray(U, dtype=np.float64)
    ns, n = U.shape
    comp = n // nH
    a = comp // 2
    h, H = L / n, L / nH
    cs = delta ** 2 + H ** 2
    s = _bstress(np.roll(U, 1, axis=-1), U, h, visc)
    Ud = _bfilt(U, *dbl)
    Um = np.roll(Ud, 1, axis=-1)  # Ud[i - 1]
    Th = _bfilt(s, *dbl) - _bstress(Um, Ud, h, visc)
    dh = (Ud - Um) / h
    Sh = -cs * np.abs(dh) * dh
    Ul, Ur = np.roll(Ud, a + 1, axis=-1), np.roll(Ud, -a, axis=-1)  # Ud[i - a - 1], Ud[i + a]
    TH = _bfilt(s, *les) - _bstress(Ul, Ur, H, visc)
    dH = (Ur - Ul) / H
    SH = -cs * np.abs(dH) * dH
    return np.array([np.sum(Sh * Th), np.sum(Sh * Sh), np.sum(SH * TH), np.sum(SH * SH)])


def smagorinsky_solve(Ustart, nH, L, visc, C, tstop, cfl, dbl):
    """solve_smagorinsky (src/Burgers.jl:682-772) for each sample: the coarse
    face samples of D * u, then the reference's RK4 loop (its axpy! stages
    accumulate on the running state) with dt = min(cfl min(H / max|u|,
    H^2 / visc), tstop - t); FP64."""
    Ustart = np.asarray(Ustart, dtype=np.float64)
    ns, n = Ustart.shape
    comp = n // nH
    H = L / nH
    out = np.empty((ns, nH))

    def rhs(u):
        um = np.roll(u, 1)
        d = (u - um) / H
        sig = _bstress(um, u, H, visc) + C * np.abs(d) * d  # s - (-C |du| du)
        return (np.roll(sig, -1) - sig) / H

    for q in range(ns):
        ub = _bfilt(Ustart[q], *dbl)[comp - 1::comp].copy()  # faces J comp (1-based)
        t = 0.0
        while t < tstop:
            dt = min(cfl * min(H / np.abs(ub).max(), H * H / visc), tstop - t)
            t += dt
            u0 = ub.copy()
            k1 = rhs(ub)
            ub = ub + dt / 2 * k1
            k2 = rhs(ub)
            ub = ub + dt / 2 * k2
            k3 = rhs(ub)
            ub = ub + dt * k3
            k4 = rhs(ub)
            ub = u0 + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        out[q] = ub
    return out


def fft1d(x, sign=-1):
    x = np.array(x, dtype=np.complex128)
    lib().or_fft1d(len(x), sign, _p(x))
    return x


def randn(count, seed, offset=0):
    out = np.empty(count, dtype=np.float64)
    lib().or_randn(count, seed, offset, _p(out))
    return out


def randomfield(n, l=1.0, profile="peak", param=5.0, totalenergy=0.5, seed=123):
    """randomfield (src/N.jl:394-476) with the shared counter-based generator."""
    pid = {"peak": 0, "power": 1}[profile]
    u = [field(n) for _ in range(3)]
    lib().or_randomfield(n, l, pid, param, totalenergy, seed, *[_p(x) for x in u])
    return u


def propose_timestep(u, visc, cfl, l=1.0):
    """propose_timestep (src/N.jl:348-356)."""
    n = u[0].shape[0]
    h = l / n
    umax = max(float(np.max(np.abs(x))) for x in u)
    dt_adv = cfl * h / umax
    dt_visc = cfl * h ** 2 / visc / (2 * 3)
    return min(dt_adv, dt_visc)


def maxabs(x):
    x = np.ascontiguousarray(x.reshape(-1, order="A"))
    return lib().or_maxabs(x.size, _p(x))


# --------------------------------------------------------------------------
# Filter kernels (src/N.jl:592-637, src/Burgers.jl:146-162)
# --------------------------------------------------------------------------

def tophat(compression):
    """tophat (src/N.jl:592-597); odd compression only, like the reference."""
    assert compression % 2 == 1
    return np.full(compression, 1.0 / compression)


def gaussian(h, delta, nsigma):
    """gaussian (src/N.jl:599-607).  Delta = 0 returns [1] (the reference would
    return NaN there; src/Burgers.jl:149-150 special-cases it the same way)."""
    if delta < 1e-15:
        return np.ones(1)
    sigma = delta / math.sqrt(12)
    r = int(np.rint(nsigma * sigma / h))  # Julia round: half to even
    x = h * np.arange(-r, r + 1)
    w = np.exp(-x ** 2 / (2 * sigma ** 2))
    return w / w.sum()


def composekernel(F, G):
    """composekernel (src/N.jl:609-625): full discrete convolution."""
    I = len(F) // 2
    J = len(G) // 2
    K = I + J
    out = np.zeros(2 * K + 1)
    for k in range(-K, K + 1):
        s = 0.0
        for j in range(-J, J + 1):
            i = k - j
            if abs(i) <= I:
                s += F[I + i] * G[J + j]
        out[k + K] = s
    return out


def tophat_weights(comp):
    """tophat_weights (src/Burgers.jl:158-162): (R, w), R = comp div 2, w = 1/comp
    over comp points (odd comp; the reference has no even form)."""
    assert comp % 2 == 1
    return comp // 2, [1.0 / comp] * comp


def trapezoid_weights(comp):
    """The even-compression finite-volume kernel used where the reference's
    tophat_weights would index past its array (c + 1 faces, ends 1/(2c)):
    the box of c cells sampled with the trapezoid rule."""
    assert comp % 2 == 0
    w = [1.0 / comp] * (comp + 1)
    w[0] = w[comp] = 0.5 / comp
    return comp // 2, w


def gaussian_weights(h, delta, nsigma=3):
    """gaussian_weights (src/Burgers.jl:146-156): sigma^2 = (delta/h)^2 / 12 in
    cells, R = round(nsigma sigma) with Julia's ties-to-even, weights
    exp(-r^2 / (2 sigma^2)) normalised; delta < 1e-15 gives [1]."""
    s2 = (delta / h) ** 2 / 12
    R = round(nsigma * math.sqrt(s2))  # Python round: ties to even, as Julia's
    if delta < 1e-15:
        return R, [1.0]
    w = [math.exp(-r * r / (2 * s2)) for r in range(-R, R + 1)]
    t = sum(w)
    return R, [x / t for x in w]


def build_doublekernel(F, G):
    """build_doublekernel (src/Burgers.jl:217-233): the full convolution of the
    two kernels, radius I + J, by the definition sum_j f[k - j] g[j]."""
    I, f = F
    J, g = G
    K = I + J
    out = []
    for k in range(-K, K + 1):
        s = 0.0
        for j in range(-J, J + 1):
            if abs(k - j) <= I:
                s += f[I + k - j] * g[J + j]
        out.append(s)
    return K, out


def burgers_les_kernels(nh, nH, L, delta_ratio, lesfiltertype="tophat"):
    """The LES kernel and the double kernel of compute_fractions
    (src/Burgers.jl:1030-1048): Delta = H delta_ratio; the finite-volume kernel
    over comp = nh div nH fine cells (tophat_weights; the trapezoid for even
    comp); the LES top-hat of 2 (round(Delta/h) div 2) + 1 cells, or the
    Gaussian of gaussian_weights; and their double kernel.  Returns numpy
    arrays."""
    h, H = L / nh, L / nH
    delta = H * delta_ratio
    comp = nh // nH
    fvm = tophat_weights(comp) if comp % 2 else trapezoid_weights(comp)
    if lesfiltertype == "tophat":
        R = round(delta / h) // 2
        les = tophat_weights(2 * R + 1)
    elif lesfiltertype == "gaussian":
        les = gaussian_weights(h, delta)
    else:
        raise ValueError(lesfiltertype)
    dbl = build_doublekernel(fvm, les)
    return (les[0], np.array(les[1])), (dbl[0], np.array(dbl[1]))


def applyfilter(u, w1, w2, w3):
    """applyfilter!(ubar, u, kernelproduct(w1, w2, w3)) (src/N.jl:639-653)."""
    u = _f(u)
    n = u.shape[0]
    out = field(n)
    ws = [np.ascontiguousarray(w, dtype=np.float64) for w in (w1, w2, w3)]
    lib().or_applyfilter(n, _p(ws[0]), len(ws[0]) // 2, _p(ws[1]), len(ws[1]) // 2,
                         _p(ws[2]), len(ws[2]) // 2, _p(u), _p(out))
    return out


def coarsegrain(u, nbar, stag):
    """coarsegrain! (src/N.jl:670-684)."""
    u = _f(u)
    n = u.shape[0]
    out = field(nbar)
    st = (ctypes.c_int * 3)(*[int(bool(s)) for s in stag])
    lib().or_coarsegrain(n, nbar, ctypes.cast(st, ctypes.c_void_p), _p(u), _p(out))
    return out


def filter_coarse(u, nbar, f1, f2, f3):
    """Separable filter evaluated at the coarse points; each f = (weights, offset)
    relative to the coarse face index c*(J+1)-1 (lazycoarsegrain of lazyfilter)."""
    u = _f(u)
    n = u.shape[0]
    out = field(nbar)
    ws = [np.ascontiguousarray(f[0], dtype=np.float64) for f in (f1, f2, f3)]
    lib().or_filter_coarse(n, nbar, _p(ws[0]), len(ws[0]), int(f1[1]), _p(ws[1]), len(ws[1]),
                           int(f2[1]), _p(ws[2]), len(ws[2]), int(f3[1]), _p(u), _p(out))
    return out


# --------------------------------------------------------------------------
# Closure quantities (src/N.jl:999-1135) for any compression
# --------------------------------------------------------------------------
# One-dimensional filter pieces, offsets relative to the coarse face f = c(J+1)-1.

def _box(c):
    return np.full(c, 1.0 / c), -(c - 1)


def _face(c):
    if c % 2:
        return np.full(c, 1.0 / c), -(c // 2)
    w = np.full(c + 1, 1.0 / c)
    w[0] = w[-1] = 0.5 / c
    return w, -(c // 2)


def _centre(c):
    if c % 2:
        return np.ones(1), -(c // 2)
    return np.array([0.5, 0.5]), -(c // 2)


def _point():
    return np.ones(1), 0


def _with(g, piece):
    w, off = piece
    if len(g) == 1:
        return np.asarray(w, dtype=np.float64), off
    return composekernel_any(w, g), off - len(g) // 2


def composekernel_any(w, g):
    return np.convolve(np.asarray(w, dtype=np.float64), np.asarray(g, dtype=np.float64))


def closure_filters(c, kind="va", g=None):
    """1D filters of every closure quantity of one filter set.

    Returns dict name -> (source, [(w, off)] * 3) with sources
    'u1'..'u3', 's11','s22','s33','s12','s23','s31', 'r11','r22','r33' (r = s + p I).

    kind 'va' is the volume average (get_kernels with FH = box/trapezoid and FDelta = g,
    src/N.jl:1008-1019): ubar (GDH at faces), sigDH (get_sigDH, :1088-1115) and the
    face-filtered rDHi (get_rDHi, :1038-1073) in column-major 11,21,...,33 order.
    kind 'fa' is the face (surface) average: ubar_a is the c^2 box of the coarse face,
    r_aa the box at the coarse centre plane pair, r_ab the line average.
    """
    g = np.ones(1) if g is None else np.asarray(g, dtype=np.float64)
    q = {}
    sym = {(0, 0): "s11", (1, 1): "s22", (2, 2): "s33", (0, 1): "s12", (1, 0): "s12",
           (1, 2): "s23", (2, 1): "s23", (2, 0): "s31", (0, 2): "s31"}
    rnm = {(0, 0): "r11", (1, 1): "r22", (2, 2): "r33"}
    if kind == "va":
        for a in range(3):
            q[f"ubar{a + 1}"] = (f"u{a + 1}", [_with(g, _face(c)) if d == a else _with(g, _box(c))
                                             for d in range(3)])
        for (a, b) in [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]:
            if a == b:
                fs = [_with(g, _box(c))] * 3
            else:
                fs = [_with(g, _face(c)) if d in (a, b) else _with(g, _box(c)) for d in range(3)]
            q[f"sigDH{a + 1}{b + 1}"] = (sym[(a, b)], fs)
        for b in range(3):
            for a in range(3):
                fs = []
                for d in range(3):
                    if d == b:
                        fs.append(_with(g, _centre(c)) if a == b else _with(g, _point()))
                    elif d == a:
                        fs.append(_with(g, _face(c)))
                    else:
                        fs.append(_with(g, _box(c)))
                src = rnm[(a, b)] if a == b else sym[(a, b)]
                q[f"rDHi{a + 1}{b + 1}"] = (src, fs)
    elif kind == "fa":
        assert len(g) == 1
        for a in range(3):
            q[f"ubar{a + 1}"] = (f"u{a + 1}", [_point() if d == a else _box(c) for d in range(3)])
        for b in range(3):
            for a in range(3):
                fs = []
                for d in range(3):
                    if a == b:
                        fs.append(_centre(c) if d == a else _box(c))
                    else:
                        fs.append(_point() if d in (a, b) else _box(c))
                src = rnm[(a, b)] if a == b else sym[(a, b)]
                q[f"rDHi{a + 1}{b + 1}"] = (src, fs)
    else:
        raise ValueError(kind)
    return q


NONSYM_ORDER = ["11", "21", "31", "12", "22", "32", "13", "23", "33"]


def closure_sources(u, p, visc, l=1.0):
    s11, s22, s33, s12, s23, s31 = stresstensor(u, visc, l)
    return {"u1": _f(u[0]), "u2": _f(u[1]), "u3": _f(u[2]), "s11": s11, "s22": s22,
            "s33": s33, "s12": s12, "s23": s23, "s31": s31,
            "r11": s11 + p, "r22": s22 + p, "r33": s33 + p}


def closure_quantities(u, p, visc, nbar, c, kind="va", g=None, l=1.0, names=None):
    src = closure_sources(u, p, visc, l)
    out = {}
    for name, (s, fs) in closure_filters(c, kind, g).items():
        if names is not None and name not in names:
            continue
        out[name] = filter_coarse(src[s], nbar, *fs)
    return out


def les_stress_sym(u, visc, l=1.0):
    return stresstensor(u, visc, l)


def commutator(q, ubar, visc, l=1.0):
    """tau_ab = (S r)_ab - sigma_H(ubar)_ab, nonsym order (the exact-closure flux)."""
    s = stresstensor(ubar, visc, l)
    sym = {"11": s[0], "22": s[1], "33": s[2], "12": s[3], "21": s[3], "23": s[4],
           "32": s[4], "31": s[5], "13": s[5]}
    return [q["rDHi" + ab] - sym[ab] for ab in NONSYM_ORDER]


# --------------------------------------------------------------------------
# Sub-filter stresses and their dissipation (sfs_tensors_volume and
# sfs_tensors_surface, src/NavierStokes.jl:470-593; process.jl:107-193).
# The filters there are Turbulox's, a package outside the reference tree:
# they are restated as the volume, face and line averages of the pieces above.
# --------------------------------------------------------------------------

def sfs_filters(c, kind):
    """name -> (source, [(w, off)] * 3) of the DNS quantities of one
    sfs_tensors_* call:
      fu{a}   the LES velocity filter (kind 'volume': volumefilter!,
              'surface': surfacefilter!);
      fr{ab}  the classic filter of r = sigma + p I (volume: volumefilter!
              of the tensor; surface: surfacefilter!(r, true), the face
              normal to a, the filter of u_a);
      sr{ab}  the exact-closure flux filter (volume: surfacefilter!(r,
              false), faces normal to the derivative direction b; surface:
              the diagonal of fr and linefilter!(r) off the diagonal), i.e.
              the rDHi of closure_filters.
    Tensors in nonsym order 11, 21, 31, 12, 22, 32, 13, 23, 33."""
    base = closure_filters(c, "va" if kind == "volume" else "fa")
    sym = {(0, 1): "s12", (1, 0): "s12", (1, 2): "s23", (2, 1): "s23", (2, 0): "s31", (0, 2): "s31"}
    q = {}
    for a in range(3):
        q[f"fu{a + 1}"] = base[f"ubar{a + 1}"]
    for ab in NONSYM_ORDER:
        a, b = int(ab[0]) - 1, int(ab[1]) - 1
        if a == b:
            fs = [_box(c)] * 3 if kind == "volume" else base[f"rDHi{ab}"][1]
            q[f"fr{ab}"] = (f"r{ab}", fs)
        elif kind == "volume":
            q[f"fr{ab}"] = (sym[(a, b)], [_face(c) if d in (a, b) else _box(c) for d in range(3)])
        else:
            q[f"fr{ab}"] = (sym[(a, b)], [_point() if d == a else (_face(c) if d == b else _box(c))
                                          for d in range(3)])
        q[f"sr{ab}"] = base[f"rDHi{ab}"]
    return q


def sfs_tensors(u, nbar, visc, experiment, l=1.0):
    """sfs_tensors_volume (experiment 'volavg': doproject false;
    'project_volavg': doproject true) and sfs_tensors_surface ('surfavg')
    (src/NavierStokes.jl:470-593): the DNS pressure from the projected
    right-hand side, the filtered velocity fu (projected for
    'project_volavg'), the LES pressure of fu the same way, and
    sigma_classic = fr - r_H(fu), sigma_swapfil = sr - r_H(fu) and its
    symmetric part, r_H = sigma_H + p_les I.  Returns dict fu (3),
    classic, swapfil, swapfil_symm (9 each, nonsym order)."""
    u = [_f(x) for x in u]
    n = u[0].shape[0]
    c = n // nbar
    assert c * nbar == n
    kind = "surface" if experiment == "surfavg" else "volume"
    assert experiment in ("volavg", "project_volavg", "surfavg"), experiment
    du = rhs(u, visc, l)
    p = project(du, l)
    src = closure_sources(u, p, visc, l)
    q = {name: filter_coarse(src[s], nbar, *fs) for name, (s, fs) in sfs_filters(c, kind).items()}
    fu = [q["fu1"], q["fu2"], q["fu3"]]
    if experiment == "project_volavg":
        project(fu, l)
    dul = rhs(fu, visc, l)
    pl = project(dul, l)
    s = stresstensor(fu, visc, l)
    rh = {"11": s[0] + pl, "22": s[1] + pl, "33": s[2] + pl, "12": s[3], "21": s[3], "23": s[4], "32": s[4],
          "31": s[5], "13": s[5]}
    classic = [q["fr" + ab] - rh[ab] for ab in NONSYM_ORDER]
    swapfil = [q["sr" + ab] - rh[ab] for ab in NONSYM_ORDER]
    m = {ab: i for i, ab in enumerate(NONSYM_ORDER)}
    symm = [(swapfil[m[ab]] + swapfil[m[ab[::-1]]]) / 2 for ab in NONSYM_ORDER]
    return {"fu": fu, "classic": classic, "swapfil": swapfil, "swapfil_symm": symm}


def tensordissipation(sigma9, u, l=1.0):
    """Turbulox.tensordissipation_staggered!(d, sigma, u) (process.jl:160-180),
    restated: d = sigma : grad u at the cell centres of the staggered grid.
    sigma_aa delta_a u_a is formed at the centre; sigma_ab delta_b u_a at its
    edge (a, b) (x_J + (e_a + e_b) h / 2, where tensordivergence_nonsym reads
    sigma_ab) and averaged over the four edges around the centre."""
    u = [_f(x) for x in u]
    n = u[0].shape[0]
    h = l / n
    S = dict(zip(NONSYM_ORDER, [_f(x) for x in sigma9]))
    d = field(n)
    for a in range(3):
        d += S[f"{a + 1}{a + 1}"] * ((u[a] - np.roll(u[a], 1, axis=a)) / h)
        for b in range(3):
            if b == a:
                continue
            e = S[f"{a + 1}{b + 1}"] * ((np.roll(u[a], -1, axis=b) - u[a]) / h)
            d += (e + np.roll(e, 1, axis=a) + np.roll(e, 1, axis=b) + np.roll(np.roll(e, 1, axis=a), 1, axis=b)) / 4
    return d


def sfs_dissipation(u, nbar, visc, experiment, l=1.0):
    """The dissipation fields of process.jl:107-193 for one LES grid and
    experiment: tensordissipation of sigma_classic, sigma_swapfil and
    sigma_swapfil_symm with the filtered velocity (the kernel density
    estimate of process.jl is not part of this)."""
    t = sfs_tensors(u, nbar, visc, experiment, l)
    return {k: tensordissipation(t[k], t["fu"], l) for k in ("classic", "swapfil", "swapfil_symm")}


def compute_errors(ubar, uref):
    """compute_errors (src/N.jl:1334-1349) for one LES field."""
    D = 3
    s = sum(float(np.sum((a - b) ** 2)) / float(np.sum(b ** 2)) for a, b in zip(ubar, uref))
    return math.sqrt(s / D)


# --------------------------------------------------------------------------
# Drivers
# --------------------------------------------------------------------------

def step_forwardeuler(u, visc, dt, l=1.0):
    """step_forwardeuler! (src/N.jl:309-318)."""
    du = rhs(u, visc, l)
    for a in range(3):
        u[a] += dt * du[a]
    project(u, l)
    return u


def step_wray3(u, visc, dt, l=1.0):
    """step_wray3! (src/N.jl:320-346)."""
    a = (8 / 15, 5 / 12, 3 / 4)
    b = (1 / 4, 0 / 1)
    u0 = [x.copy(order="F") for x in u]
    for i in range(3):
        du = rhs(u, visc, l)
        for c in range(3):
            if i != 0:
                u[c][...] = u0[c]
            u[c] += a[i] * dt * du[c]
        project(u, l)
        if i != 2:
            for c in range(3):
                u0[c] += b[i] * dt * du[c]
    return u


RK4_A = (0.5, 0.5, 1.0)
RK4_B = (1.0, 2.0, 2.0, 1.0)


def les_rhs_exact(u_les, tau, visc, l=1.0):
    """-div_H sigma(u_les) - div_H^ns tau, then projected (step_DHpi_classic_flux_div!,
    src/N.jl:1189-1206, with tau = rDHi - sigma_H(ubar))."""
    du = tensordivergence(stresstensor(u_les, visc, l), l=l)
    du = tensordivergence_nonsym(tau, du, doadd=True, l=l)
    p = project(du, l)
    return du, p


def dnsaid_rk4(u, sets, visc, dt, nsteps, l=1.0, les=None, closure="flux"):
    """DNS-aided LES with classical RK4 and the exact closure at every stage.

    sets: list of (c, kind).  For each set an LES field starts from
    P_H(V u) and is advanced with k_les = P_H(-div_H sigma(u_les) - div^ns_H tau),
    tau = S r(u_s) - sigma_H(P_H V u_s), r = sigma + p I with p from the DNS
    right-hand-side projection (src/N.jl:1305-1319).  Returns (u, les, taus) with
    taus the closure of the last stage evaluated.

    closure="forcing": the same closure in divergence form, the LES driven by
    F_H(u_les) + m with m = V(k) - F_H(P_H V u_s), k the projected DNS
    right-hand side (taus then holds the three components of m).
    """
    n = u[0].shape[0]
    names = ["ubar1", "ubar2", "ubar3"] + ["rDHi" + ab for ab in NONSYM_ORDER]

    def filtered(us, p, c, kind):
        q = closure_quantities(us, p, visc, n // c, c, kind, l=l, names=names)
        ub = [q["ubar1"].copy(order="F"), q["ubar2"].copy(order="F"), q["ubar3"].copy(order="F")]
        project(ub, l)
        return q, ub

    if les is None:
        les = []
        for (c, kind) in sets:
            ub = [filter_coarse(_f(u[a]), n // c, *closure_filters(c, kind)[f"ubar{a + 1}"][1])
                  for a in range(3)]
            project(ub, l)
            les.append(ub)
    taus = None
    for _ in range(nsteps):
        u0 = [x.copy(order="F") for x in u]
        l0 = [[x.copy(order="F") for x in v] for v in les]
        acc = [field(n) for _ in range(3)]
        lacc = [[np.zeros_like(x) for x in v] for v in les]
        for st in range(4):
            k = rhs(u, visc, l)
            p = project(k, l)
            taus = []
            kl = []
            for si, (c, kind) in enumerate(sets):
                if closure == "forcing":
                    F = closure_filters(c, kind)
                    ub = [filter_coarse(_f(u[a]), n // c, *F[f"ubar{a + 1}"][1]) for a in range(3)]
                    project(ub, l)
                    vk = [filter_coarse(_f(k[a]), n // c, *F[f"ubar{a + 1}"][1]) for a in range(3)]
                    fh = rhs(ub, visc, l)
                    m = [vk[a] - fh[a] for a in range(3)]
                    taus.append(m)
                    kk = rhs(les[si], visc, l)
                    kk = [kk[a] + m[a] for a in range(3)]
                    project(kk, l)
                    kl.append(kk)
                    continue
                q, ub = filtered(u, p, c, kind)
                tau = commutator(q, ub, visc, l)
                taus.append(tau)
                kk, _ = les_rhs_exact(les[si], tau, visc, l)
                kl.append(kk)
            for a in range(3):
                acc[a] += RK4_B[st] * k[a]
                for si in range(len(sets)):
                    lacc[si][a] += RK4_B[st] * kl[si][a]
            if st < 3:
                for a in range(3):
                    u[a][...] = u0[a] + RK4_A[st] * dt * k[a]
                    for si in range(len(sets)):
                        les[si][a][...] = l0[si][a] + RK4_A[st] * dt * kl[si][a]
            else:
                for a in range(3):
                    u[a][...] = u0[a] + dt / 6 * acc[a]
                    for si in range(len(sets)):
                        les[si][a][...] = l0[si][a] + dt / 6 * lacc[si][a]
    return u, les, taus


def apply_DHpi(u, nbar, c, g=None, l=1.0):
    """apply_DHpi (src/N.jl:1228-1237): filter, coarse-grain at faces, project."""
    q = closure_filters(c, "va", g)
    ub = [filter_coarse(_f(u[a]), nbar, *q[f"ubar{a + 1}"][1]) for a in range(3)]
    project(ub, l)
    return ub


def dnsaid_project(u, n_les, visc, delta_ratio=2.0, nsigma=2, twarm=0.1, tstop=0.1, cfl=0.4, l=1.0):
    """dnsaid_project (src/N.jl:1239-1332) with its time control: forward-Euler
    DNS warm-up to twarm, then the DNS and the five LES to tstop, each step
    dt = min(propose_timestep(u_dns, visc, cfl), t_end - t).  Returns the LES
    fields as dnsaid_project_fe and the number of steps of the main loop."""
    t = 0.0
    while t < twarm:
        dt = min(propose_timestep(u, visc, cfl, l), twarm - t)
        t += dt
        step_forwardeuler(u, visc, dt, l)
    state = {"t": 0.0, "k": 0}

    def next_dt(uu):
        if state["t"] >= tstop:
            return None
        dt = min(propose_timestep(uu, visc, cfl, l), tstop - state["t"])
        state["t"] += dt
        state["k"] += 1
        return dt

    les = dnsaid_project_fe(u, n_les, visc, None, next_dt, delta_ratio, nsigma, l)
    return les, state["k"]


def dnsaid_project_fe(u, n_les, visc, nsteps, dt, delta_ratio=0.0, nsigma=2, l=1.0):
    """dnsaid_project main loop (src/N.jl:1296-1332) with fixed dt and nsteps
    (the reference chooses dt by propose_timestep and stops at t = 0.1; a
    callable dt(u) giving the next step or None to stop does that, see
    dnsaid_project).

    Returns the five LES fields and the filtered DNS, keyed like the reference:
    u_les, u_nomo, u_c, u_cf, u_cfd, u_cfd_symm.
    """
    n = u[0].shape[0]
    c = n // n_les
    h = l / n
    H = l / n_les
    g = gaussian(h, delta_ratio * H, nsigma)
    filt = closure_filters(c, "va", g)
    ubar = apply_DHpi(u, n_les, c, g, l)
    les = {k: [x.copy(order="F") for x in ubar] for k in ("u_nomo", "u_c", "u_cf", "u_cfd",
                                                          "u_cfd_symm")}
    # FDH as a full-grid kernel for the "classic" variant's stress of the
    # fine-grid filtered field (src/N.jl:1143-1145).
    if c % 2:
        FDH = composekernel(tophat(c), g)
    else:
        FDH = None
    step_dt = dt if callable(dt) else None
    k = 0
    while True:
        if step_dt is not None:
            dt = step_dt(u)
            if dt is None:
                break
        elif k >= nsteps:
            break
        k += 1
        du_dns = rhs(u, visc, l)
        p = project(du_dns, l)
        src = closure_sources(u, p, visc, l)
        ubar = apply_DHpi(u, n_les, c, g, l)
        q = {name: filter_coarse(src[s], n_les, *fs) for name, (s, fs) in filt.items()
             if not name.startswith("ubar")}
        sigDH = [q["sigDH11"], q["sigDH22"], q["sigDH33"], q["sigDH12"], q["sigDH23"],
                 q["sigDH31"]]
        # no model
        step_forwardeuler(les["u_nomo"], visc, dt, l)
        # classic: sigma_les + sigDH - coarse(sigma(filter(u)))
        if FDH is not None:
            uf = [applyfilter(x, FDH, FDH, FDH) for x in u]
            sf = stresstensor(uf, visc, l)
            pos = [(0, 0, 0), (0, 0, 0), (0, 0, 0), (1, 1, 0), (0, 1, 1), (1, 0, 1)]
            s2 = [coarsegrain(sf[m], n_les, pos[m]) for m in range(6)]
            sl = stresstensor(les["u_c"], visc, l)
            du = tensordivergence([sl[m] + sigDH[m] - s2[m] for m in range(6)], l=l)
            project(du, l)
            for a in range(3):
                les["u_c"][a] += dt * du[a]
        # classic + flux: sigma_les + sigDH - sigma_H(ubar)
        sl = stresstensor(les["u_cf"], visc, l)
        s2 = stresstensor(ubar, visc, l)
        du = tensordivergence([sl[m] + sigDH[m] - s2[m] for m in range(6)], l=l)
        project(du, l)
        for a in range(3):
            les["u_cf"][a] += dt * du[a]
        # classic + flux + div (exact)
        rD = [q["rDHi" + ab] for ab in NONSYM_ORDER]
        sl = stresstensor(les["u_cfd"], visc, l)
        du = tensordivergence([sl[m] - s2[m] for m in range(6)], l=l)
        du = tensordivergence_nonsym(rD, du, doadd=True, l=l)
        project(du, l)
        for a in range(3):
            les["u_cfd"][a] += dt * du[a]
        # symmetrised
        sl = stresstensor(les["u_cfd_symm"], visc, l)
        du = tensordivergence([sl[m] - s2[m] for m in range(6)], l=l)
        R = dict(zip(NONSYM_ORDER, rD))
        symm = [R["11"], R["22"], R["33"], (R["21"] + R["12"]) / 2, (R["32"] + R["23"]) / 2,
                (R["31"] + R["13"]) / 2]
        du = tensordivergence(symm, du, doadd=True, l=l)
        project(du, l)
        for a in range(3):
            les["u_cfd_symm"][a] += dt * du[a]
        # DNS
        for a in range(3):
            u[a] += dt * du_dns[a]
    les["u_les"] = apply_DHpi(u, n_les, c, g, l)
    return les


# --------------------------------------------------------------------------
# Burgers (src/Burgers.jl)
# --------------------------------------------------------------------------

def burgers_stress(u, visc, L=2 * math.pi):
    u = np.ascontiguousarray(u, dtype=np.float64)
    s = np.empty_like(u)
    lib().or_burgers_stress(len(u), L, visc, _p(u), _p(s))
    return s


def burgers_rhs(u, visc, L=2 * math.pi):
    u = np.ascontiguousarray(u, dtype=np.float64)
    du = np.empty_like(u)
    lib().or_burgers_rhs(len(u), L, visc, _p(u), _p(du))
    return du


def burgers_rk4(u, visc, dt, nsteps, L=2 * math.pi):
    u = np.array(u, dtype=np.float64)
    lib().or_burgers_rk4(len(u), L, visc, dt, nsteps, _p(u))
    return u


def burgers_cfl(u, visc, L=2 * math.pi):
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().or_burgers_cfl(len(u), L, visc, _p(u))


def burgers_ic(n, seed, kmax=10):
    u = np.empty(n, dtype=np.float64)
    lib().or_burgers_ic(n, kmax, seed, _p(u))
    return u


def burgers_filter_coarse(u, nbar, w, off):
    u = np.ascontiguousarray(u, dtype=np.float64)
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.empty(nbar)
    lib().or_burgers_filter_coarse(len(u), nbar, _p(w), len(w), off, _p(u), _p(out))
    return out


def burgers_dnsaid_rk4(u, c, visc, dt, nsteps, L=2 * math.pi):
    u = np.array(u, dtype=np.float64)
    nH = len(u) // c
    us = np.empty(nH)
    um = np.empty(nH)
    lib().or_burgers_dnsaid_rk4(len(u), c, L, visc, dt, nsteps, _p(u), _p(us), _p(um))
    return u, us, um


_M64 = (1 << 64) - 1


def _sm64(x):
    x = (x + 0x9E3779B97F4A7C15) & _M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & _M64
    return x ^ (x >> 31)


def _phase(seed, k):
    """The counter-based uniform phase of mode k for a sample seed (splitmix64
    finaliser of seed ^ finaliser(k), top 53 bits)."""
    return (_sm64((seed ^ _sm64(k)) & _M64) >> 11) * 2.0 ** -53


def burgers_randomfield(n, kpeak, totalenergy, seed, L=2 * math.pi):
    """randomfield with bump_profile (src/Burgers.jl:192-202): c_k =
    (k/kp)^4 exp(-2 (k/kp)^2 + 2 pi i phi_k) for k = 0 .. n/2, a = irfft(c n,
    n) (numpy; the whole spectrum), u = sqrt(E / etot) a with etot = sum a^2 /
    2 h.  The phases phi_k replace Julia's rand(Xoshiro(i)) draws."""
    k = np.arange(n // 2 + 1)
    amp = (k / kpeak) ** 4 * np.exp(-2 * (k / kpeak) ** 2)
    ph = np.array([_phase(seed, int(kk)) for kk in k])
    a = np.fft.irfft(amp * np.exp(2j * np.pi * ph) * n, n)
    etot = np.sum(a ** 2) / 2 * (L / n)
    return np.sqrt(totalenergy / etot) * a


def burgers_dns_aided_les(ustart, nH, L, visc, les, dbl, tstop, cfl_factor=0.3):
    """dns_aided_les (src/Burgers.jl:350-445) for one DNS initial field, FP64:
    forward Euler with dt = min(cfl_factor min(h / max|u|, h^2 / visc), tstop -
    t); per step the DNS stress s, fsu = C_c(F * s), vsu = C_c(D * s), vu =
    C_f(D * u), svu = s_H(vu), sVU = C_c(s_h(D * u)); the LES stresses of
    nomodel (s_H(u)), class_m (+ vsu - sVU), class_p (+ vsu - svu) and
    swapfil (+ fsu - svu) and stepfromstress! for every LES and then the DNS
    (src/Burgers.jl:338-347).  Filters and coarse-grain indices as in
    burgers_fractions.  Returns (fields, steps)."""
    u = np.array(ustart, dtype=np.float64)
    n = len(u)
    comp = n // nH
    a = comp // 2
    h, H = L / n, L / nH
    Rl, wl = les
    K, wd = dbl

    def stress(x, sp):
        xa = np.roll(x, 1)
        return -((xa + x) ** 2) / 8 + visc * (x - xa) / sp

    def filt(x, R, w):
        out = np.zeros_like(x)
        for r in range(-R, R + 1):
            out += w[r + R] * np.roll(x, -r)
        return out

    J1 = np.arange(1, nH + 1)
    centre = J1 * comp - a - 1
    face = J1 * comp - 1
    uH = filt(u, K, wd)[face]
    les_u = {k: uH.copy() for k in ("nomodel", "class_m", "class_p", "swapfil")}
    t, steps = 0.0, 0
    while t < tstop:
        dt = cfl_factor * min(h / np.abs(u).max(), h * h / visc)
        dt = min(dt, tstop - t)
        t += dt
        steps += 1
        s = stress(u, h)
        fsu = filt(s, Rl, wl)[centre]
        vsu = filt(s, K, wd)[centre]
        VU = filt(u, K, wd)
        vu = VU[face]
        svu = stress(vu, H)
        sVU = stress(VU, h)[centre]
        sig = {"nomodel": stress(les_u["nomodel"], H),
               "class_m": stress(les_u["class_m"], H) + vsu - sVU,
               "class_p": stress(les_u["class_p"], H) + vsu - svu,
               "swapfil": stress(les_u["swapfil"], H) + fsu - svu}
        for key, sg in sig.items():
            les_u[key] = les_u[key] + dt * ((np.roll(sg, -1) - sg) / H)
        u = u + dt * ((np.roll(s, -1) - s) / h)
    VU = filt(u, K, wd)
    out = {"dns_ref": u, "dns_ref_fil": VU, "dns_fil": VU[face]}
    out.update(les_u)
    return out, steps
