"""Drop-in mirror of the ``NavierStokes`` module of ExactClosure.jl (src/N.jl).

Same function names and argument meaning as the reference; Julia's ``!``
suffix becomes a trailing underscore (``project!`` -> ``project_``).  Every
loop runs in libexactclosure on the GPU; fields are torch tensors used only
as device memory.  There is no CPU fallback: without the library or a GPU
the calls raise.

Layout: a ``Field`` holds a contiguous tensor of shape (n, n, n) indexed
[k, j, i] -- x fastest, identical in memory to a Julia ``Array{T,3}``.
``Field.numpy()`` returns the Julia-ordered [i, j, k] view on the host.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

from . import _abi
from ._abi import P3, P6, P9, ptr, ptrs

# --------------------------------------------------------------------------
# Grid and fields (src/N.jl:22-92)
# --------------------------------------------------------------------------


@dataclass(frozen=True)
class Grid:
    """Grid{D}(l, n) (src/N.jl:22-32); D = 3."""

    l: float
    n: int
    D: int = 3


def spacing(g):
    return g.l / g.n


def dim(g):
    return g.D


def wavenumber_int(g, i):
    """src/N.jl:32 with the reference's 1-based index i."""
    return i - 1 - (0 if i <= (g.n + 1) >> 1 else g.n)


class Field:
    """Scalar field on a grid (src/N.jl:35-43), device storage."""

    def __init__(self, grid, data=None, dtype=None):
        import torch

        self.grid = grid
        if data is None:
            dtype = dtype or torch.float64
            data = torch.zeros((grid.n,) * 3, dtype=dtype, device="cuda")
        assert data.is_contiguous() and tuple(data.shape) == (grid.n,) * 3
        self.data = data

    @property
    def dtype(self):
        return self.data.dtype

    def numpy(self):
        """Host copy in Julia index order [i, j, k] (Fortran-ordered float64)."""
        return np.asfortranarray(self.data.double().cpu().numpy().transpose(2, 1, 0))

    @staticmethod
    def from_numpy(grid, a, dtype=None):
        """Field from a Julia-ordered [i, j, k] host array."""
        import torch

        t = torch.from_numpy(np.ascontiguousarray(np.asarray(a).transpose(2, 1, 0)))
        return Field(grid, t.to(device="cuda", dtype=dtype or torch.float64).contiguous())

    def __repr__(self):
        return f"Field({self.grid}, {self.dtype})"


def vectorfield(g, dtype=None):
    return tuple(Field(g, dtype=dtype) for _ in range(3))


def tensorfield(g, dtype=None):
    return tuple(Field(g, dtype=dtype) for _ in range(6))


def _dt(f):
    return _abi.dtype_code(f.data)


def _ctx():
    return _abi.context(0)


# --------------------------------------------------------------------------
# Operators (src/N.jl:94-307)
# --------------------------------------------------------------------------


def stresstensor(u, visc):
    """Materialised stresstensor(u, visc): (11, 22, 33, 12, 23, 31) (src/N.jl:94-118).

    The reference returns lazy fields; tensordivergence_ fuses the same
    evaluation without materialising (see rhs_)."""
    g = u[0].grid
    s = tuple(Field(g, dtype=u[0].dtype) for _ in range(6))
    _ctx().call("ec_ns_stresstensor", _dt(u[0]), g.n, g.l, visc, ptrs([x.data for x in u], P3),
                ptrs([x.data for x in s], P6))
    return s


def tensordivergence_(du, sigma, doadd):
    """tensordivergence!(du, sigma, doadd) (src/N.jl:120-160)."""
    g = du[0].grid
    _ctx().call("ec_ns_tensordivergence", _dt(du[0]), g.n, g.l, ptrs([x.data for x in sigma], P6),
                ptrs([x.data for x in du], P3), int(bool(doadd)))
    return du


def tensordivergence_nonsym_(du, sigma9, doadd):
    """tensordivergence_nonsym!(du, sigma, doadd) (src/N.jl:162-200); sigma may be FP64."""
    g = du[0].grid
    sd = _abi.dtype_code(sigma9[0].data if isinstance(sigma9[0], Field) else sigma9[0])
    data = [x.data if isinstance(x, Field) else x for x in sigma9]
    _ctx().call("ec_ns_tensordivergence_nonsym", _dt(du[0]), g.n, g.l, sd, ptrs(data, P9),
                ptrs([x.data for x in du], P3), int(bool(doadd)))
    return du


def rhs_(du, u, visc, doadd=False):
    """tensordivergence!(du, stresstensor(u, visc), doadd) fused in one kernel."""
    g = u[0].grid
    _ctx().call("ec_ns_rhs", _dt(u[0]), g.n, g.l, visc, ptrs([x.data for x in u], P3),
                ptrs([x.data for x in du], P3), int(bool(doadd)))
    return du


def divergence_(d, u):
    """divergence!(d, u) (src/N.jl:222-238)."""
    g = d.grid
    _ctx().call("ec_ns_divergence", _dt(d), g.n, g.l, ptrs([x.data for x in u], P3), ptr(d.data))
    return d


def pressuregradient_(u, p):
    """pressuregradient!(u, p) (src/N.jl:240-254)."""
    g = p.grid
    _ctx().call("ec_ns_pressuregradient", _dt(p), g.n, g.l, ptrs([x.data for x in u], P3), ptr(p.data))
    return u


class PoissonSolver:
    """poissonsolver(grid, backend) (src/N.jl:202-220): FFT plan, spectrum and p."""

    def __init__(self, grid, dtype=None):
        import torch

        dtype = dtype or torch.float64
        self.grid = grid
        self.torch_dtype = dtype
        self.ctx = _ctx()
        h = _abi.c_void_p()
        code = _abi.EC_F64 if dtype == torch.float64 else _abi.EC_F32
        self.ctx.call("ec_poisson_create", code, grid.n, grid.l, _abi.ctypes.byref(h))
        self.h = h
        n = grid.n
        # p aliases the solver's own pressure buffer (poisson.p of the reference)
        self._p_ptr = self.ctx.lib.ec_poisson_pressure(h)
        self._ptensor = None

    @property
    def p(self):
        """The solver's pressure as a Field (a copy synchronised with the stream)."""
        import torch

        n = self.grid.n
        t = torch.empty((n, n, n), dtype=self.torch_dtype, device="cuda")
        self.ctx.call("ec_d2d", _abi.c_void_p(t.data_ptr()), _abi.c_void_p(self._p_ptr),
                      t.numel() * t.element_size())
        return Field(self.grid, t)

    @property
    def p_ptr(self):
        return self._p_ptr

    def close(self):
        if self.h:
            self.ctx.lib.ec_poisson_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poissonsolver(grid, dtype=None):
    return PoissonSolver(grid, dtype)


def project_(u, poisson):
    """project!(u, poisson) (src/N.jl:256-307); the pressure stays in poisson.p."""
    _ctx().call("ec_ns_project", poisson.h, ptrs([x.data for x in u], P3))
    return u


def rfft(x, poisson):
    """plan_rfft * x (unnormalised), shape (n, n, n/2+1) complex tensor [kz, ky, kx]."""
    import torch

    n = poisson.grid.n
    _check_packed(x, (n, n, n), poisson)
    cdt = torch.complex128 if x.dtype == torch.float64 else torch.complex64
    out = torch.empty((n, n, n // 2 + 1), dtype=cdt, device="cuda")
    _ctx().call("ec_fft_r2c", poisson.h, ptr(x), ptr(out))
    return out


def _check_packed(t, shape, poisson):
    """The library reads and writes packed (contiguous) arrays of the solver's
    precision; a strided view (e.g. a spectrum from another FFT library) is
    refused rather than read with the wrong layout."""
    import torch

    if tuple(t.shape) != shape or not t.is_contiguous():
        raise ValueError(f"expected a contiguous {shape} tensor, got shape {tuple(t.shape)} "
                         f"strides {t.stride()} (call .contiguous())")
    real = torch.float64 if poisson.torch_dtype == torch.float64 else torch.float32
    if t.dtype not in (real, torch.complex128 if real == torch.float64 else torch.complex64):
        raise ValueError(f"dtype {t.dtype} does not match the solver's {poisson.torch_dtype}")


def irfft(X, poisson):
    """ldiv!(x, plan, X): normalised inverse."""
    import torch

    n = poisson.grid.n
    _check_packed(X, (n, n, n // 2 + 1), poisson)
    rdt = torch.float64 if X.dtype == torch.complex128 else torch.float32
    out = torch.empty((n, n, n), dtype=rdt, device="cuda")
    _ctx().call("ec_fft_c2r", poisson.h, ptr(X), ptr(out))
    return out


def propose_timestep(u, visc, cfl):
    """propose_timestep (src/N.jl:348-356), FP64 device reduction."""
    g = u[0].grid
    h = spacing(g)
    umax = max(maxabs(x) for x in u)
    return min(cfl * h / umax, cfl * h ** 2 / visc / (2 * 3))


def maxabs(f):
    out = _abi.c_double()
    t = f.data if isinstance(f, Field) else f
    _ctx().call("ec_reduce_maxabs", _abi.dtype_code(t), t.numel(), ptr(t), _abi.ctypes.byref(out))
    return out.value


def sumsq(f):
    out = _abi.c_double()
    t = f.data if isinstance(f, Field) else f
    _ctx().call("ec_reduce_sumsq", _abi.dtype_code(t), t.numel(), ptr(t), _abi.ctypes.byref(out))
    return out.value


def diffsq(a, b):
    out = _abi.c_double()
    ta = a.data if isinstance(a, Field) else a
    tb = b.data if isinstance(b, Field) else b
    _ctx().call("ec_reduce_diffsq", _abi.dtype_code(ta), ta.numel(), ptr(ta), ptr(tb), _abi.ctypes.byref(out))
    return out.value


def peak_profile(k, kpeak):
    """peak_profile (src/N.jl:366)."""
    return k ** 4 * math.exp(-2 * (k / kpeak) ** 2)


def randomfield(grid, poisson, profile="peak", param=5.0, totalenergy=0.5, seed=123, dtype=None):
    """randomfield (src/N.jl:394-476) with a counter-based generator.

    profile "peak": peak_profile(k; kpeak=param); "power": (k > 0) k^param."""
    u = vectorfield(grid, dtype=dtype or poisson.torch_dtype)
    pid = {"peak": 0, "power": 1}[profile]
    _ctx().call("ec_ns_randomfield", poisson.h, pid, param, totalenergy, seed,
                ptrs([x.data for x in u], P3))
    return u


def spectral_stuff(grid, npoint=None):
    """spectral_stuff (src/N.jl:547-574): the output shells 1..kcut, kcut =
    div(2 kmax, 3), kmax = div(n, 2) (or npoint log-spaced integer shells of
    that range), and their wavenumbers 2 pi / l * shell.  The shells
    themselves (getshells, src/N.jl:478-545) are formed on the device by
    ec_ns_spectrum."""
    kmax = grid.n // 2
    kcut = (2 * kmax) // 3
    if npoint is None:
        kuse = list(range(1, kcut + 1))
    else:
        kuse = sorted(set(int(round(x)) for x in np.geomspace(1.0, float(kcut), npoint)))
    return {"shells": kuse, "k": [2 * np.pi / grid.l * s for s in kuse]}


def spectrum(u, stuff, poisson):
    """spectrum(u, stuff, poisson) (src/N.jl:576-590): (k, s) with s[i] the
    kinetic energy of shell stuff["shells"][i], i.e. the sum of |u^/n^3|^2 / 2
    over the wavenumbers with shell^2 <= |k|^2 < (shell + 1)^2, conjugate
    half included, summed over the components (FP64).  Overwrites the
    solver's spectrum buffer, as the reference overwrites poisson.phat."""
    shells = stuff["shells"]
    if not shells:  # n < 3: spectral_stuff selects no shell
        return {"k": [], "s": []}
    nshell = max(shells) + 1
    out = (ctypes.c_double * nshell)()
    _ctx().call("ec_ns_spectrum", poisson.h, ptrs([x.data for x in u], P3), nshell, out)
    return {"k": list(stuff["k"]), "s": [out[i] for i in shells]}


def step_forwardeuler_(u, du, poisson, visc, dt):
    """step_forwardeuler! (src/N.jl:309-318)."""
    rhs_(du, u, visc, False)
    for a, b in zip(u, du):
        _axpy(dt, b, a)
    return project_(u, poisson)


def step_wray3_(u, du, u0, poisson, visc, dt):
    """step_wray3! (src/N.jl:320-346)."""
    a = (8 / 15, 5 / 12, 3 / 4)
    b = (1 / 4, 0 / 1)
    for x, x0 in zip(u, u0):
        x0.data.copy_(x.data)
    for i in range(3):
        rhs_(du, u, visc, False)
        for d, x, x0 in zip(du, u, u0):
            if i != 0:
                x.data.copy_(x0.data)
            _axpy(a[i] * dt, d, x)
        project_(u, poisson)
        if i != 2:
            for d, x0 in zip(du, u0):
                _axpy(b[i] * dt, d, x0)
    return u


def _axpy(a, x, y):
    _ctx().call("ec_axpy", _dt(y), y.data.numel(), a, ptr(x.data), ptr(y.data))


# --------------------------------------------------------------------------
# Filters (src/N.jl:592-698)
# --------------------------------------------------------------------------


def tophat(grid, compression):
    """tophat (src/N.jl:592-597): odd compression, weights 1/c."""
    assert compression % 2 == 1, "tophat needs odd compression (trapezoid_face covers even)"
    return np.full(compression, 1.0 / compression)


def trapezoid_face(compression):
    """Face-centred volume-average weights for any compression: odd c is the
    top-hat; even c spans c+1 faces with half weight at both ends."""
    c = compression
    if c % 2:
        return np.full(c, 1.0 / c)
    w = np.full(c + 1, 1.0 / c)
    w[0] = w[-1] = 0.5 / c
    return w


def gaussian(grid, delta, nsigma):
    """gaussian (src/N.jl:599-607); delta = 0 gives [1] (the reference's NaN
    case, handled as src/Burgers.jl:149-150 does)."""
    if delta < 1e-15:
        return np.ones(1)
    h = spacing(grid)
    sigma = delta / math.sqrt(12)
    r = int(np.rint(nsigma * sigma / h))
    x = h * np.arange(-r, r + 1)
    w = np.exp(-x ** 2 / (2 * sigma ** 2))
    return w / w.sum()


def composekernel(F, G):
    """composekernel (src/N.jl:609-625): full discrete convolution."""
    return np.convolve(np.asarray(F, dtype=np.float64), np.asarray(G, dtype=np.float64))


class KernelProduct:
    """kernelproduct(grid, kernels) (src/N.jl:627-637), kept separable."""

    def __init__(self, kernels):
        self.kernels = [np.ascontiguousarray(k, dtype=np.float64) for k in kernels]

    def dense(self):
        w1, w2, w3 = self.kernels
        return w1[:, None, None] * w2[None, :, None] * w3[None, None, :]


def kernelproduct(grid, kernels):
    return KernelProduct(kernels)


def _quantity(source, pieces, out):
    q = _abi.FilterQuantity()
    q.source = _abi.SRC[source]
    keep = []
    for d, (w, off) in enumerate(pieces):
        w = np.ascontiguousarray(w, dtype=np.float64)
        keep.append(w)
        q.len[d] = len(w)
        q.off[d] = int(off)
        q.w[d] = w.ctypes.data_as(_abi.ctypes.POINTER(_abi.c_double))
    q.out = out.data_ptr()
    return q, keep


def filter_coarse(u, p, gbar, visc, quantities):
    """Evaluate lazycoarsegrain(lazyfilter(src, kernel)) for several sources.

    quantities: list of (source, [(w, off)] * 3) with offsets relative to the
    coarse face index c*(J+1)-1.  Returns FP64 tensors (nbar, nbar, nbar)."""
    import torch

    g = u[0].grid
    outs = []
    qs = []
    keep = []
    for src, pieces in quantities:
        o = torch.empty((gbar.n,) * 3, dtype=torch.float64, device="cuda")
        q, k = _quantity(src, pieces, o)
        outs.append(o)
        qs.append(q)
        keep.append(k)
    arr = (_abi.FilterQuantity * len(qs))(*qs)
    if p is None:
        pp = None
    elif isinstance(p, Field):
        pp = ptr(p.data)
    else:
        pp = _abi.c_void_p(p)  # a device pointer, e.g. PoissonSolver.p_ptr
    _ctx().call("ec_filter_coarse", _dt(u[0]), g.n, gbar.n, g.l, visc, ptrs([x.data for x in u], P3),
                pp, len(qs), arr)
    return outs


def coarse_offset(compression, stag, R):
    """Offset of a (2R+1)-tap centred kernel at the coarse-grain point
    c*J - !stag*(c/2) (coarsegrain!, src/N.jl:670-684), relative to c*(J+1)-1."""
    return -R - (0 if stag else compression // 2)


def applyfilter_coarsegrain(ubar, u, kernel, stag):
    """coarsegrain!(ubar, lazyfilter(u, kernel), stag) fused (src/N.jl:655-684)."""
    g = u.grid
    c = g.n // ubar.grid.n
    pieces = [(k, coarse_offset(c, s, len(k) // 2)) for k, s in zip(kernel.kernels, stag)]
    dummy = (u, u, u)
    (out,) = filter_coarse(dummy, None, ubar.grid, 0.0, [("u1", pieces)])
    ubar.data.copy_(out.to(ubar.dtype))
    return ubar


def coarsegrain_(ubar, u, stag):
    """coarsegrain!(ubar, u, stag) (src/N.jl:670-684): point decimation."""
    one = KernelProduct([np.ones(1)] * 3)
    return applyfilter_coarsegrain(ubar, u, one, stag)


def applyfilter_(ubar, u, kernel):
    """applyfilter!(ubar, u, kernelproduct(w1, w2, w3)) (src/N.jl:627-653) on
    the full grid, as three periodic 1D passes with FP64 sums."""
    g = u.grid
    ks = [np.ascontiguousarray(k, dtype=np.float64) for k in kernel.kernels]
    lens = (_abi.c_int * 3)(*[len(k) for k in ks])
    dp = _abi.ctypes.POINTER(_abi.c_double)
    ws = (dp * 3)(*[k.ctypes.data_as(dp) for k in ks])
    _ctx().call("ec_filter_apply", _dt(u), g.n, lens, ws, ptr(u.data), ptr(ubar.data))
    return ubar


# Staggered positions of get_positions (src/N.jl:999-1006): faces e_i, the
# cell centre and the three edge positions corners3D = (F,T,T), (T,F,T), (T,T,F)
CENTER = (False, False, False)
CORNER23, CORNER31, CORNER12 = (False, True, True), (True, False, True), (True, True, False)


def face(i):
    return tuple(d == i for d in range(3))


def dnsaid_quantities(c, FD, FDH):
    """The coarse quantities of one DNS-aided LES step with the kernels of
    get_kernels, as (name, source, [(w, off)] * 3) for filter_coarse:
      ubar{a}    lazyfilter(u_a, GDH) at face a (apply_DHpi, src/N.jl:1228-1237);
      sigDH{ab}  lazyfilter(sigma_ab, GDH) at the stress positions (get_sigDH,
                 src/N.jl:1088-1115), symmetric order 11, 22, 33, 12, 23, 31;
      rDHi{ab}   lazyfilter(r_ab, GDHi[b]) (get_rDHi, src/N.jl:1038-1073) in
                 nonsym order 11, 21, 31, 12, 22, 32, 13, 23, 33, r = sigma + p I.
    GDH = FDH x FDH x FDH, GDHi[i] = FD along i and FDH elsewhere."""
    GDH = (FDH, FDH, FDH)
    GDHi = [tuple(FD if d == i else FDH for d in range(3)) for i in range(3)]

    def pieces(kernels, stag):
        return [(np.asarray(k, dtype=np.float64), coarse_offset(c, s, len(k) // 2)) for k, s in zip(kernels, stag)]

    q = []
    for a in range(3):
        q.append((f"ubar{a + 1}", f"u{a + 1}", pieces(GDH, face(a))))
    for name, src, pos in (("11", "s11", CENTER), ("22", "s22", CENTER), ("33", "s33", CENTER),
                           ("12", "s12", CORNER12), ("23", "s23", CORNER23), ("31", "s31", CORNER31)):
        q.append((f"sigDH{name}", src, pieces(GDH, pos)))
    # r_ab filtered with GDHi[b], b the derivative direction of -div^ns
    for name, src, b, pos in (("11", "r11", 0, CENTER), ("21", "s12", 0, CORNER12), ("31", "s31", 0, CORNER31),
                              ("12", "s12", 1, CORNER12), ("22", "r22", 1, CENTER), ("32", "s23", 1, CORNER23),
                              ("13", "s31", 2, CORNER31), ("23", "s23", 2, CORNER23), ("33", "r33", 2, CENTER)):
        q.append((f"rDHi{name}", src, pieces(GDHi[b], pos)))
    return q


def stress_positions(c):
    """lazytensorcoarsegrain (src/N.jl:1117-1135) as one-tap quantities: the six
    stresses of a fine field sampled at their LES positions."""
    one = np.ones(1)
    return [(f"s{name}", f"s{name}", [(one, coarse_offset(c, st, 0)) for st in pos])
            for name, pos in (("11", CENTER), ("22", CENTER), ("33", CENTER), ("12", CORNER12),
                              ("23", CORNER23), ("31", CORNER31))]


def get_kernels(g_dns, g_les, delta_ratio, nsigma):
    """get_kernels (src/N.jl:1008-1020) as separable factors."""
    comp = g_dns.n // g_les.n
    FH = tophat(g_dns, comp)
    FD = gaussian(g_dns, delta_ratio * spacing(g_les), nsigma)
    FDH = composekernel(FH, FD)
    GDH = KernelProduct([FDH] * 3)
    GDHi = tuple(KernelProduct([FD if i == j else FDH for j in range(3)]) for i in range(3))
    return GDH, GDHi


# --------------------------------------------------------------------------
# Exact closure (src/N.jl:999-1237)
# --------------------------------------------------------------------------

NONSYM = ["11", "21", "31", "12", "22", "32", "13", "23", "33"]


class ClosureSet:
    """One LES grid of the closure: kind 'va' (volume average) or 'fa' (face
    average).  Holds the FP64 outputs ubar (3) and S r (9, nonsym order)."""

    def __init__(self, g_dns, nbar, kind="va"):
        import torch

        self.g_les = Grid(g_dns.l, nbar)
        self.kind = kind
        self.ubar = [torch.empty((nbar,) * 3, dtype=torch.float64, device="cuda") for _ in range(3)]
        self.flux = [torch.empty((nbar,) * 3, dtype=torch.float64, device="cuda") for _ in range(9)]
        self.tau = [torch.empty((nbar,) * 3, dtype=torch.float64, device="cuda") for _ in range(9)]

    def struct(self):
        s = _abi.ClosureSet()
        s.nbar = self.g_les.n
        s.kind = 0 if self.kind == "va" else 1
        for a in range(3):
            s.ubar[a] = self.ubar[a].data_ptr()
        for m in range(9):
            s.flux[m] = self.flux[m].data_ptr()
        return s


def closure_(sets, u, p, visc):
    """Filtered velocity V u and face-filtered fluxes S r, r = sigma(u) + p I
    (get_r, get_rDHi, apply_DHpi: src/N.jl:1022-1073, 1228-1237), every set in
    one call, FP64 outputs."""
    g = u[0].grid
    arr = (_abi.ClosureSet * len(sets))(*[s.struct() for s in sets])
    pp = ptr(p.data) if isinstance(p, Field) else _abi.c_void_p(p)
    _ctx().call("ec_ns_closure", _dt(u[0]), g.n, g.l, visc, ptrs([x.data for x in u], P3), pp, len(sets), arr)
    return sets


def commutator_(cset, ubar, visc):
    """tau = S r - sigma_H(ubar) in FP64 (the closure of step_DHpi_classic_flux_div!,
    src/N.jl:1189-1206, with sigma_les cancelling)."""
    g = cset.g_les
    _ctx().call("ec_ns_commutator", g.n, g.l, visc, ptrs(ubar, P3), ptrs(cset.flux, P9), ptrs(cset.tau, P9))
    return cset.tau


def compute_errors(ubar, uref):
    """compute_errors (src/N.jl:1334-1349) for one LES field, FP64 reductions."""
    s = 0.0
    for a, b in zip(ubar, uref):
        s += diffsq(a, b) / sumsq(b)
    return math.sqrt(s / 3)


# --------------------------------------------------------------------------
# Sub-filter stresses and their dissipation (src/NavierStokes.jl:470-593,
# process.jl:107-193)
# --------------------------------------------------------------------------

SFS_EXPERIMENTS = ("volavg", "project_volavg", "surfavg")
SFS_TENSORS = ("classic", "swapfil", "swapfil_symm")


def sfs_filter(u, p, g_les, visc, kind):
    """ec_ns_sfs_filter: the DNS-side filters of sfs_tensors_volume (kind
    "volume": volumefilter! of u and of r = sigma + p I, surfacefilter!(r,
    false)) or sfs_tensors_surface ("surface": surfacefilter! of u,
    surfacefilter!(r, true), linefilter!(r) off the diagonal) in one launch.
    p: the DNS pressure (Field or device pointer).  Returns FP64 tensors
    (fu[3], fr[9], sr[9]), nonsym order."""
    import torch

    g = u[0].grid
    k = {"volume": 0, "surface": 1}[kind]
    shape = (g_les.n,) * 3

    def new(m):
        return [torch.empty(shape, dtype=torch.float64, device="cuda") for _ in range(m)]

    fu, fr, sr = new(3), new(9), new(9)
    pp = ptr(p.data) if isinstance(p, Field) else _abi.c_void_p(p)
    _ctx().call("ec_ns_sfs_filter", _dt(u[0]), g.n, g_les.n, g.l, visc, ptrs([x.data for x in u], P3), pp, k,
                ptrs(fu, P3), ptrs(fr, P9), ptrs(sr, P9))
    return fu, fr, sr


def tensordissipation(sigma9, u):
    """Turbulox.tensordissipation_staggered!(d, sigma, u) as used in
    process.jl:160-180: d = sigma : grad u at the cell centres (FP64), sigma
    in nonsym order, u FP64."""
    import torch

    g = u[0].grid
    data = [x.data if isinstance(x, Field) else x for x in sigma9]
    d = torch.empty((g.n,) * 3, dtype=torch.float64, device="cuda")
    _ctx().call("ec_ns_tensordissipation", g.n, g.l, ptrs(data, P9), ptrs([x.data for x in u], P3), ptr(d))
    return Field(g, d)


def sfs_tensors(u, g_les, visc, experiment, ps_dns=None, ps_les=None):
    """sfs_tensors_volume (experiment "volavg": doproject false,
    "project_volavg": doproject true) and sfs_tensors_surface ("surfavg")
    (src/NavierStokes.jl:470-593) of the DNS field u:
      the DNS pressure from the projected right-hand side; fu the LES
      velocity filter (projected for "project_volavg"); the LES pressure of
      fu the same way, r_H(fu) = sigma_H(fu) + p_les I;
      classic = fr - r_H(fu), swapfil = sr - r_H(fu) and its symmetric part.
    Returns {"fu": 3 FP64 Fields, "classic"/"swapfil"/"swapfil_symm": 9 FP64
    tensors each, nonsym order}.  ps_dns/ps_les: Poisson solvers of the two
    grids (made here when omitted; ps_les is FP64)."""
    import torch

    if experiment not in SFS_EXPERIMENTS:
        raise ValueError(f"experiment {experiment!r} (one of {SFS_EXPERIMENTS})")
    g = u[0].grid
    if g.n % g_les.n:
        raise ValueError(f"n_les = {g_les.n} does not divide n_dns = {g.n}")
    ps_dns = ps_dns or PoissonSolver(g, u[0].dtype)
    ps_les = ps_les or PoissonSolver(g_les, torch.float64)
    du = vectorfield(g, u[0].dtype)
    rhs_(du, u, visc)
    project_(du, ps_dns)
    del du
    fu, fr, sr = sfs_filter(u, ps_dns.p_ptr, g_les, visc, "surface" if experiment == "surfavg" else "volume")
    fuF = tuple(Field(g_les, x) for x in fu)
    if experiment == "project_volavg":
        project_(fuF, ps_les)
    dul = vectorfield(g_les, torch.float64)
    rhs_(dul, fuF, visc)
    project_(dul, ps_les)
    pl = ps_les.p_ptr
    out = {"fu": fuF}
    ctx = _ctx()
    for key, q in (("classic", fr), ("swapfil", sr)):
        tau = [torch.empty_like(x) for x in q]
        ctx.call("ec_ns_commutator", g_les.n, g_les.l, visc, ptrs(fu, P3), ptrs(q, P9), ptrs(tau, P9))
        for a in range(3):  # r_H = sigma_H + p_les I on the diagonal
            ctx.call("ec_axpy", _abi.EC_F64, tau[4 * a].numel(), -1.0, _abi.c_void_p(pl), ptr(tau[4 * a]))
        out[key] = tau
    sw = out["swapfil"]
    symm = []
    for m, ab in enumerate(NONSYM):
        t = sw[m].clone()
        if ab[0] != ab[1]:
            ctx.call("ec_axpby", _abi.EC_F64, t.numel(), 0.5, ptr(sw[NONSYM.index(ab[::-1])]), 0.5, ptr(t))
        symm.append(t)
    out["swapfil_symm"] = symm
    return out


def sfs_dissipation(u, g_les, visc, experiment, ps_dns=None, ps_les=None):
    """The dissipation fields of one LES grid and experiment of
    process.jl:107-193: tensordissipation of sigma_classic, sigma_swapfil and
    sigma_swapfil_symm with the filtered velocity fu (FP64 Fields).  The
    kernel density estimate process.jl forms from them is left to the
    caller."""
    t = sfs_tensors(u, g_les, visc, experiment, ps_dns, ps_les)
    return {k: tensordissipation(t[k], t["fu"]) for k in SFS_TENSORS}
