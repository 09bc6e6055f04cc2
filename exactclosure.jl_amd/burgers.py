"""Drop-in mirror of the ``Burgers`` module of ExactClosure.jl (src/Burgers.jl)
for the batched DNS-aided LES path.

``BurgersBatch`` replaces the serial loop of run_dns_aided_les
(src/Burgers.jl:267-328): many instances of uneven size and compression run
in one persistent launch (optionally replayed from a captured hipGraph).
"""
from __future__ import annotations

import ctypes
import math

import numpy as np

from . import _abi


def spacing(L, n):
    return L / n


def tophat_weights(comp):
    """tophat_weights (src/Burgers.jl:158-162) for odd comp; even comp uses the
    trapezoid over comp+1 faces (the reference would index past its array)."""
    if comp % 2:
        return comp // 2, np.full(comp, 1.0 / comp)
    w = np.full(comp + 1, 1.0 / comp)
    w[0] = w[-1] = 0.5 / comp
    return comp // 2, w


def gaussian_weights(h, delta, nsigma=3):
    """gaussian_weights (src/Burgers.jl:146-156)."""
    s2 = (delta / h) ** 2 / 12
    R = int(np.rint(nsigma * math.sqrt(s2)))
    if delta < 1e-15:
        return R, np.ones(1)
    w = np.exp(-np.arange(-R, R + 1) ** 2 / (2 * s2))
    return R, w / w.sum()


def build_doublekernel(F, G):
    """build_doublekernel (src/Burgers.jl:217-233)."""
    I, f = F
    J, g = G
    return I + J, np.convolve(f, g)


def les_kernels(nh, nH, L, delta_ratio, lesfiltertype="tophat"):
    """The LES kernel and the double kernel of compute_fractions
    (src/Burgers.jl:1030-1048): Delta = H * delta_ratio; the finite-volume
    top-hat over comp = nh / nH fine cells; the LES top-hat (width rounded to
    the nearest odd count of fine cells) or Gaussian; their convolution."""
    h, H = L / nh, L / nH
    delta = H * delta_ratio
    fvm = tophat_weights(nh // nH)
    if lesfiltertype == "tophat":
        R = int(np.rint(delta / h)) // 2
        les = tophat_weights(2 * R + 1)
    elif lesfiltertype == "gaussian":
        les = gaussian_weights(h, delta)
    else:
        raise ValueError(f"lesfiltertype {lesfiltertype!r}")
    return les, build_doublekernel(fvm, les)


def compute_fractions(u, nH, L=2 * math.pi, visc=5e-4, delta_ratio=1.0, lesfiltertype="tophat"):
    """compute_fractions for one (nH, delta_ratio) (src/Burgers.jl:1022-1099):
    u holds DNS snapshots as a (nsamples, nh) FP64 CUDA tensor (one sample per
    row, the columns of the reference's data matrix); returns the mean
    fractions {classic, flux, div} of the closure parts, computed on the GPU
    in FP64."""
    import torch

    assert u.dtype == torch.float64 and u.is_cuda and u.dim() == 2 and u.is_contiguous()
    ns, nh = u.shape
    (Rl, wl), (K, wd) = les_kernels(nh, nH, L, delta_ratio, lesfiltertype)
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wd = np.ascontiguousarray(wd, dtype=np.float64)
    out = (ctypes.c_double * 3)()
    dp = ctypes.POINTER(ctypes.c_double)
    _abi.context().call("ec_burgers_fractions", nh, nH, L, visc, ns, _abi.c_void_p(u.data_ptr()), Rl,
                        wl.ctypes.data_as(dp), K, wd.ctypes.data_as(dp), out)
    return {"classic": out[0], "flux": out[1], "div": out[2]}


def compute_dissipation(u, nH, L=2 * math.pi, visc=5e-4, delta_ratio=1.0, lesfiltertype="tophat"):
    """compute_dissipation for one (nH, delta_ratio) (src/Burgers.jl:447-516):
    the sub-filter dissipation coefficients sigma * delta_stag(vu) at every
    coarse point of every sample for the stresses class_m (vsu - s_h(D*u)),
    class_p (vsu - s_H(vu)) and swapfil (fsu - s_H(vu)), FP64 on the GPU.
    u: (nsamples, nh) FP64 CUDA tensor.  Returns the reference's named tuple:
    nomodel (one zero) and one (nsamples * nH,) CUDA tensor per model, coarse
    point fastest, then sample (vec(d[:, m, :]))."""
    import torch

    assert u.dtype == torch.float64 and u.is_cuda and u.dim() == 2 and u.is_contiguous()
    ns, nh = u.shape
    (Rl, wl), (K, wd) = les_kernels(nh, nH, L, delta_ratio, lesfiltertype)
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wd = np.ascontiguousarray(wd, dtype=np.float64)
    out = torch.empty((3, ns * nH), dtype=torch.float64, device=u.device)
    dp = ctypes.POINTER(ctypes.c_double)
    _abi.context().call("ec_burgers_dissipation", nh, nH, L, visc, ns, _abi.c_void_p(u.data_ptr()), Rl,
                        wl.ctypes.data_as(dp), K, wd.ctypes.data_as(dp), _abi.c_void_p(out.data_ptr()))
    return {"nomodel": torch.zeros(1, dtype=torch.float64, device=u.device), "class_m": out[0],
            "class_p": out[1], "swapfil": out[2]}


def smagorinsky_coefficients(U, nH, L=2 * math.pi, visc=5e-4, delta_ratio=2.0, lesfiltertype="gaussian"):
    """theta_classic and theta_discret (smagorinsky.jl:24-36 over the fields
    of smagorinsky_fields, src/Burgers.jl:547-680): least-squares Smagorinsky
    coefficients theta = sqrt(-<S,T>/<S,S>) against the classic sub-filter
    stress on the DNS grid and the discretisation-consistent one on the LES
    stencil.  U: DNS snapshots as a (nsamples, nh) FP64 CUDA tensor.  A
    negative theta^2 raises, as Julia's sqrt does."""
    import torch

    assert U.dtype == torch.float64 and U.is_cuda and U.dim() == 2 and U.is_contiguous()
    ns, nh = U.shape
    (Rl, wl), (K, wd) = les_kernels(nh, nH, L, delta_ratio, lesfiltertype)
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wd = np.ascontiguousarray(wd, dtype=np.float64)
    out = (ctypes.c_double * 4)()
    dp = ctypes.POINTER(ctypes.c_double)
    _abi.context().call("ec_burgers_smag_fit", nh, nH, L, visc, L / nH * delta_ratio, ns,
                        _abi.c_void_p(U.data_ptr()), Rl, wl.ctypes.data_as(dp), K, wd.ctypes.data_as(dp), out)
    sums = list(out)
    return {"classic": math.sqrt(-sums[0] / sums[1]), "discret": math.sqrt(-sums[2] / sums[3]), "sums": sums}


def solve_smagorinsky(ustart, nH, theta, L=2 * math.pi, visc=5e-4, delta_ratio=2.0, lesfiltertype="gaussian",
                      tstop=0.1, cfl=0.3):
    """solve_smagorinsky for one (Delta_ratio, nH) (src/Burgers.jl:682-772):
    every sample of ustart ((nsamples, nh) FP64 CUDA tensor) filtered to the
    LES grid and advanced to tstop with the Smagorinsky closure
    C = theta^2 (Delta^2 + H^2), all samples in one launch; returns the
    (nsamples, nH) FP64 LES states."""
    import torch

    assert ustart.dtype == torch.float64 and ustart.is_cuda and ustart.dim() == 2 and ustart.is_contiguous()
    ns, nh = ustart.shape
    _, (K, wd) = les_kernels(nh, nH, L, delta_ratio, lesfiltertype)
    wd = np.ascontiguousarray(wd, dtype=np.float64)
    H = L / nH
    C = theta ** 2 * ((H * delta_ratio) ** 2 + H ** 2)
    out = torch.empty((ns, nH), dtype=torch.float64, device=ustart.device)
    _abi.context().call("ec_burgers_smag_solve", nh, nH, L, visc, C, tstop, cfl, ns,
                        _abi.c_void_p(ustart.data_ptr()), K, wd.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                        _abi.c_void_p(out.data_ptr()))
    return out


FIELDS = ("dns_ref", "dns_ref_fil", "dns_fil", "nomodel", "class_m", "class_p", "swapfil")
LES_KEYS = ("nomodel", "class_m", "class_p", "swapfil")


def getsetup(nsample=1000):
    """getsetup (src/Burgers.jl:235-270): the reference's Burgers study."""
    return {"L": 2 * math.pi, "nh": 100 * 3 ** 3 * 5, "nH": [100 * 3 ** k for k in (1, 2, 3)], "Delta_ratios": [2.0],
            "visc": 5.0e-4, "kpeak": 10, "initialenergy": 2.0, "tstop": 0.1, "nsample": nsample,
            "lesfiltertype": "gaussian"}


def randomfield(n, kpeak, totalenergy, nsamples=1, seed=1, L=2 * math.pi):
    """randomfield with bump_profile (src/Burgers.jl:192-202) for nsamples
    samples (sample q seeded by seed + q, as the reference seeds Xoshiro(i)
    per sample; the phases come from this library's counter-based hash):
    a (nsamples, n) FP64 CUDA tensor."""
    import torch

    u = torch.empty((nsamples, n), dtype=torch.float64, device="cuda")
    _abi.context().call("ec_burgers_randomfield", n, L, float(kpeak), float(totalenergy), nsamples, seed,
                        _abi.c_void_p(u.data_ptr()))
    return u


def dns_aided_les(ustart, nH, visc, delta, tstop, cfl_factor=0.3, lesfiltertype="gaussian", L=2 * math.pi):
    """dns_aided_les(ustart, gh, gH, visc; Delta, tstop, cfl_factor,
    lesfiltertype) (src/Burgers.jl:350-445) for every row of ustart ((nsamples,
    nh) FP64 CUDA tensor) in one launch: forward Euler with the CFL step, the
    DNS and the nomodel, class_m, class_p and swapfil LES.  Returns a dict of
    (nsamples, n) tensors keyed like the reference's named tuple, plus
    "steps" (list of step counts)."""
    import torch

    assert ustart.dtype == torch.float64 and ustart.is_cuda and ustart.dim() == 2 and ustart.is_contiguous()
    ns, nh = ustart.shape
    delta_ratio = delta / (L / nH)
    (Rl, wl), (K, wd) = les_kernels(nh, nH, L, delta_ratio, lesfiltertype)
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wd = np.ascontiguousarray(wd, dtype=np.float64)
    out = torch.empty((ns, 2 * nh + 5 * nH), dtype=torch.float64, device="cuda")
    steps = (ctypes.c_int * ns)()
    dp = ctypes.POINTER(ctypes.c_double)
    _abi.context().call("ec_burgers_dns_aided_les", nh, nH, L, visc, tstop, cfl_factor, ns,
                        _abi.c_void_p(ustart.data_ptr()), Rl, wl.ctypes.data_as(dp), K, wd.ctypes.data_as(dp),
                        _abi.c_void_p(out.data_ptr()), steps)
    res = {"dns_ref": out[:, :nh], "dns_ref_fil": out[:, nh:2 * nh]}
    for m, key in enumerate(FIELDS[2:]):
        res[key] = out[:, 2 * nh + m * nH:2 * nh + (m + 1) * nH]
    res["steps"] = list(steps)
    return res


def run_dns_aided_les(setup, seed=1):
    """run_dns_aided_les (src/Burgers.jl:272-325): every (nH, Delta_ratio) of
    the setup over nsample samples (sample i from randomfield seeded i), all
    samples of one (nH, Delta_ratio) in one launch.  Returns a list of
    {"nH", "Delta_ratio", "fields"} like the reference's saved results."""
    out = []
    ustart = randomfield(setup["nh"], setup["kpeak"], setup["initialenergy"], setup["nsample"], seed, setup["L"])
    for nH in setup["nH"]:
        for dr in setup["Delta_ratios"]:
            H = setup["L"] / nH
            f = dns_aided_les(ustart, nH, setup["visc"], H * dr, setup["tstop"], 0.3, setup["lesfiltertype"],
                              setup["L"])
            out.append({"nH": nH, "Delta_ratio": dr, "fields": f})
    return out


def compute_errors(series):
    """compute_errors (src/Burgers.jl:909-917): for every LES model the
    relative error norm(u - dns_fil) / norm(dns_fil) over all samples, one
    entry per (nH, Delta_ratio) of the series (FP64)."""
    return {key: [float((r["fields"][key] - r["fields"]["dns_fil"]).norm() / r["fields"]["dns_fil"].norm())
                  for r in series] for key in LES_KEYS}


class BurgersBatch:
    """count instances: n[i] points, compression c[i], RK4 with the exact
    closure and a no-model LES, FP32 or FP64 state, FP64 commutator."""

    def __init__(self, n, c, visc=5e-4, L=2 * math.pi, cfl=0.3, kmax=10, seed=1, dtype="f32"):
        self.ctx = _abi.context(0)
        self.n = np.asarray(n, dtype=np.int32)
        self.c = np.asarray(c, dtype=np.int32)
        self.count = len(self.n)
        self.L = L
        self.visc = visc
        code = _abi.EC_F64 if dtype in ("f64", "float64") else _abi.EC_F32
        self.dtype = "f64" if code == _abi.EC_F64 else "f32"
        h = _abi.c_void_p()
        nn = (ctypes.c_int * self.count)(*self.n.tolist())
        cc = (ctypes.c_int * self.count)(*self.c.tolist())
        self.ctx.call("ec_burgers_batch_create", code, self.count, nn, cc, L, visc, cfl, kmax, seed,
                      ctypes.byref(h))
        self.h = h

    def set_threads(self, threads):
        self.ctx.call("ec_burgers_batch_set_threads", self.h, threads)

    def run(self, nsteps, use_graph=True):
        self.ctx.call("ec_burgers_batch_run", self.h, nsteps, int(bool(use_graph)))

    def get(self, i, which="dns"):
        w = {"dns": 0, "swapfil": 1, "nomodel": 2}[which]
        m = int(self.n[i]) if w == 0 else int(self.n[i] // self.c[i])
        out = np.empty(m, dtype=np.float64)
        self.ctx.call("ec_burgers_batch_get", self.h, i, w, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), m)
        return out

    def dt(self, i):
        d = ctypes.c_double()
        self.ctx.call("ec_burgers_batch_dt", self.h, i, ctypes.byref(d))
        return d.value

    def close(self):
        if self.h:
            self.ctx.lib.ec_burgers_batch_destroy(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def baseline_batch(count=4096, seed=0):
    """The batched-Burgers workload: sizes {2048, 4096, 8192, 16384} split
    40/30/20/10 %, compressions drawn from {4, 8, 16, 32, 64}."""
    rng = np.random.default_rng(seed)
    counts = [int(round(count * f)) for f in (0.4, 0.3, 0.2)]
    counts.append(count - sum(counts))
    n = np.concatenate([np.full(k, s) for k, s in zip(counts, (2048, 4096, 8192, 16384))])
    c = rng.choice([4, 8, 16, 32, 64], size=count)
    perm = rng.permutation(count)
    return n[perm].astype(np.int32), c[perm].astype(np.int32)


def benchmark(count=4096, nsteps=1000, dtype="f32", tune=True, seed=0):
    """Time the batched workload: `nsteps` RK4 steps of every instance (DNS,
    exact-closure LES and no-model LES) in one graph-replayed persistent
    launch. Returns a dict with DNS point-steps per second."""
    import torch

    from . import autotune

    n, c = baseline_batch(count, seed)
    tuned = None
    if tune:
        probe = BurgersBatch(n, c, dtype=dtype, seed=seed + 1)
        tuned = autotune.tune_burgers_threads(autotune.Autotuner(), probe)
        probe.close()
    b = BurgersBatch(n, c, dtype=dtype, seed=seed)
    if tuned:
        b.set_threads(tuned)
    b.run(nsteps, use_graph=True)     # capture the graph of this length, warm up
    torch.cuda.synchronize()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    b.run(nsteps, use_graph=True)
    t1.record()
    t1.synchronize()
    ms = t0.elapsed_time(t1)
    pts = int(n.sum())
    les_pts = int((n // c).sum())
    out = {"instances": count, "points": pts, "les_points": les_pts, "steps": nsteps, "ms": ms,
           "point_steps_per_s": pts * nsteps / (ms * 1e-3), "threads": tuned, "dtype": dtype}
    b.close()
    return out
