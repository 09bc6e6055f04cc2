"""The reference's DNS-aided LES comparison on the GPU.

dnsaid_project (src/N.jl:1239-1332): one DNS and five LES fields advanced
side by side with forward Euler and a CFL time step, the LES started from
apply_DHpi of the warmed-up DNS:

  u_nomo      no model (step_forwardeuler!, src/N.jl:309-318)
  u_c         classic: sigma_les + sigma_DH - coarse(sigma(G_DH u))
              (step_DHpi_classic!, src/N.jl:1137-1162; odd compression only)
  u_cf        classic + flux: sigma_les + sigma_DH - sigma_H(ubar)
              (step_DHpi_classic_flux!, src/N.jl:1164-1187)
  u_cfd       classic + flux + div, the exact closure:
              -div_H (sigma_les - sigma_H(ubar)) - div^ns_H r_DHi
              (step_DHpi_classic_flux_div!, src/N.jl:1189-1206)
  u_cfd_symm  the same with the symmetrised r_DHi
              (step_DHpi_classic_flux_div_symm!, src/N.jl:1208-1226)

with the filter kernels of get_kernels (src/N.jl:1008-1020): the top-hat of
the compression composed with a Gaussian of width Delta = delta_ratio * H
(nsigma sigmas of support).  compute_errors (src/N.jl:1334-1349) compares
each LES with apply_DHpi of the final DNS.  Every field operation runs in a
libexactclosure kernel (FP64): the fused RHS, the FFT projection, the
separable filters sampled at the LES points, the full-grid filter of the
classic variant, stress, tensor divergences and axpy updates.
"""
from __future__ import annotations

import torch

from . import _abi
from . import navierstokes as NS

VARIANTS = ("u_nomo", "u_c", "u_cf", "u_cfd", "u_cfd_symm")


def _axpy(a, x, y):
    """y += a x (LinearAlgebra.axpy!) on device tensors or Fields."""
    xd = x.data if isinstance(x, NS.Field) else x
    yd = y.data if isinstance(y, NS.Field) else y
    NS._ctx().call("ec_axpy", _abi.dtype_code(yd), yd.numel(), float(a), _abi.c_void_p(xd.data_ptr()),
                   _abi.c_void_p(yd.data_ptr()))


def _copy(dst, src):
    dd = dst.data if isinstance(dst, NS.Field) else dst
    sd = src.data if isinstance(src, NS.Field) else src
    NS._ctx().call("ec_d2d", _abi.c_void_p(dd.data_ptr()), _abi.c_void_p(sd.data_ptr()),
                   dd.numel() * dd.element_size())


class DNSAidProject:
    """State of dnsaid_project (src/N.jl:1239-1332) on one GPU, FP64.

    g_dns, g_les: grids (n_les divides n_dns with an odd compression); visc;
    delta_ratio and nsigma of the Gaussian (delta_ratio = 0: the plain
    top-hat)."""

    def __init__(self, g_dns, g_les, visc, delta_ratio=2.0, nsigma=2):
        self.g, self.gl = g_dns, g_les
        self.visc = float(visc)
        self.c = g_dns.n // g_les.n
        assert self.c * g_les.n == g_dns.n, "n_les must divide n_dns"
        c = self.c
        # get_kernels takes tophat(g_dns, c), which asserts an odd compression;
        # even compressions use the RK4 driver of dnsaid.py (trapezoid filters)
        assert c % 2 == 1, f"dnsaid_project needs an odd compression (got {c})"
        FH = NS.tophat(g_dns, c)
        self.FD = NS.gaussian(g_dns, delta_ratio * NS.spacing(g_les), nsigma)
        self.FDH = NS.composekernel(FH, self.FD)
        self.GDH = NS.KernelProduct([self.FDH] * 3)
        self.quantities = NS.dnsaid_quantities(c, self.FD, self.FDH)
        self.ps_dns = NS.PoissonSolver(g_dns, torch.float64)
        self.ps_les = NS.PoissonSolver(g_les, torch.float64)
        self.u = NS.vectorfield(g_dns, torch.float64)
        self.du = NS.vectorfield(g_dns, torch.float64)
        self.les = {k: NS.vectorfield(g_les, torch.float64) for k in VARIANTS}
        self.ubar = NS.vectorfield(g_les, torch.float64)
        self.du_les = NS.vectorfield(g_les, torch.float64)
        # the classic variant needs the full-grid filtered DNS velocity
        self.uf = NS.vectorfield(g_dns, torch.float64)
        self.t = 0.0
        self.nstep = 0

    # -- state ---------------------------------------------------------------
    def set_state(self, u):
        for a in range(3):
            _copy(self.u[a], u[a])

    def randomfield(self, profile="power", param=-5.0 / 3.0, totalenergy=1.0, seed=0):
        """The reference's initial field (src/N.jl:1245-1252): spectrum
        (k > 0) k^(-5/3), total energy 1 (counter-based normal variates)."""
        u = NS.randomfield(self.g, self.ps_dns, profile, param, totalenergy, seed=seed)
        self.set_state(u)

    def apply_DHpi(self, out):
        """apply_DHpi (src/N.jl:1228-1237): filter, coarse-grain at the faces, project."""
        qs = [(src, pcs) for name, src, pcs in self.quantities if name.startswith("ubar")]
        outs = NS.filter_coarse(self.u, None, self.gl, self.visc, qs)
        for a in range(3):
            _copy(out[a], outs[a])
        NS.project_(out, self.ps_les)
        return out

    def init_les(self):
        self.apply_DHpi(self.ubar)
        for k in VARIANTS:
            for a in range(3):
                _copy(self.les[k][a], self.ubar[a])

    def propose_timestep(self, cfl):
        return NS.propose_timestep(self.u, self.visc, cfl)

    def warmup(self, twarm, cfl):
        """The DNS warm-up of dnsaid_project: forward Euler to twarm."""
        t = 0.0
        while t < twarm:
            dt = min(self.propose_timestep(cfl), twarm - t)
            t += dt
            NS.step_forwardeuler_(self.u, self.du, self.ps_dns, self.visc, dt)

    # -- one step --------------------------------------------------------------
    def _les_fe(self, u_les, du, dt):
        """project!(du), then u += dt du (the end of every LES step)."""
        NS.project_(du, self.ps_les)
        for a in range(3):
            _axpy(dt, du[a], u_les[a])

    def step(self, dt):
        visc, gl = self.visc, self.gl
        # DNS right-hand side, projected; its pressure stays in ps_dns.p
        NS.rhs_(self.du, self.u, visc, False)
        NS.project_(self.du, self.ps_dns)
        # every coarse quantity of the step in one call (r = sigma + p I with
        # the pressure of that projection, get_r, src/N.jl:1022-1036)
        qs = [(src, pcs) for _, src, pcs in self.quantities]
        outs = NS.filter_coarse(self.u, self.ps_dns.p_ptr, gl, visc, qs)
        q = {name: o for (name, _, _), o in zip(self.quantities, outs)}
        for a in range(3):
            _copy(self.ubar[a], q[f"ubar{a + 1}"])
        NS.project_(self.ubar, self.ps_les)
        sigDH = [q["sigDH" + ab] for ab in ("11", "22", "33", "12", "23", "31")]
        rD = [q["rDHi" + ab] for ab in NS.NONSYM]
        sH = NS.stresstensor(self.ubar, visc)
        du = self.du_les
        # no model
        NS.rhs_(du, self.les["u_nomo"], visc, False)
        self._les_fe(self.les["u_nomo"], du, dt)
        # classic: sigma_les + sigma_DH - coarse(sigma(G_DH u))
        for a in range(3):
            NS.applyfilter_(self.uf[a], self.u[a], self.GDH)
        s2 = NS.filter_coarse(self.uf, None, gl, visc, [(src, pcs) for _, src, pcs in NS.stress_positions(self.c)])
        sl = NS.stresstensor(self.les["u_c"], visc)
        for m in range(6):
            _axpy(1.0, sigDH[m], sl[m])
            _axpy(-1.0, s2[m], sl[m])
        NS.tensordivergence_(du, sl, False)
        self._les_fe(self.les["u_c"], du, dt)
        # classic + flux: sigma_les + sigma_DH - sigma_H(ubar)
        sl = NS.stresstensor(self.les["u_cf"], visc)
        for m in range(6):
            _axpy(1.0, sigDH[m], sl[m])
            _axpy(-1.0, sH[m], sl[m])
        NS.tensordivergence_(du, sl, False)
        self._les_fe(self.les["u_cf"], du, dt)
        # classic + flux + div (exact): -div (sigma_les - sigma_H(ubar)) - div^ns r_DHi
        sl = NS.stresstensor(self.les["u_cfd"], visc)
        for m in range(6):
            _axpy(-1.0, sH[m], sl[m])
        NS.tensordivergence_(du, sl, False)
        NS.tensordivergence_nonsym_(du, rD, True)
        self._les_fe(self.les["u_cfd"], du, dt)
        # symmetrised: r_DHi replaced by (r_ab + r_ba) / 2 (symmetrize, src/N.jl:1075-1086)
        sl = NS.stresstensor(self.les["u_cfd_symm"], visc)
        for m in range(6):
            _axpy(-1.0, sH[m], sl[m])
        NS.tensordivergence_(du, sl, False)
        R = dict(zip(NS.NONSYM, rD))
        symm = [R["11"].clone(), R["22"].clone(), R["33"].clone(), R["21"].clone(), R["32"].clone(), R["31"].clone()]
        for m, other in ((3, "12"), (4, "23"), (5, "13")):
            NS._ctx().call("ec_axpby", _abi.EC_F64, symm[m].numel(), 0.5, _abi.c_void_p(R[other].data_ptr()), 0.5,
                           _abi.c_void_p(symm[m].data_ptr()))
        NS.tensordivergence_(du, [NS.Field(gl, x) for x in symm], True)
        self._les_fe(self.les["u_cfd_symm"], du, dt)
        # DNS
        for a in range(3):
            _axpy(dt, self.du[a], self.u[a])
        self.t += dt
        self.nstep += 1

    # -- drivers ---------------------------------------------------------------
    def run_fixed(self, dt, nsteps):
        for _ in range(nsteps):
            self.step(dt)

    def run(self, tstop, cfl):
        """The time loop of dnsaid_project: dt = propose_timestep (CFL), the
        last step cut to reach tstop; returns the number of steps."""
        t = 0.0
        k = 0
        while t < tstop:
            dt = min(self.propose_timestep(cfl), tstop - t)
            t += dt
            k += 1
            self.step(dt)
        return k

    def result(self):
        """(u_les, u_nomo, u_c, u_cf, u_cfd, u_cfd_symm) as dnsaid_project returns
        them, u_les = apply_DHpi of the current DNS."""
        out = {"u_les": self.apply_DHpi(NS.vectorfield(self.gl, torch.float64))}
        for k in VARIANTS:
            out[k] = self.les[k]
        return out

    def compute_errors(self):
        """compute_errors (src/N.jl:1334-1349) of every LES variant."""
        res = self.result()
        ref = res.pop("u_les")
        return {k: NS.compute_errors(v, ref) for k, v in res.items()}


def reference_setup(n_dns=500, n_les=100, visc=2.0e-4, delta_ratio=2.0, nsigma=2, twarm=0.1, tstop=0.1, cfl=0.4,
                    seed=0):
    """getsetup + dnsaid_project + compute_errors (navierstokes.jl:10-38,
    src/N.jl:712-723): returns the five errors, the step counts and the
    setup.  The initial field follows the reference's (k > 0) k^(-5/3)
    spectrum with energy 1; its normal variates come from this library's
    counter-based generator, not Julia's Xoshiro(0) stream."""
    g_dns, g_les = NS.Grid(1.0, n_dns), NS.Grid(1.0, n_les)
    run = DNSAidProject(g_dns, g_les, visc, delta_ratio, nsigma)
    run.randomfield("power", -5.0 / 3.0, 1.0, seed=seed)
    run.warmup(twarm, cfl)
    run.init_les()
    nsteps = run.run(tstop, cfl)
    errs = run.compute_errors()
    return {"errors": errs, "steps": nsteps,
            "setup": {"n_dns": n_dns, "n_les": n_les, "visc": visc, "delta_ratio": delta_ratio, "nsigma": nsigma,
                      "twarm": twarm, "tstop": tstop, "cfl": cfl, "seed": seed}}
