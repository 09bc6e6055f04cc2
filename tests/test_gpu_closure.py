"""GPU parity of the filters, the exact-closure fluxes and the RK4 DNS-aided
LES driver against the FP64 CPU checker, plus the exactness invariant
(navierstokes.jl:33: the classic+flux+div LES equals the filtered DNS to
rounding) and the FP32-pipeline closure error against FP64 (gate 1e-5)."""
import numpy as np
import pytest
import torch

from test_gpu_ns import dev, host, rel

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


def test_filter_coarse_gaussian(gpu, O):
    """lazycoarsegrain(lazyfilter(u, GDH)) with the reference's Gaussian-composed
    kernel at odd compression (get_kernels, src/N.jl:1008-1019)."""
    NS = gpu.navierstokes
    n, c = 15, 3
    g, gb = NS.Grid(1.0, n), NS.Grid(1.0, n // c)
    FD = NS.gaussian(g, 2 * NS.spacing(gb), 2)
    FDH = NS.composekernel(NS.tophat(g, c), FD)
    u = O.randomfield(n, seed=3)
    ref = O.coarsegrain(O.applyfilter(u[0], FDH, FDH, FDH), n // c, (True, False, False))
    ub = NS.Field(gb)
    NS.applyfilter_coarsegrain(ub, NS.Field(g, dev(u[0])), NS.KernelProduct([FDH] * 3), (True, False, False))
    assert rel(ub.numpy(), ref) < 1e-12
    pt = NS.Field(gb)
    NS.coarsegrain_(pt, NS.Field(g, dev(u[1])), (False, True, False))
    assert rel(pt.numpy(), O.coarsegrain(u[1], n // c, (False, True, False))) < 1e-15


@pytest.mark.parametrize("n,sets", [(16, [(4, "va")]), (24, [(3, "va"), (8, "va")]), (16, [(4, "fa"), (2, "va")])])
def test_dnsaid_rk4_step_fp64(gpu, O, n, sets):
    NS, D = gpu.navierstokes, gpu.dnsaid
    visc, dt = 1.0 / 500, 2e-3
    g = NS.Grid(1.0, n)
    u = O.randomfield(n, seed=5)
    drv = D.DNSAidedLES(g, visc, [(n // c, k) for c, k in sets], dtype=torch.float64)
    drv.set_state([dev(x) for x in u])
    drv.run(dt, 2)
    u_cpu, les_cpu, tau_cpu = O.dnsaid_rk4([x.copy(order="F") for x in u], sets, visc, dt, 2)
    for a in range(3):
        assert rel(host(drv.u0[a]), u_cpu[a]) < 1e-11
    for si, (L, lc) in enumerate(zip(drv.les, les_cpu)):
        for a in range(3):
            assert rel(host(L.u0[a]), lc[a]) < 1e-10
        for m in range(9):
            # closure commutator of the last stage evaluated
            assert rel(host(L.cset.tau[m]), tau_cpu[si][m]) < 1e-9
    errs = drv.les_errors()
    for (c, k), e in zip(sets, errs):
        if k == "va":
            assert e < 1e-12, (c, e)  # exactness invariant


def test_exactness_many_steps_fp64(gpu, O):
    NS, D = gpu.navierstokes, gpu.dnsaid
    n, visc = 32, 1.0 / 2000
    g = NS.Grid(1.0, n)
    ps = NS.PoissonSolver(g, torch.float64)
    u = NS.randomfield(g, ps, seed=123)
    dt = NS.propose_timestep(u, visc, 0.4)
    drv = D.DNSAidedLES(g, visc, [(8, "va"), (4, "va"), (16, "va")], dtype=torch.float64)
    drv.set_state(u)
    drv.run(dt, 10)
    assert max(drv.les_errors()) < 1e-12
    assert drv.divergence_norm() < 1e-13


def test_fp32_pipeline_closure_error(gpu, O):
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
        drv = D.DNSAidedLES(g, visc, [(16, "va"), (8, "va")], dtype=dt)
        drv.set_state(u)
        c = drv.ctx
        code = drv.dt_code
        from ctypes import c_void_p
        P3 = gpu._abi.P3
        c.call("ec_ns_rhs", code, n, 1.0, visc, P3(*[x.data_ptr() for x in drv.u0]),
               P3(*[x.data_ptr() for x in drv.du]), 0)
        c.call("ec_ns_pressure_solve", drv.ps.h, P3(*[x.data_ptr() for x in drv.du]))
        drv._closure(drv.u0)
        taus = []
        for L in drv.les:
            ub = L.cset.ubar
            c.call("ec_ns_project", L.ps64.h, P3(*[x.data_ptr() for x in ub]))
            taus.append([t.clone() for t in NS.commutator_(L.cset, ub, visc)])
        res[dt] = taus
    for t32, t64 in zip(res[torch.float32], res[torch.float64]):
        num = sum(float(((a - b) ** 2).sum()) for a, b in zip(t32, t64))
        den = sum(float((b ** 2).sum()) for b in t64)
        assert (num / den) ** 0.5 < 1e-5
