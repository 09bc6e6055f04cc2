"""DNS-aided LES with classical RK4 and the exact closure at every stage.

Host orchestration of the hot path (the role of dnsaid_project,
src/N.jl:1239-1332, and dns_aid_volavg / dns_aid_surface,
src/NavierStokes.jl:141-468) over libexactclosure:

per stage s of an RK4 step (tableau of rk4_timestep!, src/Burgers.jl:120-142,
projection at every stage as in step_wray3!, src/N.jl:320-346):
  1. du = -div sigma(u_s)                         ec_ns_rhs (fused stencil)
  2. p  = L^-1 div du                              ec_ns_pressure_solve (FFT)
  3. for every LES grid: V u_s and S r, r = sigma(u_s) + p I (FP64)
                                                    ec_ns_closure
     ubar = P_H V u_s (FP64 LES projection), tau = S r - sigma_H(ubar) (FP64)
     du_les = -div_H sigma(u_les) - div^ns_H tau, p_les = L_H^-1 div du_les
  4. k = du - grad p; RK4 update of DNS and LES     ec_ns_rk4_update

With the volume-average filter the LES reproduces P_H V u_dns to rounding
(the exactness invariant of navierstokes.jl:33).  Fields run in `dtype`
(FP32 or FP64); the closure commutator runs in FP64.

closure="forcing" (the default) evaluates the same closure in divergence
form: V(k) = -div_H^ns (S r) by the commutation identity of the filter pair,
so step 3 becomes one filter launch of k = du - grad p for all LES grids,
v = V u is carried through the RK4 stages in FP64, and the closure forcing
m = V(k) - F_H(P_H v) (the commutator, FP64) drives
du_les = F_H(u_les) + m.  closure="flux" keeps the nine-component flux of
step 3 above.
"""
from __future__ import annotations

import torch

from . import _abi
from . import navierstokes as NS
from ._abi import P3, ptr, ptrs


def _vec(n, dtype):
    return [torch.empty((n, n, n), dtype=dtype, device="cuda") for _ in range(3)]


class LESGrid:
    def __init__(self, g_dns, nbar, kind, dtype, closure="forcing", rk_states=False):
        self.g = NS.Grid(g_dns.l, nbar)
        self.kind = kind
        self.dtype = dtype
        self.cset = NS.ClosureSet(g_dns, nbar, kind) if closure == "flux" else None
        # RK4 of v = V u (FP64) in stored stage states: vs and vacc hold v1
        # and v2, v3 the third.  The FP32 LES state keeps the accumulator form:
        # its stored states would carry FP32 rounding into the new state (the
        # LES then sits ~1.5x further from the filtered DNS) for a small saving
        self.rk_states = bool(rk_states) and closure == "forcing"
        f64 = torch.float64
        if closure == "forcing":
            self.v0 = _vec(nbar, f64)    # V u at the step start (unprojected)
            self.vs = _vec(nbar, f64)    # V u at the current stage
            self.vacc = _vec(nbar, f64)
            self.v3 = _vec(nbar, f64) if self.rk_states else None
            self.g_k = _vec(nbar, f64)   # V k of the current stage
            self.ubar = _vec(nbar, f64)  # P_H V u
            self.m = _vec(nbar, f64)     # closure forcing V k - F_H(ubar)
            self.zero64 = None if self.rk_states else torch.zeros((nbar,) * 3, dtype=f64, device="cuda")
        self.ps = NS.PoissonSolver(self.g, dtype)
        self.ps64 = self.ps if dtype == torch.float64 else NS.PoissonSolver(self.g, torch.float64)
        self.u0 = _vec(nbar, dtype)
        self.us = _vec(nbar, dtype)
        self.du = _vec(nbar, dtype)
        self.acc = _vec(nbar, dtype)

    def u_stage(self, st):
        """The LES state that stage st evaluates."""
        return self.u0 if st == 0 else self.us

    def v_stage(self, st):
        """v = V u at stage st."""
        if st == 0:
            return self.v0
        return [self.vs, self.vacc, self.v3][st - 1] if self.rk_states else self.vs


class DNSAidedLES:
    """State and buffers of one DNS with exact-closure LES grids.

    sets: list of (nbar, kind) with kind 'va' (volume average) or 'fa'
    (face average).  Memory at n = 1024 FP32: 12 DNS fields + p + half
    spectrum ~ 60 GB."""

    def __init__(self, g_dns, visc, sets, dtype=torch.float32, closure="forcing", clean_divergence=True,
                 rk_states=True, les_stream=True):
        assert closure in ("forcing", "flux")
        self.closure = closure
        # the DNS RK4 in stored stage states (ec_ns_rk4_update_states: stages
        # 0..2 write u1, u2, u3, the last stage combines them; 49 fields of
        # state traffic per step against 58 with an accumulator), forcing form
        # only; False keeps the accumulator form (ec_ns_rk4_update)
        self.rk_states = bool(rk_states) and closure == "forcing"
        # fold 6/dt div_h u_3 into the last stage's Poisson right-hand side, so
        # that the step removes the divergence the state carried in (FP32
        # rounding of the projected stage tendencies would otherwise add up
        # step after step); in FP64 the term is at rounding level
        self.clean_divergence = clean_divergence
        self.g = g_dns
        self.visc = float(visc)
        self.dtype = dtype
        self.ctx = _abi.context(0)
        self.dt_code = _abi.EC_F64 if dtype == torch.float64 else _abi.EC_F32
        n = g_dns.n
        self.ps = NS.PoissonSolver(g_dns, dtype)
        self.u0 = _vec(n, dtype)
        self.us = _vec(n, dtype)   # stage state (u1 of the stored-state form)
        self.du = _vec(n, dtype)
        self.acc = _vec(n, dtype)  # accumulator (u2 of the stored-state form)
        self.u3 = _vec(n, dtype) if self.rk_states else None
        self.les = [LESGrid(g_dns, nbar, kind, dtype, closure, self.rk_states) for nbar, kind in sets]
        # objects with .stage(st, dt), advanced after every stage of the forcing
        # form once the stage's exact closure is known (a posteriori model LES)
        self.stage_hooks = []
        self._graphs = {}  # dt -> captured step (ec_capture_end)
        # les_stream (forcing form): the LES grids' stage work runs on a second
        # stream, beside the DNS right-hand side and Poisson solve of the next
        # stage (it only needs the filter outputs of its own stage, and the
        # next filter waits for it); in a captured step it is a parallel
        # branch of the graph.  Same kernels on the same values: same bits;
        # 1% off the 1024^3 step (the LES launches fill what the DNS passes
        # leave of the device).  The last stage's LES work ends the step.
        self.les_ctx = None
        if les_stream and closure == "forcing" and self.les:
            self._les_torch_stream = torch.cuda.Stream()
            self.les_ctx = _abi.Context(0, self._les_torch_stream.cuda_stream)

    # -- state -------------------------------------------------------------
    def set_state(self, u):
        for a in range(3):
            self.u0[a].copy_(u[a].data if isinstance(u[a], NS.Field) else u[a])
        self.init_les()

    def _vel_sets(self, outs):
        arr = (_abi.VelSet * len(self.les))()
        for i, (L, o) in enumerate(zip(self.les, outs)):
            arr[i].nbar = L.g.n
            arr[i].kind = 0 if L.kind == "va" else 1
            for a in range(3):
                arr[i].out[a] = o[a].data_ptr()
        return arr

    def _filter_vel(self, v, p_ptr, outs):
        """outs[s][a] = V_s(v_a - delta_a p) for every LES grid, one launch."""
        if not self.les:
            return
        self.ctx.call("ec_ns_filter_vel", self.dt_code, self.g.n, self.g.l, ptrs(v, P3),
                      _abi.c_void_p(p_ptr) if p_ptr else None, len(self.les), self._vel_sets(outs))

    def init_les(self):
        """LES fields start from P_H V u (apply_DHpi, src/N.jl:1228-1237)."""
        if self.closure == "forcing":
            self._filter_vel(self.u0, None, [L.v0 for L in self.les])
            for L in self.les:
                for a in range(3):
                    L.ubar[a].copy_(L.v0[a])
                self.ctx.call("ec_ns_project", L.ps64.h, ptrs(L.ubar, P3))
                for a in range(3):
                    L.u0[a].copy_(L.ubar[a])
            return
        self._closure(self.u0)
        for L in self.les:
            ub = L.cset.ubar
            self.ctx.call("ec_ns_project", L.ps64.h, ptrs(ub, P3))
            for a in range(3):
                L.u0[a].copy_(ub[a])

    def filtered_dns(self):
        """P_H V u of the current DNS state for every LES grid (FP64 tensors)."""
        if self.closure == "forcing":
            outs = [[torch.empty_like(L.v0[0]) for _ in range(3)] for L in self.les]
            self._filter_vel(self.u0, None, outs)
            for L, ub in zip(self.les, outs):
                self.ctx.call("ec_ns_project", L.ps64.h, ptrs(ub, P3))
            return outs
        self._closure(self.u0)
        out = []
        for L in self.les:
            ub = [x.clone() for x in L.cset.ubar]
            self.ctx.call("ec_ns_project", L.ps64.h, ptrs(ub, P3))
            out.append(ub)
        return out

    def _closure(self, u, with_p=False):
        sets = [L.cset for L in self.les]
        if not sets:
            return
        p = self.ps.p_ptr
        arr = (_abi.ClosureSet * len(sets))(*[s.struct() for s in sets])
        self.ctx.call("ec_ns_closure", self.dt_code, self.g.n, self.g.l, self.visc, ptrs(u, P3),
                      _abi.c_void_p(p), len(sets), arr)

    # -- one RK4 step --------------------------------------------------------
    def step(self, dt):
        g = self._graphs.get(dt) if self._graphs else None
        if g is not None:
            self.ctx.call("ec_graph_launch", g)
            return
        self._step_eager(dt)

    def capture(self, dt):
        """Record one RK4 step of time step dt into a HIP graph (ec_capture_*,
        all of the step's launches, the LES grids included); step(dt) then
        replays it with one launch.  The graph holds this driver's buffers and
        dt; a stage hook (host work between stages) cannot be captured."""
        assert not self.stage_hooks, "stage hooks run host code between the stages"
        if dt in self._graphs:
            return
        self.ctx.call("ec_capture_begin")
        g = _abi.c_void_p()
        try:
            self._step_eager(dt)
        finally:
            rc = self.ctx.lib.ec_capture_end(self.ctx.h, _abi.ctypes.byref(g))
        self.ctx.check(rc, "ec_capture_end")
        self._graphs[dt] = g

    def release_graphs(self):
        for g in self._graphs.values():
            self.ctx.lib.ec_graph_destroy(self.ctx.h, g)
        self._graphs = {}

    def __del__(self):
        try:
            self.release_graphs()
        except Exception:
            pass

    def _step_eager(self, dt):
        if self.closure == "forcing":
            return self._step_forcing(dt)
        c = self.ctx
        n, l = self.g.n, self.g.l
        for st in range(4):
            u = self.u0 if st == 0 else self.us
            c.call("ec_ns_rhs", self.dt_code, n, l, self.visc, ptrs(u, P3), ptrs(self.du, P3), 0)
            c.call("ec_ns_pressure_solve", self.ps.h, ptrs(self.du, P3))
            if self.les:
                self._closure(u)
            for L in self.les:
                ub = L.cset.ubar
                c.call("ec_ns_project", L.ps64.h, ptrs(ub, P3))
                tau = NS.commutator_(L.cset, ub, self.visc)
                ul = L.u0 if st == 0 else L.us
                gl = L.g
                c.call("ec_ns_rhs", self.dt_code, gl.n, gl.l, self.visc, ptrs(ul, P3), ptrs(L.du, P3), 0)
                c.call("ec_ns_tensordivergence_nonsym", self.dt_code, gl.n, gl.l, _abi.EC_F64,
                       ptrs(tau, _abi.P9), ptrs(L.du, P3), 1)
                c.call("ec_ns_pressure_solve", L.ps.h, ptrs(L.du, P3))
                c.call("ec_ns_rk4_update", self.dt_code, gl.n, gl.l, st, dt, ptrs(L.u0, P3), ptrs(L.us, P3),
                       ptrs(L.du, P3), _abi.c_void_p(L.ps.p_ptr), ptrs(L.acc, P3))
            c.call("ec_ns_rk4_update", self.dt_code, n, l, st, dt, ptrs(self.u0, P3), ptrs(self.us, P3),
                   ptrs(self.du, P3), _abi.c_void_p(self.ps.p_ptr), ptrs(self.acc, P3))

    def _les_forcing(self, L, st, c=None):
        """m = V k - F_H(P_H v) in FP64 for the stage state of LES grid L.

        The face average maps a discretely divergence-free field to a coarse
        divergence-free one (its coarse fluxes are face sums, so div_H V u is
        the cell average of div_h u), and k is the projected DNS tendency: for
        a face-average grid v is projected once per step, at stage 0 (in
        place), and the later stage states v_s = v0 + a_s dt V k stay
        projected up to the rounding of k's divergence.  The volume average
        has no such property (P_H V u differs from V u at O(1)), so its v is
        projected at every stage, as apply_DHpi does."""
        c = c or self.ctx
        gl = L.g
        v = L.v_stage(st)
        if L.kind == "fa":
            if st == 0:
                c.call("ec_ns_project", L.ps64.h, ptrs(L.v0, P3))
            ub = v
        else:
            c.call("ec_ns_project_into", L.ps64.h, ptrs(v, P3), ptrs(L.ubar, P3))
            ub = L.ubar
        # m = V k - F_H(P_H v) = -F_H(ubar) + g_k, formed in the RHS pass
        c.call("ec_ns_rhs_add", _abi.EC_F64, gl.n, gl.l, self.visc, ptrs(ub, P3), ptrs(L.m, P3), -1.0,
               ptrs(L.g_k, P3))
        return L.m

    def _stage_states(self):
        """[u1, u2, u3] of the stored-state RK4 (the buffers us, acc, u3)."""
        return [self.us, self.acc, self.u3]

    def _step_forcing(self, dt):
        c = self.ctx
        n, l = self.g.n, self.g.l
        ust = self._stage_states() if self.rk_states else None
        ust_p = ptrs([x for s in ust for x in s], _abi.P9) if ust else None
        for st in range(4):
            u = self.u0 if st == 0 else (ust[st - 1] if ust else self.us)
            # du = -div sigma(u) and its divergence in one pass, then the FFT solve
            # at the last stage u_new = u0 + dt (acc + k / 6): with d = div du +
            # (6/dt) div u_3, div k = -(6/dt) div u_3 cancels the divergence
            # that u0 (= u_3 up to projected terms) brought into the step
            rdiv = 6.0 / dt if (st == 3 and self.clean_divergence) else 0.0
            c.call("ec_ns_rhs_div", self.dt_code, n, l, self.visc, ptrs(u, P3), ptrs(self.du, P3),
                   _abi.c_void_p(self.ps.p_ptr), rdiv)
            c.call("ec_poisson_solve", self.ps.h)
            # V_s(du - grad p) for every LES grid and the DNS RK4 stage in one
            # pass over du and p (nothing below reads the DNS state)
            if self.les:
                # (stage hooks read the LES state between stages: one stream then)
                lc = self.les_ctx if not self.stage_hooks else None
                if lc is not None and st > 0:
                    c.call("ec_stream_after", lc.h)  # the previous LES stage has read g_k
                sets = self._vel_sets([L.g_k for L in self.les])
                if ust:
                    c.call("ec_ns_filter_vel_rk4_states", self.dt_code, n, l, ptrs(self.du, P3),
                           _abi.c_void_p(self.ps.p_ptr), len(self.les), sets, st, dt, ptrs(self.u0, P3), ust_p)
                else:
                    c.call("ec_ns_filter_vel_rk4", self.dt_code, n, l, ptrs(self.du, P3),
                           _abi.c_void_p(self.ps.p_ptr), len(self.les), sets, st, dt, ptrs(self.u0, P3),
                           ptrs(self.us, P3), ptrs(self.acc, P3))
                if lc is None:
                    self._les_stage(st, dt, c)
                else:
                    lc.call("ec_stream_after", c.h)  # this stage's filter outputs are written
                    self._les_stage(st, dt, lc)
                    if st == 3:
                        c.call("ec_stream_after", lc.h)  # the step ends with the LES stages
            elif ust:
                c.call("ec_ns_rk4_update_states", self.dt_code, n, l, st, dt, ptrs(self.u0, P3), ust_p,
                       ptrs(self.du, P3), _abi.c_void_p(self.ps.p_ptr))
            else:
                c.call("ec_ns_rk4_update", self.dt_code, n, l, st, dt, ptrs(self.u0, P3), ptrs(self.us, P3),
                       ptrs(self.du, P3), _abi.c_void_p(self.ps.p_ptr), ptrs(self.acc, P3))
            for h in self.stage_hooks:
                h.stage(st, dt)

    def _les_stage(self, st, dt, c):
        """Stage st of every LES grid (forcing form) on context c: the exact
        forcing from the stage's filter outputs, the LES RHS with it, the
        projection and the RK4 updates of the LES state and of v = V u."""
        for L in self.les:
            gl = L.g
            m = self._les_forcing(L, st, c)
            ul = L.u_stage(st)
            # du = F_H(u) + m in one pass (the sum in FP64, rounded once)
            c.call("ec_ns_rhs_add", self.dt_code, gl.n, gl.l, self.visc, ptrs(ul, P3), ptrs(L.du, P3), 1.0,
                   ptrs(m, P3))
            c.call("ec_ns_pressure_solve", L.ps.h, ptrs(L.du, P3))
            c.call("ec_ns_rk4_update", self.dt_code, gl.n, gl.l, st, dt, ptrs(L.u0, P3), ptrs(L.us, P3),
                   ptrs(L.du, P3), _abi.c_void_p(L.ps.p_ptr), ptrs(L.acc, P3))
            # v = V u follows the DNS through the same tableau (V k is already
            # projected: no pressure)
            if L.rk_states:
                vst = ptrs(L.vs + L.vacc + L.v3, _abi.P9)
                c.call("ec_ns_rk4_update_states", _abi.EC_F64, gl.n, gl.l, st, dt, ptrs(L.v0, P3), vst,
                       ptrs(L.g_k, P3), None)
            else:
                c.call("ec_ns_rk4_update", _abi.EC_F64, gl.n, gl.l, st, dt, ptrs(L.v0, P3), ptrs(L.vs, P3),
                       ptrs(L.g_k, P3), ptr(L.zero64), ptrs(L.vacc, P3))

    def run(self, dt, nsteps):
        for _ in range(nsteps):
            self.step(dt)

    # -- diagnostics -------------------------------------------------------
    def divergence_norm(self):
        """h * ||div_h u||_2 / ||u||_2 of the DNS state, FP64 reductions."""
        d = torch.empty_like(self.u0[0])
        self.ctx.call("ec_ns_divergence", self.dt_code, self.g.n, self.g.l, ptrs(self.u0, P3), ptr(d))
        num = NS.sumsq(d)
        den = sum(NS.sumsq(x) for x in self.u0)
        h = self.g.l / self.g.n
        return h * (num ** 0.5) / (den ** 0.5)

    def closure_terms(self, u=None):
        """The FP64 closure of state u (default: the current state) for every
        LES grid, without advancing anything: the forcing m = V k - F_H(P_H V u)
        (3 components) with closure="forcing", the flux tau (9) with "flux"."""
        if self.closure == "flux":
            return self.closure_taus(u)
        c = self.ctx
        u = self.u0 if u is None else u
        n, l = self.g.n, self.g.l
        c.call("ec_ns_rhs", self.dt_code, n, l, self.visc, ptrs(u, P3), ptrs(self.du, P3), 0)
        c.call("ec_ns_pressure_solve", self.ps.h, ptrs(self.du, P3))
        self._filter_vel(self.du, self.ps.p_ptr, [L.g_k for L in self.les])
        self._filter_vel(u, None, [L.v0 for L in self.les])
        return [[t.clone() for t in self._les_forcing(L, 0)] for L in self.les]

    def closure_taus(self, u=None):
        """Closure commutator tau (9 FP64 components per LES grid) of state u
        (default: the current state), without advancing anything."""
        c = self.ctx
        u = self.u0 if u is None else u
        n, l = self.g.n, self.g.l
        c.call("ec_ns_rhs", self.dt_code, n, l, self.visc, ptrs(u, P3), ptrs(self.du, P3), 0)
        c.call("ec_ns_pressure_solve", self.ps.h, ptrs(self.du, P3))
        self._closure(u)
        out = []
        for L in self.les:
            ub = L.cset.ubar
            c.call("ec_ns_project", L.ps64.h, ptrs(ub, P3))
            out.append([t.clone() for t in NS.commutator_(L.cset, ub, self.visc)])
        return out

    def les_errors(self):
        """compute_errors of every LES grid against P_H V u_dns."""
        ref = self.filtered_dns()
        out = []
        for L, r in zip(self.les, ref):
            s = 0.0
            for a in range(3):
                la = L.u0[a].double()
                s += float(((la - r[a]) ** 2).sum() / (r[a] ** 2).sum())
            out.append((s / 3) ** 0.5)
        return out


def precision_study(n, visc, sets, nsteps, seed=123, cfl=0.4, l=1.0, dt=None, timing=False):
    """The FP32 pipeline (FP32 fields, FP64 closure commutator and reductions)
    against the plain FP64 form, both from the same seeded initial field
    (randomfield, src/N.jl:394-476, spectrum k^4 exp(-2 (k/5)^2), energy 1/2)
    and advanced `nsteps` RK4 steps with dt from propose_timestep (src/N.jl:
    348-356) at t = 0 unless given.  Returns a dict with
      velocity_relerr   compute_errors (src/N.jl:1334-1349) of the FP32 DNS
                        velocity against the FP64 one after the steps;
      divergence_f32/f64 h ||div_h u||_2 / ||u||_2 of each final state;
      closure_relerr    ||m_32 - m_64|| / ||m_64|| over every LES grid and
                        component, the FP64 closure evaluated from the FP32
                        final state promoted to FP64 (the same input state);
      les_relerr        per LES grid, compute_errors of the FP32 LES against
                        the FP64 LES after the steps;
      exactness_f32/f64 per LES grid, compute_errors of the LES against the
                        filtered DNS of its own run.
    With timing, also the mean ms per step of each form."""
    import time

    f32, f64 = torch.float32, torch.float64
    g = NS.Grid(l, n)
    ps64 = NS.PoissonSolver(g, f64)
    u64 = NS.randomfield(g, ps64, "peak", 5.0, 0.5, seed=seed)
    if dt is None:
        dt = NS.propose_timestep(u64, visc, cfl)
    d64 = DNSAidedLES(g, visc, sets, dtype=f64)
    d64.set_state([x.data for x in u64])
    d32 = DNSAidedLES(g, visc, sets, dtype=f32)
    d32.set_state([x.data.to(f32) for x in u64])
    del u64, ps64
    out = {"n": n, "visc": visc, "dt": dt, "steps": nsteps, "sets": [f"{nb}:{k}" for nb, k in sets]}
    for name, drv in (("f64", d64), ("f32", d32)):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        drv.run(dt, nsteps)
        torch.cuda.synchronize()
        if timing:
            out[f"ms_per_step_{name}"] = 1e3 * (time.perf_counter() - t0) / max(nsteps, 1)
    u32d = [x.to(f64) for x in d32.u0]
    out["velocity_relerr"] = NS.compute_errors(u32d, d64.u0)
    out["divergence_f32"] = d32.divergence_norm()
    out["divergence_f64"] = d64.divergence_norm()
    out["les_relerr"] = [NS.compute_errors([x.to(f64) for x in L32.u0], L64.u0)
                         for L32, L64 in zip(d32.les, d64.les)]
    out["exactness_f32"] = d32.les_errors()
    out["exactness_f64"] = d64.les_errors()
    m32 = d32.closure_terms()
    m64 = d64.closure_terms(u32d)
    num = den = 0.0
    for s32, s64 in zip(m32, m64):
        for x, y in zip(s32, s64):
            num += float(((x - y) ** 2).sum())
            den += float((y ** 2).sum())
    out["closure_relerr"] = (num / den) ** 0.5
    return out
