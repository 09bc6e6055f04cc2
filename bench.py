"""DNS-aided LES benchmark on one MI355X.

Workload: 3D incompressible Navier-Stokes DNS at 1024^3 (FP32 fields,
viscosity 1/8000), classical RK4, exact closure evaluated at every stage for
four LES filter sets (volume average at 64^3, 128^3, 256^3 and face average at
256^3), each LES grid advanced with its closure.  Closure commutators and
energy reductions run in FP64.  Initial field: random, spectrum
k^4 exp(-2 (k/5)^2), energy 1/2, seed 123, projected divergence-free.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 1024] [--dtype f32]
    python bench.py --burgers       # the batched Burgers workload instead

Prints one JSON line.  value = DNS-aided LES RK4 steps per second.  Launch
shapes are picked by the project's autotuner before the timed region.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))

METRIC = "DNS-aided LES steps/s + HBM GB/s at 1024³ FP32; closure rel-err vs FP64"
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--dtype", default="f32", choices=["f32", "f64"])
    ap.add_argument("--sets", default="16:va,8:va,4:va,4:fa", help="compression:kind list")
    ap.add_argument("--visc", type=float, default=1.0 / 8000)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-closure-check", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=3,
                    help="timed steps of the plain FP64 form of the same step after the FP32 leg (0: skip)")
    ap.add_argument("--reference-setup", action="store_true",
                    help="run the reference's own comparison (dnsaid_project at n_dns 500, n_les 100, nu 2e-4, "
                         "Delta = 2H Gaussian, forward Euler, CFL 0.4, warm-up and run to t = 0.1) instead")
    ap.add_argument("--precision-study", action="store_true",
                    help="run the 512^3 FP32-vs-FP64 study (nu = 1/4000, 50 steps, 64^3 and 128^3 closures) instead")
    ap.add_argument("--cpu-n", type=int, default=256)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a plain 1 GiB read + 1 GiB write kernel first (HBM counter calibration)")
    ap.add_argument("--kernels", action="store_true", help="print the per-kernel table to stderr")
    ap.add_argument("--closure-tiles", default="16,16,128", help="tx,ty,tz block targets of the closure kernel")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--eager", action="store_true", help="time eager launches instead of one graph replay per step")
    ap.add_argument("--les-stream", type=int, default=1, choices=[0, 1],
                    help="1: the LES grids' stage work on a second stream beside the DNS passes")
    ap.add_argument("--burgers", action="store_true",
                    help="time the batched Burgers workload (4096 instances, 1000 RK4 steps) instead")
    ap.add_argument("--burgers-steps", type=int, default=1000)
    ap.add_argument("--burgers-study", action="store_true",
                    help="time the Burgers closure study (closure fractions, Smagorinsky fit and solve) "
                         "at nh = 13500, nH in {300, 900, 2700}, Delta = 2H Gaussian, 1000 samples")
    return ap.parse_args()


def traffic_from_profile(kernel):
    """HBM bytes per launch of `kernel` at the benchmark size from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE, copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_1024_f32.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0] == kernel:
            try:
                return (float(cells[1]) + float(cells[2])) * 1e9
            except ValueError:
                return None
    return None


def burgers_main(a):
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge

    ec = ge.load_package()
    r = ec.burgers.benchmark(nsteps=a.burgers_steps, tune=not a.no_autotune)
    out = {"metric": "batched Burgers DNS-aided LES point-steps/s", "value": r["point_steps_per_s"],
           "unit": "point-steps/s", "n_gpus": 1, "steps": r["steps"], "warmup": 1, "ms_per_step": r["ms"] / r["steps"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32 state + fp64 commutator",
           "data": "synthetic: modes k<=10, amplitude 1/k, uniform phase, instance seed = index",
           "config": {"workload": "4096 Burgers instances, N in {2048,4096,8192,16384} (40/30/20/10 %), "
                                  "c in {4..64}, nu=5e-4, RK4, exact + no-model LES, one graph-replayed launch",
                      "instances": r["instances"], "points": r["points"], "threads": r["threads"]}}
    print(json.dumps(out), flush=True)


def burgers_study_main(a):
    """compute_fractions (src/Burgers.jl:1022-1099), compute_dissipation
    (src/Burgers.jl:447-516), the Smagorinsky fits (smagorinsky.jl:24-36) and
    solve_smagorinsky (src/Burgers.jl:682-772) on
    the reference setup (src/Burgers.jl:236-246), all samples on the GPU in
    FP64.  The data come from the reference's own pipeline: 1000 bump-profile
    initial fields (randomfield, src/Burgers.jl:192-202, energy 2, kpeak 10)
    and dns_aided_les (src/Burgers.jl:350-445: forward Euler, CFL 0.3, to
    t = 0.1) for each nH, whose LES errors (compute_errors,
    src/Burgers.jl:909-917) are reported; the closure study then runs on the
    DNS samples at t = 0.1.  The oracle's numpy restatement on 10 samples is
    the CPU leg."""
    sys.path.insert(0, ROOT)
    import math

    import numpy as np
    import torch
    import __graft_entry__ as ge

    ec = ge.load_package()
    B = ec.burgers
    L, visc, nh, ns, dr, tstop = 2 * math.pi, 5e-4, 13500, 1000, 2.0, 0.1
    ustart = B.randomfield(nh, 10, 2.0, ns, seed=1, L=L)
    per = {}
    total = 0.0
    U = None
    les_errors = {}
    for nH in (300, 900, 2700):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f = B.dns_aided_les(ustart, nH, visc, dr * L / nH, tstop, 0.3, "gaussian", L)
        t_les = time.perf_counter() - t0
        err = B.compute_errors([{"nH": nH, "fields": f}])
        les_errors[str(nH)] = {k: v[0] for k, v in err.items()}
        per[str(nH)] = {"dns_aided_les_ms": 1e3 * t_les, "steps_mean": float(np.mean(f["steps"]))}
        log(f"[bench] dns_aided_les nH={nH}: {1e3 * t_les:.1f} ms, errors {les_errors[str(nH)]}")
        if U is None:
            U = f["dns_ref"].contiguous()
        del f
    for nH in (300, 900, 2700):
        B.compute_fractions(U, nH, L, visc, dr, "gaussian")  # warm
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fr = B.compute_fractions(U, nH, L, visc, dr, "gaussian")
        t1 = time.perf_counter()
        di = B.compute_dissipation(U, nH, L, visc, dr, "gaussian")
        t1d = time.perf_counter()
        th = B.smagorinsky_coefficients(U, nH, L, visc, dr, "gaussian")
        t2 = time.perf_counter()
        B.solve_smagorinsky(U, nH, th["classic"], L, visc, dr, "gaussian", tstop)
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        per[str(nH)].update({"fractions_ms": 1e3 * (t1 - t0), "dissipation_ms": 1e3 * (t1d - t1),
                             "fit_ms": 1e3 * (t2 - t1d), "solve_ms": 1e3 * (t3 - t2), "fractions": fr,
                             "mean_dissipation": {k: float(di[k].mean()) for k in ("class_m", "class_p", "swapfil")},
                             "theta_classic": th["classic"], "theta_discret": th["discret"]})
        del di
        total += t3 - t0
    cpu = None
    if not a.no_cpu_baseline:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O

        les, dbl = B.les_kernels(nh, 300, L, dr, "gaussian")
        Uh = U[:10].cpu().numpy()
        t0 = time.perf_counter()
        O.burgers_fractions(Uh, 300, L, visc, les, dbl)
        O.burgers_dissipation(Uh, 300, L, visc, les, dbl)
        O.smagorinsky_sums(Uh, 300, L, visc, dr * L / 300, les, dbl)
        H = L / 300
        O.smagorinsky_solve(Uh, 300, L, visc, per["300"]["theta_classic"] ** 2 * ((dr * H) ** 2 + H ** 2), tstop,
                            0.3, dbl)
        tc = time.perf_counter() - t0
        cpu = {"value": 1e3 * tc * ns / 10, "unit": "ms per 1000 samples (nH = 300 only)", "cores": 1,
               "kind": "port", "sample": f"numpy oracle on 10 samples at nH = 300: {tc:.2f} s, scaled by 100"}
    out = {"metric": "Burgers closure study: fractions + dissipation + Smagorinsky fit + solve, ms per 1000 samples",
           "value": 1e3 * total, "unit": "ms", "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": 1e3 * total,
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
           "data": "synthetic: 1000 bump-profile fields (kpeak 10, energy 2, counter-based phases), "
                   "DNS to t = 0.1 by dns_aided_les",
           "les_errors": les_errors,
           "config": {"workload": "nh=13500, nH in {300, 900, 2700}, Delta=2H Gaussian, 1000 samples, "
                                  "Smagorinsky LES to t=0.1 (CFL 0.3)", "per_nH": per},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n=256, nsteps=5):
    """The FP64 C oracle (OpenMP, all host threads) on the 256^3 DNS-aided RK4
    step with the 32^3 volume-average closure (the reference's CPU
    configuration); steps/s scaled to 1024^3 by the cell count, since the
    work per cell of every pass is independent of n (FFT: log n)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O

    visc, dt = 1.0 / 2000, 1e-3
    u = O.randomfield(n, seed=123)
    t0 = time.perf_counter()
    O.dnsaid_rk4(u, [(8, "va")], visc, dt, nsteps)
    t = time.perf_counter() - t0
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    sps = nsteps / t
    return {"value": sps * (n / 1024) ** 3, "unit": "steps/s (1024^3-equivalent)", "cores": threads,
            "kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "sample": f"{nsteps} RK4 DNS-aided steps at {n}^3 FP64 with the 32^3 closure: {t:.2f} s "
                      f"({sps:.4f} steps/s at {n}^3), scaled by (n/1024)^3"}


def copy_ceiling_gbs(abi, nbytes=1 << 31, reps=5):
    """Attainable streaming rate of this device: a plain device-to-device copy
    of nbytes (read nbytes, write nbytes) in GB/s, the faster of torch's copy
    kernel and the runtime's (hipMemcpyAsync through ec_d2d), each the mean of
    reps copies between two synchronisations.  The step's kernels each read
    and write several fields, so this (not the 8 TB/s of the data sheet) is the
    rate a perfectly streaming kernel reaches here."""
    import torch

    ctx = abi.context()
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    best = 0.0
    for f in (lambda: y.copy_(x),
              lambda: ctx.call("ec_d2d", abi.c_void_p(y.data_ptr()), abi.c_void_p(x.data_ptr()), nbytes)):
        f()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            f()
        torch.cuda.synchronize()
        best = max(best, 2.0 * nbytes * reps / (time.perf_counter() - t0) / 1e9)
    del x, y
    torch.cuda.empty_cache()
    return best


def fp64_leg(ec, n, visc, sets, dt, steps):
    """The plain FP64 form of the same DNS-aided step (the same kernels with
    FP64 fields), on the same seeded initial field: mean ms per step over
    `steps` timed steps after one warm-up step."""
    import torch

    NS, D = ec.navierstokes, ec.dnsaid
    g = NS.Grid(1.0, n)
    ps64 = NS.PoissonSolver(g, torch.float64)
    u64 = NS.randomfield(g, ps64, "peak", 5.0, 0.5, seed=123)
    del ps64
    drv = D.DNSAidedLES(g, visc, sets, dtype=torch.float64)
    drv.set_state([x.data for x in u64])
    del u64
    drv.step(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        drv.step(dt)
    torch.cuda.synchronize()
    ms = 1e3 * (time.perf_counter() - t0) / steps
    del drv
    torch.cuda.empty_cache()
    return ms


def reference_setup_main(a):
    """getsetup + dnsaid_project + compute_errors (navierstokes.jl:10-38): the
    five LES errors of the reference's own comparison, FP64, on the GPU."""
    sys.path.insert(0, ROOT)
    import torch
    import __graft_entry__ as ge

    ec = ge.load_package()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = ec.dnsaid_project.reference_setup()
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    out = {"metric": "dnsaid_project (reference setup): LES errors against the filtered DNS",
           "value": r["errors"]["u_cfd"], "unit": "relative L2 (exact closure)", "n_gpus": 1,
           "steps": r["steps"], "warmup": 0, "ms_per_step": 1e3 * t / r["steps"], "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp64",
           "data": "synthetic: spectrum (k>0) k^(-5/3), E=1, counter-based normal variates (seed 0)",
           "config": {"workload": "n_dns 500, n_les 100, nu 2e-4, Gaussian Delta = 2H (nsigma 2), forward Euler, "
                                  "CFL 0.4, warm-up to t = 0.1, then to t = 0.1", **r["setup"]},
           "errors": r["errors"], "wall_s": t,
           "reference_pins": {"u_nomo": 0.263051, "u_c": 0.211564, "u_cf": 0.0642457, "u_cfd": 5.68942e-15,
                              "u_cfd_symm": 0.0285891}}
    print(json.dumps(out), flush=True)


def precision_main(a):
    """The 512^3 FP32 pipeline against the plain FP64 form (velocity,
    divergence and closure errors after 50 steps from the same seed)."""
    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge

    ec = ge.load_package()
    r = ec.dnsaid.precision_study(512, 1.0 / 4000, [(64, "va"), (128, "va")], 50, timing=True)
    out = {"metric": "FP32 pipeline vs FP64 form at 512^3 after 50 RK4 steps: velocity rel-L2 error",
           "value": r["velocity_relerr"], "unit": "relative L2", "n_gpus": 1, "steps": 50, "warmup": 0,
           "ms_per_step": r["ms_per_step_f32"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32 fields + fp64 closure vs fp64", "data": "synthetic: spectrum k^4 exp(-2(k/5)^2), E=1/2, seed 123",
           "config": {"workload": "NS 512^3, nu=1/4000, RK4 50 steps, exact closure at 64^3 and 128^3 (volume average)"},
           "study": r}
    print(json.dumps(out), flush=True)


def main():
    a = parse()
    if a.burgers and a.gpus == 1:
        return burgers_main(a)
    if a.burgers_study and a.gpus == 1:
        return burgers_study_main(a)
    if a.precision_study and a.gpus == 1:
        return precision_main(a)
    if a.reference_setup and a.gpus == 1:
        return reference_setup_main(a)
    if a.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {a.gpus}); sharding the step over several "
              "GPUs is not part of this build", file=sys.stderr)
        sys.exit(2)
    import torch

    sys.path.insert(0, ROOT)
    import __graft_entry__ as ge

    ec = ge.load_package()
    NS, D, abi = ec.navierstokes, ec.dnsaid, ec._abi
    n = a.n
    dtype = torch.float32 if a.dtype == "f32" else torch.float64
    sets = []
    for tok in a.sets.split(","):
        c, kind = tok.split(":")
        sets.append((n // int(c), kind))
    g = NS.Grid(1.0, n)

    copy_gbs = copy_ceiling_gbs(abi)
    if a.copy_calibration:
        # a plain streaming kernel of known traffic for the HBM counters: reads
        # 1 GiB and writes 1 GiB (element-wise negation of 2^28 floats)
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.neg(x, out=y)
        torch.cuda.synchronize()
        del x, y

    # initial field in FP64, then the run's precision
    t0 = time.perf_counter()
    ps64 = NS.PoissonSolver(g, torch.float64)
    u64 = NS.randomfield(g, ps64, "peak", 5.0, 0.5, seed=123)
    dt = NS.propose_timestep(u64, a.visc, 0.4)
    u_init = [x.data.to(dtype) for x in u64]
    del u64, ps64
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log(f"[bench] IC {n}^3 ready in {time.perf_counter() - t0:.1f}s, dt={dt:.3e}")

    drv = D.DNSAidedLES(g, a.visc, sets, dtype=dtype, les_stream=bool(a.les_stream))
    if a.closure_tiles:
        drv.ctx.call("ec_closure_set_tiles", *[int(x) for x in a.closure_tiles.split(",")])
    drv.set_state(u_init)
    del u_init
    for _ in range(a.warmup):
        drv.step(dt)
    torch.cuda.synchronize()
    log(f"[bench] warmup {a.warmup} steps done")
    tuned = {}
    if not a.no_autotune:
        tuner = ec.autotune.Autotuner()
        tuned["filter_vel_rk4_zchunk"] = ec.autotune.tune_filter_vel(tuner, drv, dt)
        torch.cuda.synchronize()
        log(f"[bench] autotune: {tuner.log}")

    ctx = drv.ctx
    # the timed steps replay one captured graph per step (every launch of the
    # step, the LES grids included); a first replay runs before the timing
    graph = not a.eager
    if graph:
        drv.capture(dt)
        drv.step(dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        drv.step(dt)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    log(f"[bench] {a.steps} {'graph-replayed' if graph else 'eager'} steps: {1e3 * t / a.steps:.1f} ms/step")
    # per-launch HIP events (ec_profile_*) on the context stream: the same
    # step issued eagerly, K more steps, for the kernel table and the roofline
    drv.release_graphs()
    # (the LES stage work runs on the driver's second context when it has one)
    pctx = [ctx] + ([drv.les_ctx] if drv.les_ctx is not None else [])
    for c in pctx:
        c.call("ec_profile_begin")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        drv.step(dt)
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t0
    cap = 100000
    names, ms, by = [], [], []
    for c in pctx:
        nm_c = (abi.ctypes.c_char_p * cap)()
        ms_c = (abi.c_double * cap)()
        by_c = (abi.c_double * cap)()
        cnt = abi.c_int()
        c.call("ec_profile_end", nm_c, ms_c, by_c, cap, abi.ctypes.byref(cnt))
        k = min(cnt.value, cap)
        names += [nm_c[i] for i in range(k)]
        ms += ms_c[:k]
        by += by_c[:k]
    m = len(names)
    per = {}
    for i in range(m):
        k = names[i].decode()
        e = per.setdefault(k, [0, 0.0, 0.0])
        e[0] += 1
        e[1] += ms[i]
        e[2] += by[i]
    total_bytes = sum(v[2] for v in per.values()) / a.steps
    ms_step = 1000 * t / a.steps
    dom = max(per.items(), key=lambda kv: kv[1][1])
    dname = dom[0]
    # the dominant kernel's DNS-size launches (the LES-grid launches share its name)
    big = [(ms[i], by[i]) for i in range(m) if names[i].decode() == dname]
    bmax = max(b for _, b in big)
    big = [(t_, b) for t_, b in big if b >= 0.5 * bmax]
    dcount = len(big)
    dms = sum(t_ for t_, _ in big)
    dbytes = sum(b for _, b in big)
    d_avg_ms = dms / dcount
    d_achieved = (dbytes / dcount) / (d_avg_ms * 1e-3) / 1e9
    if a.kernels:
        for k, (c, tms, b) in sorted(per.items(), key=lambda kv: -kv[1][1]):
            log(f"  {k:22s} launches {c:6d}  total {tms:9.2f} ms  avg {tms / c:8.3f} ms  "
                f"{(b / c) / (tms / c * 1e-3) / 1e9:8.1f} GB/s")
    log(f"[bench] {a.steps} steps in {t:.3f}s: {ms_step:.1f} ms/step ({1e3 * t_eager / a.steps:.1f} eager with "
        f"events), algorithmic {total_bytes / 1e9:.1f} GB/step")

    closure_err = None
    les_err = drv.les_errors()
    div = drv.divergence_norm()
    # energy spectrum of the final DNS state (spectral_stuff + spectrum,
    # src/N.jl:547-590): three forward FFTs and the FP64 shell sums; outside
    # the timed steps (the entry point synchronises to return the sums)
    stuff = NS.spectral_stuff(g)
    uf = tuple(NS.Field(g, x) for x in drv.u0)
    spec = NS.spectrum(uf, stuff, drv.ps)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        spec = NS.spectrum(uf, stuff, drv.ps)
    t_spec = (time.perf_counter() - t0) / 3
    bsz = 4 if a.dtype == "f32" else 8
    spec_bytes = 3 * (bsz * n ** 3 + 6 * 2 * bsz * (n // 2 + 1) * n * n)
    spectrum_diag = {"ms": 1e3 * t_spec, "gbs": spec_bytes / t_spec / 1e9, "shells": len(stuff["shells"]),
                     "energy_shells_1_to_kcut": sum(spec["s"])}
    del uf
    if not a.no_closure_check:
        taus32 = drv.closure_terms()
        u_state = [x.double() for x in drv.u0]
        del drv
        torch.cuda.empty_cache()
        d64 = D.DNSAidedLES(g, a.visc, sets, dtype=torch.float64)
        taus64 = d64.closure_terms(u_state)
        num = den = 0.0
        for s32, s64 in zip(taus32, taus64):
            for x, y in zip(s32, s64):
                num += float(((x - y) ** 2).sum())
                den += float((y ** 2).sum())
        closure_err = math.sqrt(num / den)
        del d64, taus64, taus32, u_state
    else:
        del drv
    torch.cuda.empty_cache()

    fp64_ms = None
    if a.fp64_steps > 0 and a.dtype == "f32":
        log(f"[bench] FP64 form: {a.fp64_steps} timed steps ...")
        fp64_ms = fp64_leg(ec, n, a.visc, sets, dt, a.fp64_steps)
        log(f"[bench] FP64 form {fp64_ms:.1f} ms/step against {ms_step:.1f} ms/step in FP32")

    cpu = None
    if not a.no_cpu_baseline:
        log("[bench] CPU baseline (oracle, OpenMP) ...")
        cpu = cpu_baseline(a.cpu_n)

    out = {
        "metric": METRIC,
        "value": a.steps / t,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_step,
        "step_launch": "one hipGraph replay per step" if graph else "eager launches",
        "eager_ms_per_step": 1e3 * t_eager / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 fields + fp64 closure" if a.dtype == "f32" else "fp64",
        "data": "synthetic: random divergence-free field, spectrum k^4 exp(-2(k/5)^2), E=1/2, seed 123",
        "config": {"workload": f"DNS-aided LES, NS {n}^3, RK4, nu={a.visc:.3g}, dt={dt:.3e} "
                               f"(propose_timestep, CFL 0.4), closure sets "
                               + ", ".join(f"{nb}^3 {k}" for nb, k in sets),
                   "grid": n, "les": [f"{nb}:{k}" for nb, k in sets], "parallelism": "single GPU"},
        "hbm_gbs": total_bytes / (t / a.steps) / 1e9,
        "hbm_frac_step": total_bytes / (t / a.steps) / 1e9 / PEAK_HBM_GBS,
        "closure_relerr_vs_fp64": closure_err,
        "divergence": div,
        "les_error_vs_filtered_dns": les_err,
        "fp64_ms_per_step": fp64_ms,
        "speedup_vs_fp64": (fp64_ms / ms_step) if fp64_ms else None,
        "spectrum": spectrum_diag,
        "roofline": {"bound": "hbm", "kernel": dname, "achieved": d_achieved, "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": d_achieved / PEAK_HBM_GBS, "launches": dcount,
                     "avg_ms": d_avg_ms, "bytes_per_launch": dbytes / dcount,
                     "traffic": traffic_from_profile(dname) if n == 1024 and a.dtype == "f32" else None,
                     "copy_ceiling": copy_gbs, "frac_of_copy_ceiling": d_achieved / copy_gbs},
        "autotuned": tuned,
        "kernels": {k: {"launches": c, "ms": round(tms, 3), "gbs": round((b / c) / (tms / c * 1e-3) / 1e9, 1)}
                    for k, (c, tms, b) in per.items()},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
