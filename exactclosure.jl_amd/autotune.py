"""Launch-shape autotuner.

Times each candidate shape of a tunable kernel on the device it will run on
(HIP events on the library's stream, median of several repetitions) and
keeps the fastest. Results are cached per (device, kernel, workload key) in
a JSON file, so a process tunes a given shape once. Tunables:

* ``filter_vel_rk4_zchunk``: planes per block of the one-launch velocity
  filter with the fused DNS RK4 stage (``ec_filter_vel_set_tiles``);
* ``closure_tiles``: x/y/z block targets of the flux-form closure kernel
  (``ec_closure_set_tiles``);
* ``burgers_threads``: threads per instance block of the persistent batched
  Burgers kernel (``ec_burgers_batch_set_threads``).
"""
from __future__ import annotations

import json
import os
import statistics

import numpy as np
import torch

DEFAULT_CACHE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_tune_cache.json")


class Autotuner:
    def __init__(self, path=None, reps=3, warmup=1):
        self.path = path or os.environ.get("EC_TUNE_CACHE", DEFAULT_CACHE)
        self.reps = reps
        self.warmup = warmup
        self.device = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
        self.cache = {}
        if os.path.exists(self.path):
            try:
                with open(self.path) as f:
                    self.cache = json.load(f)
            except (OSError, ValueError):
                self.cache = {}
        self.log = []

    def _key(self, kernel, key):
        return f"{self.device}|{kernel}|{key}"

    def time_ms(self, run):
        """Median wall time of `run()` in ms, measured with events on the current stream."""
        for _ in range(self.warmup):
            run()
        ts = []
        for _ in range(self.reps):
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record()
            run()
            b.record()
            b.synchronize()
            ts.append(a.elapsed_time(b))
        return statistics.median(ts)

    def tune(self, kernel, key, candidates, apply, run):
        """Pick the fastest candidate: `apply(c)` installs it, `run()` executes
        the workload once. The winner stays applied and is returned."""
        k = self._key(kernel, key)
        hit = self.cache.get(k)
        if hit is not None and hit["best"] in [_norm(c) for c in candidates]:
            best = _denorm(hit["best"], candidates)
            apply(best)
            self.log.append({"kernel": kernel, "key": key, "best": hit["best"], "cached": True})
            return best
        times = {}
        for c in candidates:
            apply(c)
            times[_norm(c)] = self.time_ms(run)
        best_n = min(times, key=times.get)
        best = _denorm(best_n, candidates)
        apply(best)
        self.cache[k] = {"best": best_n, "ms": times}
        self.log.append({"kernel": kernel, "key": key, "best": best_n, "ms": times, "cached": False})
        try:
            with open(self.path, "w") as f:
                json.dump(self.cache, f, indent=1, sort_keys=True)
        except OSError:
            pass
        return best


def _norm(c):
    return ",".join(str(x) for x in c) if isinstance(c, (tuple, list)) else str(c)


def _denorm(s, candidates):
    for c in candidates:
        if _norm(c) == s:
            return c
    raise KeyError(s)


def tune_filter_vel(tuner, drv, dt=None, candidates=(64, 128, 256, 512)):
    """Planes per block of the forcing-form filter with the fused DNS RK4
    stage, on the driver's own fields.  With dt, one timed run is a whole
    DNS-aided step (it advances the driver; call it between steps): the chunk
    changes how the filter's halo lines meet the other kernels' data in L2,
    and isolated launches have ranked the chunks differently from whole steps
    (stage 0 alone once chose 512 planes, 2 ms per step slower in the step
    than 256).  Without dt, one run is stages 0, 1 and 2 with dt = 0,
    which write only stage buffers (us and acc, or the stored states u1..u3;
    the next step rewrites them) and leave the DNS state u0 untouched."""
    from . import _abi
    from ._abi import P3, ptrs

    ctx = drv.ctx
    outs = [L.g_k for L in drv.les]
    n, l = drv.g.n, drv.g.l

    def run_stages():
        for st in range(3):
            if drv.rk_states:
                ust = ptrs([x for s in drv._stage_states() for x in s], _abi.P9)
                ctx.call("ec_ns_filter_vel_rk4_states", drv.dt_code, n, l, ptrs(drv.du, P3),
                         _abi.c_void_p(drv.ps.p_ptr), len(drv.les), drv._vel_sets(outs), st, 0.0, ptrs(drv.u0, P3),
                         ust)
            else:
                ctx.call("ec_ns_filter_vel_rk4", drv.dt_code, n, l, ptrs(drv.du, P3), _abi.c_void_p(drv.ps.p_ptr),
                         len(drv.les), drv._vel_sets(outs), st, 0.0, ptrs(drv.u0, P3), ptrs(drv.us, P3),
                         ptrs(drv.acc, P3))

    run = run_stages if dt is None else (lambda: drv.step(dt))
    cands = [z for z in candidates if n % z == 0] or [n]
    key = f"n{n}|" + ",".join(f"{L.g.n}{L.kind}" for L in drv.les) + ("|step" if dt is not None else "|stages")
    return tuner.tune("filter_vel_rk4_zchunk", key, cands, lambda z: ctx.call("ec_filter_vel_set_tiles", 16, int(z)),
                      run)


def tune_burgers_threads(tuner, batch, nsteps=20, candidates=(128, 256, 512, 1024)):
    """Threads per instance of the persistent Burgers kernel (advances the batch).
    A block size the library refuses for this batch (threads that cannot
    cover the largest instance with its largest points per thread) is
    skipped; the library's own check decides, so the rule lives in one place.
    The cache key holds the state precision and the compression mix, which
    set the LDS per block and with it the occupancy."""
    from ._abi import ECError

    nmax = int(batch.n.max())
    cands = []
    for t in candidates:
        try:
            batch.set_threads(t)
        except ECError:
            continue
        cands.append(t)
    cs = np.asarray(batch.c)
    mix = "-".join(f"{int(c)}x{int((cs == c).sum())}" for c in np.unique(cs))
    key = f"count{batch.count}|n{nmax}|{getattr(batch, 'dtype', 'f32')}|c{mix}"
    return tuner.tune("burgers_threads", key, cands, batch.set_threads, lambda: batch.run(nsteps, use_graph=True))
