"""Train the learned closure on pairs generated by the DNS-aided run itself.

    python exactclosure.jl_amd/train_closure.py [--n 1024] [--nbar 128] [--pairs 2000]
        [--every 1] [--epochs 20] [--batch 8] [--lr 1e-3] [--post-steps 500]
        [--store bf16] [--fallback-tol auto] [--out results.json]

1. DNS-aided LES (FP32 DNS, forcing-form exact closure, volume average) at
   n^3 with one LES grid of nbar^3; every `every` steps the pair
   (P_H V u, exact closure forcing m) of that grid is kept on the GPU, in BF16
   by default (2000 pairs at 128^3: 50 GB).
2. The CNN (3 -> 32 -> 32 -> 32 -> 3, 3x3x3, GELU, periodic) is trained with
   Adam on 7/8 of the pairs (the earliest), in BF16 with FP32 accumulation.
3. The closure error of the held-out (latest) pairs is measured against the exact
   closure (relative L2; the LES without a model scores 1).
4. A posteriori: from the filtered DNS state, three LES run `post-steps` RK4
   steps beside the DNS: no model, the network, and the network with the
   exact-closure fallback (a stage whose prediction misses the exact closure
   by more than --fallback-tol uses the exact closure; "auto" takes the median
   held-out closure error, so the fallback is exercised).
Prints one JSON line with the results (and writes it to --out when given).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--nbar", type=int, default=128)
    ap.add_argument("--visc", type=float, default=1.0 / 8000)
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--every", type=int, default=1)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--post-steps", type=int, default=500)
    ap.add_argument("--fallback-tol", default="auto", help="relative closure error, or auto (held-out median)")
    ap.add_argument("--store", choices=("bf16", "f32", "f64"), default="bf16", help="precision of the stored pairs")
    ap.add_argument("--out", default=None, help="also write the JSON result to this file")
    ap.add_argument("--seed", type=int, default=123)
    a = ap.parse_args()

    import torch

    sys.path.insert(0, os.path.dirname(HERE))
    import __graft_entry__ as ge

    ec = ge.load_package()
    NS, D, CN = ec.navierstokes, ec.dnsaid, ec.cnn
    log = lambda *x: print(*x, file=sys.stderr, flush=True)  # noqa: E731

    g = NS.Grid(1.0, a.n)
    ps64 = NS.PoissonSolver(g, torch.float64)
    u64 = NS.randomfield(g, ps64, "peak", 5.0, 0.5, seed=a.seed)
    dt = NS.propose_timestep(u64, a.visc, 0.4)
    u32 = [x.data.float() for x in u64]
    del u64, ps64
    torch.cuda.empty_cache()
    drv = D.DNSAidedLES(g, a.visc, [(a.nbar, "va")], dtype=torch.float32)
    drv.set_state(u32)
    del u32
    t0 = time.perf_counter()
    store = {"bf16": torch.bfloat16, "f32": torch.float32, "f64": None}[a.store]
    X, Y = CN.generate_pairs(drv, 0, dt, a.pairs * a.every, every=a.every, log=log, dtype=store)
    torch.cuda.synchronize()
    t_gen = time.perf_counter() - t0
    log(f"[train] {len(X)} pairs in {t_gen:.1f} s ({a.pairs * a.every} DNS steps at {a.n}^3)")

    ntr = len(X) * 7 // 8
    net = CN.CNNClosure(seed=a.seed)
    t0 = time.perf_counter()
    hist = net.fit(X[:ntr], Y[:ntr], epochs=a.epochs, batch=a.batch, lr=a.lr, seed=0, log=log)
    torch.cuda.synchronize()
    t_fit = time.perf_counter() - t0
    errs = [CN.closure_error(net.predict(CN._f64(x)), y) for x, y in zip(X[ntr:], Y[ntr:])]
    log(f"[train] fit {t_fit:.1f} s, held-out closure error {sum(errs) / len(errs):.3f}")
    del X, Y
    torch.cuda.empty_cache()
    tol = sorted(errs)[len(errs) // 2] if a.fallback_tol == "auto" else float(a.fallback_tol)

    plain = CN.LearnedLES(drv, 0, None)
    learned = CN.LearnedLES(drv, 0, net, fallback_tol=None)
    guarded = CN.LearnedLES(drv, 0, net, fallback_tol=tol)
    drv.stage_hooks = [plain, learned, guarded]
    post = []
    for s in range(a.post_steps):
        drv.step(dt)
        if (s + 1) % max(1, a.post_steps // 10) == 0:
            post.append({"step": s + 1, "no_model": plain.error(), "cnn": learned.error(),
                         "cnn_fallback": guarded.error()})
            log(f"[post] {post[-1]}")
    drv.stage_hooks = []
    flops = 2 * a.nbar ** 3 * 27 * (3 * 32 + 2 * 32 * 32 + 32 * 3)
    out = {"dns": a.n, "les": a.nbar, "pairs": len(errs) + ntr, "train_pairs": ntr, "epochs": a.epochs,
           "batch": a.batch, "lr": a.lr, "store": a.store, "generate_s": t_gen, "fit_s": t_fit,
           "train_tflops": 3 * flops * ntr * a.epochs / t_fit / 1e12,
           "loss_first_epoch": hist[0], "loss_last_epoch": hist[-1],
           "heldout_closure_error": sum(errs) / len(errs), "heldout_closure_error_no_model": 1.0,
           "heldout_closure_error_range": [min(errs), max(errs)], "fallback_tol": tol,
           "a_posteriori": post, "fallback_stages": guarded.fallbacks, "model_stages": guarded.stages,
           "post_steps": a.post_steps, "dt": dt, "visc": a.visc}
    line = json.dumps(out)
    print(line, flush=True)
    if a.out:
        with open(a.out, "w") as fh:
            fh.write(line + "\n")


if __name__ == "__main__":
    main()
