"""Trains the Sinkhorn warm-start network on the GPU and evaluates it.

Training data: pairs of 3-D point clouds from 4-component Gaussian mixtures,
sizes n, m uniform in [256, 4096], uniform weights, eps = k * median cost with
k in {0.1, 0.01} (the median estimated from 16384 seeded random (i, j)
samples per pair), solved cold by the project's own Sinkhorn to an L1
marginal error of 1e-5; the converged potentials f are the targets.  The
network is trained by its own BF16 MFMA forward/backward kernels with Adam.
Evaluation: held-out pairs from the same distribution (different seed),
iterations and wall time to tolerance, cold start vs warm start (prediction
time included).

    python -m torchmorph_amd.train_warmstart --pairs 100000 --out weights.pt
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np
import torch

from .optimal_transport import sinkhorn_points
from .warmstart import WarmStartNet, gmm_clouds, solve


def make_pairs(rng: np.random.Generator, count: int, lo=256, hi=4096, fracs=(0.1, 0.01),
               device="cuda"):
    sizes = rng.integers(lo, hi + 1, size=(count, 2))
    xs = [torch.from_numpy(c).to(device) for c in gmm_clouds(rng, sizes[:, 0])]
    ys = [torch.from_numpy(c).to(device) for c in gmm_clouds(rng, sizes[:, 1])]
    eps = []
    g = torch.Generator(device=device).manual_seed(int(rng.integers(1 << 31)))
    for k, (x, y) in enumerate(zip(xs, ys)):
        i = torch.randint(0, len(x), (16384,), generator=g, device=device)
        j = torch.randint(0, len(y), (16384,), generator=g, device=device)
        med = ((x[i] - y[j]) ** 2).sum(1).median()
        eps.append(float(fracs[k % len(fracs)]) * float(med))
    return xs, ys, eps


def generate(count, seed, chunk=1024, tol=1e-5, max_iter=5000, log=print):
    rng = np.random.default_rng(seed)
    data = []
    t0 = time.perf_counter()
    for start in range(0, count, chunk):
        k = min(chunk, count - start)
        xs, ys, eps = make_pairs(rng, k)
        res = sinkhorn_points(xs, ys, torch.tensor(eps), tol=tol, max_iter=max_iter)
        fs = res.split_f()
        it = res.iterations.cpu()
        err = res.err.cpu()
        for p in range(k):
            if float(err[p]) <= tol:
                data.append((xs[p], ys[p], eps[p], fs[p].clone(), int(it[p])))
        log(json.dumps({"generate": start + k, "of": count, "kept": len(data),
                        "mean_iters": float(it.float().mean()),
                        "s": round(time.perf_counter() - t0, 1)}))
    return data


def train(net: WarmStartNet, data, epochs=8, batch=64, lr=1e-3, lr_min=1e-5, seed=0,
          loss_alpha=1.0, log=print):
    rng = np.random.default_rng(seed)
    steps_per_epoch = math.ceil(len(data) / batch)
    total = epochs * steps_per_epoch
    step = 0
    t0 = time.perf_counter()
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(data))
        losses = []
        for b in range(steps_per_epoch):
            idx = order[b * batch:(b + 1) * batch]
            xs = [data[i][0] for i in idx]
            ys = [data[i][1] for i in idx]
            eps = [data[i][2] for i in idx]
            f = torch.cat([data[i][3] for i in idx])
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / total))
            losses.append(net.train_step(xs, ys, eps, f, lr=cur, loss_alpha=loss_alpha))
            step += 1
        ep_loss = float(torch.stack(losses).mean())
        hist.append(ep_loss)
        log(json.dumps({"epoch": ep + 1, "loss": ep_loss, "s": round(time.perf_counter() - t0, 1)}))
    return hist


def evaluate(net: WarmStartNet, count=1024, seed=12345, tol=1e-5, max_iter=20000, chunk=256,
             log=print):
    """Cold vs warm on held-out pairs: iterations and wall time to `tol`."""
    rng = np.random.default_rng(seed)
    out = {"pairs": 0, "cold_iters": 0, "warm_iters": 0, "cold_s": 0.0, "warm_s": 0.0,
           "cold_err_max": 0.0, "warm_err_max": 0.0, "warm_fallbacks": 0}
    for start in range(0, count, chunk):
        k = min(chunk, count - start)
        xs, ys, eps = make_pairs(rng, k)
        e = torch.tensor(eps)
        # warm up both paths once (allocator, code objects)
        if start == 0:
            solve(None, xs[:2], ys[:2], e[:2], tol=tol)
            solve(net, xs[:2], ys[:2], e[:2], tol=tol)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        cold = solve(None, xs, ys, e, tol=tol, max_iter=max_iter)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        warm = solve(net, xs, ys, e, tol=tol, max_iter=max_iter)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        out["pairs"] += k
        out["cold_iters"] += int(cold.iterations.sum())
        out["warm_iters"] += int(warm.iterations.sum())
        out["cold_s"] += t1 - t0
        out["warm_s"] += t2 - t1
        out["cold_err_max"] = max(out["cold_err_max"], float(cold.err.max()))
        out["warm_err_max"] = max(out["warm_err_max"], float(warm.err.max()))
        log(json.dumps({"eval": out["pairs"], "cold_iters": out["cold_iters"],
                        "warm_iters": out["warm_iters"]}))
    out["iter_ratio"] = out["cold_iters"] / max(1, out["warm_iters"])
    out["cold_solves_per_s"] = out["pairs"] / out["cold_s"]
    out["warm_solves_per_s"] = out["pairs"] / out["warm_s"]
    out["speedup"] = out["warm_solves_per_s"] / out["cold_solves_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=100000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--loss-alpha", type=float, default=1.0)
    ap.add_argument("--eval-pairs", type=int, default=1024)
    ap.add_argument("--out", default="warmstart.pt")
    ap.add_argument("--summary", default=None)
    args = ap.parse_args()
    log = lambda s: print(s, flush=True)
    data = generate(args.pairs, args.seed, log=log)
    net = WarmStartNet.create(seed=args.seed)
    hist = train(net, data, epochs=args.epochs, batch=args.batch, lr=args.lr,
                 loss_alpha=args.loss_alpha, log=log)
    net.save(args.out)
    ev = evaluate(net, count=args.eval_pairs, log=log)
    summary = {"train_pairs": len(data), "epochs": args.epochs, "loss_history": hist,
               "eval": ev}
    log(json.dumps(summary))
    if args.summary:
        with open(args.summary, "w") as f:
            json.dump(summary, f, indent=1)


if __name__ == "__main__":
    main()
