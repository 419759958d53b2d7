"""Point-cloud Sinkhorn at the benchmark workloads' sizes, under parity.

* Batch of C4 pairs: 24 pairs of 1024 x 1024 points in R^3 from 4-component
  Gaussian mixtures, squared-Euclidean cost, eps = 0.05 x median cost, uniform
  marginals, stopped at 8e-6 in float32.
* Learned warm start: the shipped network (torchmorph_amd/weights/) on
  held-out pairs with n, m uniform in 256..4096 (one pair exactly 4096 x 4096,
  the column-staging limit), eps in {0.1, 0.01} x median; warm and cold to
  1e-5.

Checks: row and column L1 marginal errors of every returned plan recomputed
in float64 on the GPU from the returned (f, g) (bound 1e-5 for C4 and 3e-5 at
tol 1e-5, float32 potentials); potentials within 2e-3 x eps of the float64
oracle (up to the gauge constant) on small pairs of the same distribution at
both eps; and the warm start's measured gain in iterations and solves/s,
printed and asserted against the gain the shipped network reaches (the
north-star target of 3x / 2x is not met; see the design notes).
"""
import os
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WEIGHTS = os.path.join(ROOT, "torchmorph_amd", "weights", "warmstart_gmm3d.pt")


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _fp64_marginal_errors(x, y, f, g, eps, a=None, b=None):
    """(row, column) L1 marginal errors of the plan of (f, g), in float64."""
    x, y = x.double(), y.double()
    n, m = len(x), len(y)
    a = torch.full((n,), 1.0 / n, dtype=torch.float64, device=x.device) if a is None else a
    b = torch.full((m,), 1.0 / m, dtype=torch.float64, device=x.device) if b is None else b
    C = ((x[:, None] - y[None]) ** 2).sum(-1)
    P = torch.exp((f.double()[:, None] + g.double()[None] - C) / float(eps)) * a[:, None] * b[None]
    return float((P.sum(1) - a).abs().sum()), float((P.sum(0) - b).abs().sum())


def _median_eps(x, y, frac):
    return frac * float(((x[:, None] - y[None]) ** 2).sum(-1).flatten().median())


def test_c4_batch_fp64_marginals(tm):
    from torchmorph_amd.warmstart import gmm_clouds

    rng = np.random.default_rng(11)
    B, N = 24, 1024
    xs = torch.from_numpy(np.stack(gmm_clouds(rng, [N] * B))).cuda()
    ys = torch.from_numpy(np.stack(gmm_clouds(rng, [N] * B))).cuda()
    eps = torch.tensor([_median_eps(xs[p], ys[p], 0.05) for p in range(B)])
    res = tm.sinkhorn_points(xs, ys, eps, tol=8e-6)
    assert bool((res.err <= 8e-6).all())
    f = res.f.view(B, N)
    g = res.g.view(B, N)
    worst = 0.0
    for p in range(B):
        row, col = _fp64_marginal_errors(xs[p], ys[p], f[p], g[p], eps[p])
        worst = max(worst, row, col)
    print(f"C4 batch: {B} pairs {N}x{N}, iterations mean {float(res.iterations.float().mean()):.1f}"
          f" max {int(res.iterations.max())}, fp64 marginal error max {worst:.3g}")
    assert worst <= 1e-5


@pytest.mark.parametrize("frac", [0.1, 0.01])
def test_potentials_vs_fp64_oracle_at_config_eps(tm, oracle_mod, frac):
    """Small pairs of the warm-start distribution at eps = frac x median."""
    from torchmorph_amd.warmstart import gmm_clouds

    rng = np.random.default_rng(int(frac * 1000) + 5)
    x, y = gmm_clouds(rng, [257, 300])
    eps = _median_eps(torch.from_numpy(x).double(), torch.from_numpy(y).double(), frac)
    res = tm.sinkhorn_points(torch.from_numpy(x[None]).cuda(), torch.from_numpy(y[None]).cuda(),
                             eps, tol=1e-5)
    of, og, _, _ = oracle_mod.sinkhorn_points(x.astype(np.float64), y.astype(np.float64),
                                              np.full(257, 1 / 257), np.full(300, 1 / 300), eps,
                                              tol=1e-11)
    f = res.f.double().cpu().numpy()
    g = res.g.double().cpu().numpy()
    c = np.mean(f - of)
    assert np.abs(f - of - c).max() <= 2e-3 * eps
    assert np.abs(g - og + c).max() <= 2e-3 * eps


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason="no trained weights shipped")
def test_warm_start_gain_and_fp64_marginals():
    from torchmorph_amd.train_warmstart import make_pairs
    from torchmorph_amd.warmstart import WarmStartNet, solve

    net = WarmStartNet.load(WEIGHTS)
    rng = np.random.default_rng(777)  # held out: training used seed 0, evaluation 12345
    xs, ys, eps = make_pairs(rng, 16)
    # one pair at the 4096-point limit on both sides
    from torchmorph_amd.warmstart import gmm_clouds

    big = gmm_clouds(rng, [4096, 4096])
    xs[0] = torch.from_numpy(big[0]).cuda()
    ys[0] = torch.from_numpy(big[1]).cuda()
    eps[0] = _median_eps(xs[0][::4], ys[0][::4], 0.01)
    e = torch.tensor(eps)
    solve(None, xs[:2], ys[:2], e[:2])  # first calls: code objects, allocator
    solve(net, xs[:2], ys[:2], e[:2])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    cold = solve(None, xs, ys, e, tol=1e-5)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    warm = solve(net, xs, ys, e, tol=1e-5)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    for res in (cold, warm):
        assert bool((res.err <= 1e-5).all())
        worst = 0.0
        for p, (f, g) in enumerate(zip(res.split_f(), res.split_g())):
            row, col = _fp64_marginal_errors(xs[p], ys[p], f, g, eps[p])
            worst = max(worst, row, col)
        assert worst <= 3e-5, worst
    iter_gain = float(cold.iterations.sum()) / float(warm.iterations.sum())
    speed_gain = (t1 - t0) / (t2 - t1)
    print(f"warm start on 16 held-out pairs (256..4096 points, one 4096x4096): iterations "
          f"{int(cold.iterations.sum())} cold / {int(warm.iterations.sum())} warm = "
          f"{iter_gain:.3f}x, solves/s {16 / (t1 - t0):.1f} cold / {16 / (t2 - t1):.1f} warm = "
          f"{speed_gain:.3f}x (prediction included)")
    assert iter_gain >= 1.1
    assert speed_gain >= 1.0
