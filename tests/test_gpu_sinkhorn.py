"""GPU Sinkhorn: the HIP histogram path against the CPU path (reference
tolerances), and the point-cloud solver against the FP64 oracle."""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _make_positive(shape, seed):
    generator = torch.Generator().manual_seed(seed)
    return torch.rand(shape, generator=generator) + 0.1


@pytest.mark.parametrize("log_space", [False, True])
@pytest.mark.parametrize("n", [1, 3, 20])
def test_hist_gpu_matches_cpu(tm, log_space, n):
    """Reference test_optimal_transport.py:155-175 (rtol/atol 1e-4)."""
    solver = tm.SinkhornSolver(epsilon=1.5, max_iter=1000, log_space=log_space)
    s = _make_positive((n, 6), 503)
    t = _make_positive((n, 6), 607)
    cd, cp = solver(s, t), solver.plan(s, t)
    cf, cg = solver.potentials(s, t)
    gd, gp = solver(s.cuda(), t.cuda()), solver.plan(s.cuda(), t.cuda())
    gf, gg = solver.potentials(s.cuda(), t.cuda())
    assert torch.allclose(gd.cpu(), cd, rtol=1e-4, atol=1e-5)
    assert torch.allclose(gp.cpu(), cp, rtol=1e-4, atol=1e-5)
    assert torch.allclose(gf.cpu(), cf, rtol=1e-4, atol=1e-4)
    assert torch.allclose(gg.cpu(), cg, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("log_space", [False, True])
def test_hist_short_run_and_fp64(tm, log_space):
    solver = tm.SinkhornSolver(epsilon=1.5, max_iter=60, log_space=log_space)
    s = _make_positive((2, 6), 1409)
    t = _make_positive((2, 6), 1423)
    assert torch.allclose(solver(s.cuda(), t.cuda()).cpu(), solver(s, t), rtol=1e-4, atol=1e-5)
    solver = tm.SinkhornSolver(epsilon=1.5, max_iter=2000, log_space=log_space)
    ref = solver(s.double(), t.double())
    got = solver(s.cuda().double(), t.cuda().double())
    assert torch.allclose(got.cpu(), ref, rtol=1e-8, atol=1e-8)


def test_hist_graph_capture_failure_falls_back(tm, monkeypatch):
    solver = tm.SinkhornSolver(epsilon=1.5, max_iter=137, log_space=True)
    s = _make_positive((3, 8), 5).cuda()
    t = _make_positive((3, 8), 6).cuda()
    a = solver.potentials(s, t)
    monkeypatch.setenv("TORCHMORPH_AMD_FORCE_GRAPH_FAILURE", "1")
    b = solver.potentials(s, t)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


@pytest.mark.parametrize("log_space", [False, True])
def test_hist_backward_matches_cpu(tm, log_space):
    solver = tm.SinkhornSolver(epsilon=1.5, max_iter=1000, log_space=log_space)
    s = _make_positive((3, 6), 1213)
    t = _make_positive((3, 6), 1217)
    cs, ct = s.clone().requires_grad_(True), t.clone().requires_grad_(True)
    solver(cs, ct).sum().backward()
    gs, gt = s.cuda().requires_grad_(True), t.cuda().requires_grad_(True)
    solver(gs, gt).sum().backward()
    assert torch.allclose(gs.grad.cpu(), cs.grad, rtol=1e-4, atol=1e-5)
    assert torch.allclose(gt.grad.cpu(), ct.grad, rtol=1e-4, atol=1e-5)


def test_hist_small_epsilon_log_space(tm):
    solver = tm.SinkhornSolver(epsilon=0.1, max_iter=5000, log_space=True)
    cost = tm.build_cost_matrix((8, 8), device="cuda")
    s = _make_positive((1, 64), 31).cuda()
    t = _make_positive((1, 64), 37).cuda()
    f, g = solver.potentials(s, t, cost)
    assert torch.isfinite(f).all() and torch.isfinite(g).all()
    plan = solver.plan(s, t, cost)[0]
    a, b, _ = solver.data_preprocess(s, t, cost)
    assert torch.allclose(plan.sum(dim=0), b[0], atol=1e-5)
    assert torch.allclose(plan.sum(dim=1), a[0], atol=5e-3)


def _asym_cost(d, seed, dtype=torch.float64):
    """A non-symmetric cost: a grid distance plus a random skew, so a
    transposed operand anywhere changes the answer."""
    g = torch.Generator().manual_seed(seed)
    base = torch.cdist(torch.arange(float(d))[:, None], torch.arange(float(d))[:, None])
    return (base + 2.0 * torch.rand((d, d), generator=g, dtype=torch.float64).float()).to(dtype)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("log_space", [False, True])
def test_hist_asymmetric_cost_vs_fp64_oracle(tm, oracle_mod, dtype, log_space):
    """Non-symmetric cost against the FP64 oracle iterations directly:
    potentials, the fused value <P, C> and the plan."""
    n, d, eps, iters = 5, 13, 0.7, 300
    s = _make_positive((n, d), 71).to(dtype)
    t = _make_positive((n, d), 73).to(dtype)
    cost = _asym_cost(d, 5, dtype)
    assert not torch.equal(cost, cost.mT)
    solver = tm.SinkhornSolver(epsilon=eps, max_iter=iters, log_space=log_space)
    a, b, c = solver.data_preprocess(s, t, cost)
    form = oracle_mod.sinkhorn_log_hist if log_space else oracle_mod.sinkhorn_scaling_hist
    olu, olv = form(a.numpy(), b.numpy(), c.numpy(), eps, iters)
    oplan = np.exp(olu[:, :, None] - c.double().numpy()[None] / eps + olv[:, None, :])
    oval = (oplan * c.double().numpy()[None]).sum((1, 2))
    of = eps * olu - (eps * olu).mean(1, keepdims=True)
    og = eps * olv - (eps * olv).mean(1, keepdims=True)
    tol = 2e-4 if dtype == torch.float32 else 1e-9
    val = solver(s.cuda(), t.cuda(), cost.cuda())
    f, g = solver.potentials(s.cuda(), t.cuda(), cost.cuda())
    plan = solver.plan(s.cuda(), t.cuda(), cost.cuda())
    assert val.dtype == dtype and plan.shape == (n, d, d)
    np.testing.assert_allclose(val.double().cpu().numpy(), oval, rtol=tol, atol=tol)
    np.testing.assert_allclose(f.double().cpu().numpy(), of, rtol=tol, atol=tol)
    np.testing.assert_allclose(g.double().cpu().numpy(), og, rtol=tol, atol=tol)
    np.testing.assert_allclose(plan.double().cpu().numpy(), oplan, rtol=tol, atol=tol)
    # the transposed problem is a different one: the answer must move
    val_t = solver(s.cuda(), t.cuda(), cost.mT.contiguous().cuda())
    assert np.abs(val_t.double().cpu().numpy() - oval).max() > 1e-2


@pytest.mark.parametrize("log_space", [False, True])
def test_hist_grid_d1024_vs_fp64_oracle(tm, oracle_mod, log_space):
    """A 32 x 32 grid cost (d = 1024, the reference benchmark's shape) with
    several batch rows: float64 against the FP64 oracle, float32 against the
    CPU path at the reference's rtol / atol 1e-4."""
    cost = tm.build_cost_matrix((32, 32))
    eps, iters = 4.0, 120
    s = _make_positive((3, 1024), 91)
    t = _make_positive((3, 1024), 97)
    solver = tm.SinkhornSolver(epsilon=eps, max_iter=iters, log_space=log_space)
    a, b, c = solver.data_preprocess(s.double(), t.double(), cost.double())
    form = oracle_mod.sinkhorn_log_hist if log_space else oracle_mod.sinkhorn_scaling_hist
    olu, _ = form(a.numpy(), b.numpy(), c.numpy(), eps, iters)
    of = eps * olu - (eps * olu).mean(1, keepdims=True)
    f64, _ = solver.potentials(s.double().cuda(), t.double().cuda(), cost.double().cuda())
    np.testing.assert_allclose(f64.cpu().numpy(), of, rtol=1e-8, atol=1e-8)
    g32 = solver(s.cuda(), t.cuda(), cost.cuda()).cpu()
    c32 = solver(s, t, cost)
    assert torch.allclose(g32, c32, rtol=1e-4, atol=1e-4)


@pytest.mark.parametrize("log_space", [False, True])
def test_hist_fp64_threshold_stops_like_cpu(tm, log_space):
    """float64 CUDA honours threshold (the reference runs its torch
    iterations there): same stopping iteration and values as the CPU path."""
    from torchmorph_amd.optimal_transport import _hist_solve

    solver = tm.SinkhornSolver(epsilon=1.5, max_iter=5000, threshold=1e-9, log_space=log_space)
    s = _make_positive((4, 10), 211).double()
    t = _make_positive((4, 10), 223).double()
    a, b, c = solver.data_preprocess(s, t)
    cpu = _hist_solve(a, b, c, 1.5, 5000, 1e-9, log_space, ("potentials", "value"))
    gpu = _hist_solve(a.cuda(), b.cuda(), c.cuda(), 1.5, 5000, 1e-9, log_space,
                      ("potentials", "value"))
    assert 1 < cpu.iterations < 5000 and cpu.iterations % 10 == 1
    assert gpu.iterations == cpu.iterations
    assert torch.allclose(gpu.value.cpu(), cpu.value, rtol=1e-9, atol=1e-12)
    assert torch.allclose(gpu.f.cpu(), cpu.f, rtol=1e-9, atol=1e-12)
    assert torch.allclose(solver(s.cuda(), t.cuda()).cpu(), solver(s, t), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("dtype", [torch.float16, torch.bfloat16])
def test_hist_half_precision_matches_cpu(tm, dtype):
    """float16 / bfloat16 CUDA inputs run (solved in float32, cast back)
    and agree with the CPU path on the same inputs at half precision."""
    solver = tm.SinkhornSolver(epsilon=1.5, max_iter=400, threshold=1e-7, log_space=True)
    s = _make_positive((3, 8), 301).to(dtype)
    t = _make_positive((3, 8), 307).to(dtype)
    got = solver(s.cuda(), t.cuda())
    f, g = solver.potentials(s.cuda(), t.cuda())
    assert got.dtype == dtype and f.dtype == dtype
    ref = solver(s.float(), t.float())
    rf, rg = solver.potentials(s.float(), t.float())
    tol = 1e-2 if dtype == torch.float16 else 4e-2
    assert torch.allclose(got.float().cpu(), ref, rtol=tol, atol=tol)
    assert torch.allclose(f.float().cpu(), rf, rtol=tol, atol=tol)
    assert torch.allclose(g.float().cpu(), rg, rtol=tol, atol=tol)


def test_hist_float32_ignores_threshold(tm):
    """float32 CUDA runs exactly max_iter iterations, as the reference's
    fused kernels do, whatever the threshold."""
    from torchmorph_amd.optimal_transport import _hist_solve

    s = _make_positive((2, 6), 401).cuda()
    t = _make_positive((2, 6), 409).cuda()
    sol = _hist_solve(s / s.sum(1, keepdim=True), t / t.sum(1, keepdim=True),
                      tm.build_cost_matrix((6,), device="cuda"), 1.5, 137, 1e-3, True)
    assert sol.iterations == 137


def test_hist_misaligned_views_and_value_consistency(tm):
    """Rows that do not start on a 16-byte boundary take the scalar loads;
    the fused value equals <plan, C> from the plan kernel."""
    solver = tm.SinkhornSolver(epsilon=0.9, max_iter=150, log_space=True)
    big = _make_positive((1, 3 * 16 + 1), 501).cuda()
    s = big[0, 1:].reshape(3, 16)
    t = _make_positive((3, 16), 503).cuda()
    cost = _asym_cost(16, 9, torch.float32).cuda()
    assert s.data_ptr() % 16 != 0
    val = solver(s, t, cost)
    plan = solver.plan(s, t, cost)
    assert torch.allclose(val, (plan * cost).sum((1, 2)), rtol=1e-5, atol=1e-6)
    ref = solver(s.cpu(), t.cpu(), cost.cpu())
    assert torch.allclose(val.cpu(), ref, rtol=1e-4, atol=1e-5)


# ------------------------------------------------------------- point clouds

def _gmm_pair(rng, n, m, d=3):
    def cloud(k):
        means = rng.normal(0, 2.0, size=(4, d))
        sig = rng.uniform(0.3, 1.0, size=4)
        w = rng.dirichlet(np.ones(4))
        comp = rng.choice(4, size=k, p=w)
        return means[comp] + rng.normal(size=(k, d)) * sig[comp, None]
    return cloud(n), cloud(m)


def _marginal_err64(x, y, a, b, f, g, eps):
    C = ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)
    P = np.exp((f[:, None] + g[None, :] - C) / eps) * a[:, None] * b[None, :]
    return np.abs(P.sum(1) - a).sum(), np.abs(P.sum(0) - b).sum()


@pytest.mark.parametrize("n,m", [(64, 64), (200, 150), (37, 301)])
def test_points_vs_fp64_oracle(tm, oracle_mod, n, m):
    rng = np.random.default_rng(n + m)
    x, y = _gmm_pair(rng, n, m)
    a = np.full(n, 1.0 / n)
    b = np.full(m, 1.0 / m)
    C = ((x[:, None] - y[None]) ** 2).sum(-1)
    eps = 0.05 * float(np.median(C))
    res = tm.sinkhorn_points(torch.tensor(x[None], dtype=torch.float32, device="cuda"),
                             torch.tensor(y[None], dtype=torch.float32, device="cuda"), eps,
                             tol=1e-5)
    f = res.f.double().cpu().numpy()
    g = res.g.double().cpu().numpy()
    it = int(res.iterations[0])
    assert float(res.err[0]) <= 1e-5
    # the returned state satisfies the tolerance when re-measured in FP64
    # (row error, up to float32 rounding of the potentials) and its columns
    # are balanced by the last g-step
    row, col = _marginal_err64(x, y, a, b, f, g, eps)
    assert row <= 3e-5 and col <= 3e-5, (row, col, it)
    # potentials agree with a tightly converged FP64 solve (gauge: f+c, g-c)
    of, og, oit, _ = oracle_mod.sinkhorn_points(x, y, a, b, eps, tol=1e-12)
    c = np.mean(f - of)
    assert np.abs(f - of - c).max() <= 2e-3 * eps, np.abs(f - of - c).max() / eps
    assert np.abs(g - og + c).max() <= 2e-3 * eps
    # stopping iteration close to the FP64 solver's at the same tolerance
    # (the float32 error estimate carries ~1e-6 noise near the threshold)
    _, _, oit5, _ = oracle_mod.sinkhorn_points(x, y, a, b, eps, tol=1e-5)
    assert abs(it - oit5) <= max(2, 0.15 * oit5), (it, oit5)


def test_points_ragged_batch_equals_single_pairs(tm):
    rng = np.random.default_rng(0)
    sizes = [(64, 80), (130, 33), (257, 256), (40, 400)]
    xs, ys, eps = [], [], []
    for n, m in sizes:
        x, y = _gmm_pair(rng, n, m)
        xs.append(torch.tensor(x, dtype=torch.float32, device="cuda"))
        ys.append(torch.tensor(y, dtype=torch.float32, device="cuda"))
        eps.append(0.05 * float(torch.cdist(xs[-1], ys[-1]).square().median()))
    res = tm.sinkhorn_points(xs, ys, torch.tensor(eps), tol=1e-5)
    for p in range(len(sizes)):
        one = tm.sinkhorn_points([xs[p]], [ys[p]], eps[p], tol=1e-5)
        assert torch.equal(res.split_f()[p], one.f)
        assert torch.equal(res.split_g()[p], one.g)
        assert int(res.iterations[p]) == int(one.iterations[0])


def test_points_graph_and_plain_launches_agree(tm, monkeypatch):
    rng = np.random.default_rng(1)
    x, y = _gmm_pair(rng, 100, 90)
    X = torch.tensor(x[None], dtype=torch.float32, device="cuda")
    Y = torch.tensor(y[None], dtype=torch.float32, device="cuda")
    a = tm.sinkhorn_points(X, Y, 0.5, tol=1e-6)
    monkeypatch.setenv("TORCHMORPH_AMD_NO_GRAPH", "1")
    b = tm.sinkhorn_points(X, Y, 0.5, tol=1e-6)
    assert torch.equal(a.f, b.f) and torch.equal(a.g, b.g)
    assert torch.equal(a.iterations, b.iterations)


def test_points_compacted_schedule_equals_single_pairs(tm):
    """A batch whose work-list is compacted several times (64-, 16- and
    4-row tiles as pairs converge, replays pipelined one chunk ahead) gives
    every pair the bits and iteration count of solving it alone."""
    rng = np.random.default_rng(8)
    B = 24
    sizes = [(int(rng.integers(100, 1025)), int(rng.integers(100, 1025))) for _ in range(B)]
    fracs = rng.choice([0.01, 0.03, 0.1], size=B)
    xs, ys, eps = [], [], []
    for (n, m), fr in zip(sizes, fracs):
        x, y = _gmm_pair(rng, n, m)
        xs.append(torch.tensor(x, dtype=torch.float32, device="cuda"))
        ys.append(torch.tensor(y, dtype=torch.float32, device="cuda"))
        eps.append(float(fr) * float(torch.cdist(xs[-1], ys[-1]).square().median()))
    res = tm.sinkhorn_points(xs, ys, torch.tensor(eps), tol=1e-5)
    assert int(res.iterations.max()) > 4 * int(res.iterations.min())
    for p in range(B):
        one = tm.sinkhorn_points([xs[p]], [ys[p]], eps[p], tol=1e-5)
        assert torch.equal(res.split_f()[p], one.f), p
        assert torch.equal(res.split_g()[p], one.g), p
        assert int(res.iterations[p]) == int(one.iterations[0]), p


def test_points_warm_start_from_solution_stops_immediately(tm):
    rng = np.random.default_rng(2)
    x, y = _gmm_pair(rng, 120, 110)
    X = torch.tensor(x[None], dtype=torch.float32, device="cuda")
    Y = torch.tensor(y[None], dtype=torch.float32, device="cuda")
    cold = tm.sinkhorn_points(X, Y, 0.3, tol=1e-5)
    warm = tm.sinkhorn_points(X, Y, 0.3, tol=1e-5, f_init=cold.f)
    assert int(warm.iterations[0]) <= 1 < int(cold.iterations[0])


def test_points_max_iter_stops_unconverged(tm):
    rng = np.random.default_rng(3)
    x, y = _gmm_pair(rng, 80, 70)
    X = torch.tensor(x[None], dtype=torch.float32, device="cuda")
    Y = torch.tensor(y[None], dtype=torch.float32, device="cuda")
    res = tm.sinkhorn_points(X, Y, 0.01, tol=1e-9, max_iter=7)
    assert int(res.iterations[0]) == 7
    assert float(res.err[0]) > 1e-9


def test_point_cloud_value_gradients_vs_fp64_fd(tm, oracle_mod):
    rng = np.random.default_rng(4)
    x, y = _gmm_pair(rng, 30, 25)
    a = np.full(30, 1 / 30)
    b = np.full(25, 1 / 25)
    eps = 1.0
    X = torch.tensor(x, dtype=torch.float32, device="cuda", requires_grad=True)
    Y = torch.tensor(y, dtype=torch.float32, device="cuda", requires_grad=True)
    A = torch.tensor(a, dtype=torch.float32, device="cuda", requires_grad=True)
    B = torch.tensor(b, dtype=torch.float32, device="cuda", requires_grad=True)
    val = tm.SinkhornPointCloud.apply(X, Y, A, B, eps, 1e-7, 20000)
    val.backward()

    def value64(xx, yy, aa, bb):
        f, g, _, _ = oracle_mod.sinkhorn_points(xx, yy, aa, bb, eps, tol=1e-13)
        return (f * aa).sum() + (g * bb).sum()

    v0 = value64(x, y, a, b)
    assert abs(float(val) - v0) <= 1e-4 * max(1.0, abs(v0))
    h = 1e-5
    d = rng.normal(size=x.shape)
    fd = (value64(x + h * d, y, a, b) - value64(x - h * d, y, a, b)) / (2 * h)
    assert abs(float((X.grad.cpu().double().numpy() * d).sum()) - fd) <= 2e-3 * max(1, abs(fd))
    d = rng.normal(size=y.shape)
    fd = (value64(x, y + h * d, a, b) - value64(x, y - h * d, a, b)) / (2 * h)
    assert abs(float((Y.grad.cpu().double().numpy() * d).sum()) - fd) <= 2e-3 * max(1, abs(fd))
    # mass-preserving direction on a
    d = rng.normal(size=a.shape)
    d -= d.mean()
    fd = (value64(x, y, a + h * d, b) - value64(x, y, a - h * d, b)) / (2 * h)
    assert abs(float((A.grad.cpu().double().numpy() * d).sum()) - fd) <= 2e-3 * max(1, abs(fd))
