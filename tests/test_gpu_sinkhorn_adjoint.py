"""The HIP side of the Sinkhorn adjoint: plan products formed on the fly
(tm_sinkhorn_points_planmat) against the float64 materialised plan, and
sinkhorn_potentials' backward (float32 solve, HIP plan products) against
the float64 adjoint on the oracle's solution and float64 central
differences."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _pair(seed, n, m, d=3):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, d))
    y = rng.normal(size=(m, d)) * 0.8 + 0.4
    a = rng.random(n) + 0.5
    b = rng.random(m) + 0.5
    return x, y, a / a.sum(), b / b.sum()


def _plan64(x, y, a, b, f, g, eps):
    C = ((x[:, None] - y[None]) ** 2).sum(-1)
    return a[:, None] * b[None] * np.exp((f[:, None] + g[None] - C) / eps)


@pytest.mark.parametrize("n,m,k,d", [(300, 257, 8, 3), (1000, 64, 3, 3), (77, 4096, 5, 2)])
def test_hip_plan_products_match_fp64(tm, n, m, k, d):
    from torchmorph_amd.optimal_transport import HipPlanOps

    x, y, a, b = _pair(n + m, n, m, d)
    eps = 0.5
    X, Y = torch.tensor(x, dtype=torch.float32).cuda(), torch.tensor(y, dtype=torch.float32).cuda()
    A, B = torch.tensor(a, dtype=torch.float32).cuda(), torch.tensor(b, dtype=torch.float32).cuda()
    res = tm.sinkhorn_points(X[None], Y[None], eps, A[None], B[None], tol=1e-6)
    f, g = res.f, res.g
    ops = HipPlanOps(X, Y, A, B, f, g, eps)
    P = _plan64(X.double().cpu().numpy(), Y.double().cpu().numpy(), A.double().cpu().numpy(),
                B.double().cpu().numpy(), f.double().cpu().numpy(), g.double().cpu().numpy(), eps)
    rng = np.random.default_rng(1)
    V = rng.normal(size=(m, k))
    U = rng.normal(size=(n, k))
    got_r = ops.rows(torch.tensor(V, dtype=torch.float32).cuda()).double().cpu().numpy()
    got_c = ops.cols(torch.tensor(U, dtype=torch.float32).cuda()).double().cpu().numpy()
    want_r = P @ V.astype(np.float32).astype(np.float64)
    want_c = P.T @ U.astype(np.float32).astype(np.float64)
    scale_r = np.abs(P) @ np.abs(V)
    scale_c = np.abs(P.T) @ np.abs(U)
    assert np.all(np.abs(got_r - want_r) <= 1e-4 * scale_r + 1e-12)
    assert np.all(np.abs(got_c - want_c) <= 1e-4 * scale_c + 1e-12)


@pytest.mark.parametrize("mode,steps", [("implicit", 300), ("unrolled", 2000)])
def test_potentials_backward_matches_fp64(tm, oracle_mod, mode, steps):
    from tests.test_ot_adjoint_cpu import TorchPlanOps
    from torchmorph_amd.optimal_transport import potentials_backward

    x, y, a, b = _pair(42, 60, 45)
    eps = 0.7
    rng = np.random.default_rng(3)
    wf = rng.normal(size=60)
    wg = rng.normal(size=45)
    dev = "cuda"
    X = torch.tensor(x, dtype=torch.float32, device=dev, requires_grad=True)
    Y = torch.tensor(y, dtype=torch.float32, device=dev, requires_grad=True)
    A = torch.tensor(a, dtype=torch.float32, device=dev, requires_grad=True)
    B = torch.tensor(b, dtype=torch.float32, device=dev, requires_grad=True)
    f, g = tm.sinkhorn_potentials(X, Y, eps, A, B, tol=1e-7, mode=mode, steps=steps)
    L = (torch.tensor(wf, dtype=torch.float32, device=dev) * f).sum() + \
        (torch.tensor(wg, dtype=torch.float32, device=dev) * g).sum()
    L.backward()
    # float64 reference: the same adjoint on the oracle's solution
    f0, g0, _, _ = oracle_mod.sinkhorn_points(x, y, a, b, eps, tol=1e-14)
    c = ((g0 * b).sum() - (f0 * a).sum()) / 2
    f64, g64 = torch.tensor(f0 + c), torch.tensor(g0 - c)
    assert np.abs(f.detach().double().cpu().numpy() - f64.numpy()).max() <= 2e-3 * eps
    x64, y64, a64, b64 = map(torch.tensor, (x, y, a, b))
    ops = TorchPlanOps(torch.tensor(_plan64(x, y, a, b, f64.numpy(), g64.numpy(), eps)))
    gx, gy, ga, gb = potentials_backward(ops, x64, y64, a64, b64, f64, g64, eps,
                                         torch.tensor(wf), torch.tensor(wg), "implicit", 300)
    for got, want in ((X.grad, gx), (Y.grad, gy), (A.grad, ga), (B.grad, gb)):
        got = got.double().cpu()
        err = float((got - want).abs().max())
        assert err <= 2e-3 * max(1.0, float(want.abs().max())), err


def test_plan_loss_gradient_vs_fp64_fd(tm, oracle_mod):
    """A loss of the plan built from the differentiable potentials: the
    direct terms come from autograd, the implicit ones from the adjoint."""
    x, y, a, b = _pair(7, 40, 35, 2)
    eps = 0.6
    rng = np.random.default_rng(5)
    Q = rng.normal(size=(40, 35))
    dev = "cuda"
    X = torch.tensor(x, dtype=torch.float32, device=dev, requires_grad=True)
    Y = torch.tensor(y, dtype=torch.float32, device=dev)
    A = torch.tensor(a, dtype=torch.float32, device=dev)
    B = torch.tensor(b, dtype=torch.float32, device=dev)
    f, g = tm.sinkhorn_potentials(X, Y, eps, A, B, tol=1e-7)
    C = ((X[:, None] - Y[None]) ** 2).sum(-1)
    P = A[:, None] * B[None] * torch.exp((f[:, None] + g[None] - C) / eps)
    (P * torch.tensor(Q, dtype=torch.float32, device=dev)).sum().backward()

    def value(xv):
        f0, g0, _, _ = oracle_mod.sinkhorn_points(xv, y, a, b, eps, tol=1e-14)
        return float((_plan64(xv, y, a, b, f0, g0, eps) * Q).sum())

    h = 1e-5
    d = rng.normal(size=x.shape)
    fd = (value(x + h * d) - value(x - h * d)) / (2 * h)
    an = float((X.grad.double().cpu().numpy() * d).sum())
    assert abs(an - fd) <= 3e-3 * max(1.0, abs(fd)), (an, fd)
