"""The implicit adjoint of the point-cloud Sinkhorn fixed point, on the CPU
in float64: the package's adjoint math (solve_adjoint, adjoint_param_grads,
potentials_backward) with the plan materialised in torch, against

* autograd through an unrolled float64 log-domain Sinkhorn (3000 iterations
  from zero, the same gauge fix), and
* float64 central differences of the oracle's converged solution,

for a loss of the potentials and a loss of the plan, in both adjoint modes
("implicit" conjugate gradients, "unrolled" reverse mode from the converged
checkpoint), plus the envelope identity of the OT value."""
import numpy as np
import pytest
import torch

from torchmorph_amd.optimal_transport import (PlanOps, adjoint_param_grads, potentials_backward,
                                              solve_adjoint)


class TorchPlanOps(PlanOps):
    def __init__(self, P):
        self.P = P

    def rows(self, V):
        return self.P @ V

    def cols(self, U):
        return self.P.T @ U


def _cost(x, y):
    return ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)


def _unrolled_potentials(x, y, a, b, eps, iters=3000):
    """Float64 log-domain Sinkhorn from zero, differentiable by autograd."""
    C = _cost(x, y)
    la, lb = a.log(), b.log()
    f = torch.zeros(x.shape[0], dtype=x.dtype)
    g = torch.zeros(y.shape[0], dtype=x.dtype)
    for _ in range(iters):
        f = -eps * torch.logsumexp((g[None, :] - C) / eps + lb[None, :], dim=1)
        g = -eps * torch.logsumexp((f[:, None] - C) / eps + la[:, None], dim=0)
    c = ((g * b).sum() - (f * a).sum()) / 2
    return f + c, g - c


def _plan(x, y, a, b, f, g, eps):
    return a[:, None] * b[None, :] * torch.exp((f[:, None] + g[None, :] - _cost(x, y)) / eps)


def _problem(seed, n=11, m=9, d=3):
    rng = np.random.default_rng(seed)
    x = torch.tensor(rng.normal(size=(n, d)))
    y = torch.tensor(rng.normal(size=(m, d)) + 0.3)
    a = torch.tensor(rng.random(n) + 0.5)
    b = torch.tensor(rng.random(m) + 0.5)
    return x, y, a / a.sum(), b / b.sum()


def _losses(rng, n, m):
    wf = torch.tensor(rng.normal(size=n))
    wg = torch.tensor(rng.normal(size=m))
    Q = torch.tensor(rng.normal(size=(n, m)))

    def potentials_loss(x, y, a, b, f, g, eps):
        return (wf * f).sum() + (wg * g).sum() + 0.3 * (f ** 2).sum()

    def plan_loss(x, y, a, b, f, g, eps):
        return (_plan(x, y, a, b, f, g, eps) * Q).sum() + ((_plan(x, y, a, b, f, g, eps)) ** 2).sum()

    return {"potentials": potentials_loss, "plan": plan_loss}


def _adjoint_grads(oracle_mod, x, y, a, b, eps, loss, mode, steps):
    f0, g0, _, _ = oracle_mod.sinkhorn_points(x.numpy(), y.numpy(), a.numpy(), b.numpy(), eps,
                                              tol=1e-14)
    f0, g0 = torch.tensor(f0), torch.tensor(g0)
    c = ((g0 * b).sum() - (f0 * a).sum()) / 2
    f, g = (f0 + c).requires_grad_(True), (g0 - c).requires_grad_(True)
    xx, yy = x.clone().requires_grad_(True), y.clone().requires_grad_(True)
    aa, bb = a.clone().requires_grad_(True), b.clone().requires_grad_(True)
    L = loss(xx, yy, aa, bb, f, g, eps)  # direct terms by autograd
    dx, dy, da, db, gf, gg = torch.autograd.grad(L, (xx, yy, aa, bb, f, g), allow_unused=True)
    zero = lambda t, like: torch.zeros_like(like) if t is None else t  # noqa: E731
    dx, dy, da, db = zero(dx, x), zero(dy, y), zero(da, a), zero(db, b)
    ops = TorchPlanOps(_plan(x, y, a, b, f.detach(), g.detach(), eps))
    gx, gy, ga, gb = potentials_backward(ops, x, y, a, b, f.detach(), g.detach(), eps, gf, gg,
                                         mode, steps)
    return (gx + dx, gy + dy, (ga + da) - (ga + da).mean(), (gb + db) - (gb + db).mean())


def _unrolled_grads(x, y, a, b, eps, loss):
    xx, yy = x.clone().requires_grad_(True), y.clone().requires_grad_(True)
    aa, bb = a.clone().requires_grad_(True), b.clone().requires_grad_(True)
    f, g = _unrolled_potentials(xx, yy, aa, bb, eps)
    L = loss(xx, yy, aa, bb, f, g, eps)
    gx, gy, ga, gb = torch.autograd.grad(L, (xx, yy, aa, bb))
    return gx, gy, ga - ga.mean(), gb - gb.mean()


@pytest.mark.parametrize("which", ["potentials", "plan"])
@pytest.mark.parametrize("mode,steps", [("implicit", 200), ("unrolled", 3000)])
def test_adjoint_matches_autograd_through_unrolled_fp64(oracle_mod, which, mode, steps):
    x, y, a, b = _problem(3)
    eps = 0.8
    loss = _losses(np.random.default_rng(7), len(x), len(y))[which]
    got = _adjoint_grads(oracle_mod, x, y, a, b, eps, loss, mode, steps)
    want = _unrolled_grads(x, y, a, b, eps, loss)
    for name, g_, w_ in zip(("x", "y", "a", "b"), got, want):
        err = float((g_ - w_).abs().max())
        scale = float(w_.abs().max())
        assert err <= 1e-7 * max(1.0, scale), (name, err, scale)


@pytest.mark.parametrize("which", ["potentials", "plan"])
def test_adjoint_matches_fp64_finite_differences(oracle_mod, which):
    x, y, a, b = _problem(5, n=8, m=10, d=2)
    eps = 0.5
    rng = np.random.default_rng(11)
    loss = _losses(rng, len(x), len(y))[which]

    def value(xv, yv, av, bv):
        f0, g0, _, _ = oracle_mod.sinkhorn_points(xv.numpy(), yv.numpy(), av.numpy(),
                                                  bv.numpy(), eps, tol=1e-14)
        f0, g0 = torch.tensor(f0), torch.tensor(g0)
        c = ((g0 * bv).sum() - (f0 * av).sum()) / 2
        return float(loss(xv, yv, av, bv, f0 + c, g0 - c, eps))

    gx, gy, ga, gb = _adjoint_grads(oracle_mod, x, y, a, b, eps, loss, "implicit", 200)
    h = 1e-6
    for grad, which_arg in ((gx, 0), (gy, 1), (ga, 2), (gb, 3)):
        args = [x, y, a, b]
        d = torch.tensor(rng.normal(size=args[which_arg].shape))
        if which_arg >= 2:
            d = d - d.mean()  # mass-preserving direction
        plus = [t.clone() for t in args]
        minus = [t.clone() for t in args]
        plus[which_arg] = plus[which_arg] + h * d
        minus[which_arg] = minus[which_arg] - h * d
        fd = (value(*plus) - value(*minus)) / (2 * h)
        an = float((grad * d).sum())
        assert abs(an - fd) <= 1e-6 * max(1.0, abs(fd)), (which_arg, an, fd)


def test_ot_value_gradient_is_the_envelope(oracle_mod):
    """For L = <f, a> + <g, b> the implicit part vanishes w.r.t. the points:
    dL/dx_i = 2 sum_j P_ij (x_i - y_j) (envelope theorem)."""
    x, y, a, b = _problem(9, n=7, m=12, d=3)
    eps = 0.6

    def value_loss(xx, yy, aa, bb, f, g, e):
        return (f * aa).sum() + (g * bb).sum()

    gx, gy, _, _ = _adjoint_grads(oracle_mod, x, y, a, b, eps, value_loss, "implicit", 200)
    f0, g0, _, _ = oracle_mod.sinkhorn_points(x.numpy(), y.numpy(), a.numpy(), b.numpy(), eps,
                                              tol=1e-14)
    P = _plan(x, y, a, b, torch.tensor(f0), torch.tensor(g0), eps)
    env_x = 2 * (P.sum(1, keepdim=True) * x - P @ y)
    env_y = 2 * (P.sum(0)[:, None] * y - P.T @ x)
    assert torch.allclose(gx, env_x, atol=1e-9) and torch.allclose(gy, env_y, atol=1e-9)


def test_adjoint_system_residual_and_gauge(oracle_mod):
    """The implicit solution satisfies a alpha + P beta = vf, P^T alpha +
    b beta = vg; the unrolled series approaches it."""
    x, y, a, b = _problem(13, n=9, m=9, d=2)
    eps = 1.0
    f0, g0, _, _ = oracle_mod.sinkhorn_points(x.numpy(), y.numpy(), a.numpy(), b.numpy(), eps,
                                              tol=1e-14)
    P = _plan(x, y, a, b, torch.tensor(f0), torch.tensor(g0), eps)
    ops = TorchPlanOps(P)
    rng = np.random.default_rng(2)
    vf = torch.tensor(rng.normal(size=9))
    vg = torch.tensor(rng.normal(size=9))
    vg = vg - (vg.sum() - vf.sum()) / 9  # gauge invariant right-hand side
    al, be = solve_adjoint(ops, a, b, vf, vg, "implicit", 200, 1e-12)
    assert torch.allclose(a * al + P @ be, vf, atol=1e-10)
    assert torch.allclose(P.T @ al + b * be, vg, atol=1e-10)
    al2, be2 = solve_adjoint(ops, a, b, vf, vg, "unrolled", 4000)
    # equal up to the null direction (1, -1)
    k = float(((al2 - al).mean() - (be2 - be).mean()) / 2)
    assert torch.allclose(al2 - k, al, atol=1e-7) and torch.allclose(be2 + k, be, atol=1e-7)
    gx, gy, ga, gb = adjoint_param_grads(ops, x, y, a, b, eps, al, be)
    assert gx.shape == x.shape and gy.shape == y.shape and ga.shape == a.shape
