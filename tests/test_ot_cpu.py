"""Sinkhorn API checks that run on the CPU (reference
test/test_optimal_transport.py, CPU parametrisation)."""
import numpy as np
import pytest
import torch

from torchmorph_amd import SinkhornSolver, build_cost_matrix


def _make_positive(shape, device, seed):
    generator = torch.Generator(device=device).manual_seed(seed)
    return torch.rand(shape, generator=generator, device=device) + 0.1


def test_build_cost_matrix_known_answers():
    c1 = build_cost_matrix((5,))
    assert c1.shape == (5, 5) and torch.allclose(c1, c1.T)
    assert torch.allclose(c1.diag(), torch.zeros(5))
    assert c1[0, 1].item() == pytest.approx(1.0)
    c2 = build_cost_matrix((3, 4))
    assert c2.shape == (12, 12) and c2[0, 1].item() == pytest.approx(1.0)
    assert build_cost_matrix((2, 2), p=1)[0, 3].item() == pytest.approx(2.0)
    assert build_cost_matrix((2, 2), p=2)[0, 3].item() == pytest.approx(2.0 ** 0.5)


@pytest.mark.parametrize("log_space", [False, True])
@pytest.mark.parametrize("n,d", [(1, 4), (3, 6), (2, 16)])
def test_cpu_forward_shapes_and_marginals(n, d, log_space):
    solver = SinkhornSolver(epsilon=1.5, max_iter=300, log_space=log_space)
    source = _make_positive((n, d), "cpu", 11)
    target = _make_positive((n, d), "cpu", 23)
    distance = solver(source, target)
    plan = solver.plan(source, target)
    f, g = solver.potentials(source, target)
    assert distance.shape == (n,) and plan.shape == (n, d, d) and f.shape == (n, d)
    a, b, cost = solver.data_preprocess(source, target)
    assert torch.allclose(distance, (plan * cost).sum(dim=(-2, -1)), rtol=1e-5, atol=1e-6)
    assert torch.allclose(plan.sum(dim=-1), a, atol=1e-4)
    assert torch.allclose(plan.sum(dim=-2), b, atol=1e-4)


def test_validation_messages():
    solver = SinkhornSolver(epsilon=1.5, max_iter=10)
    with pytest.raises(ValueError, match=r"\(n, d\)"):
        solver(_make_positive((4,), "cpu", 1), _make_positive((4,), "cpu", 2))
    with pytest.raises(ValueError, match="cost_matrix"):
        solver(_make_positive((2, 4), "cpu", 5), _make_positive((2, 4), "cpu", 7),
               cost_matrix=torch.zeros(3, 3))
    for kw, msg in [(dict(epsilon=0), "epsilon"), (dict(max_iter=0), "max_iter"),
                    (dict(threshold=-1), "threshold")]:
        with pytest.raises(ValueError, match=msg):
            SinkhornSolver(**kw)


@pytest.mark.parametrize("log_space", [False, True])
def test_cpu_backward_matches_numeric_directional_derivative(log_space):
    """Reference test_optimal_transport.py:254-286 on the CPU path."""
    n, d = 3, 8
    solver = SinkhornSolver(epsilon=1.5, max_iter=3000, log_space=log_space)
    cost = build_cost_matrix((d,)).double()
    source = _make_positive((n, d), "cpu", 1009).double().requires_grad_(True)
    target = _make_positive((n, d), "cpu", 2003).double().requires_grad_(True)
    direction = _make_positive((n, d), "cpu", 3001).double() - 0.6
    solver(source, target, cost).sum().backward()

    def entropic_cost(src, tgt):
        plan = solver.plan(src, tgt, cost)
        entropy = (plan * plan.clamp(min=1e-300).log()).sum()
        return (plan * cost).sum() + solver.epsilon * entropy

    eps_fd = 1e-4
    src, tgt = source.detach(), target.detach()
    for grad, plus, minus in [
        (source.grad, (src + eps_fd * direction, tgt), (src - eps_fd * direction, tgt)),
        (target.grad, (src, tgt + eps_fd * direction), (src, tgt - eps_fd * direction)),
    ]:
        analytic = (grad * direction).sum()
        numeric = (entropic_cost(*plus) - entropic_cost(*minus)) / (2 * eps_fd)
        assert torch.allclose(analytic, numeric, rtol=1e-5, atol=1e-7)


def test_cpu_path_agrees_with_fp64_oracle(oracle_mod):
    solver = SinkhornSolver(epsilon=1.5, max_iter=500, log_space=True)
    s = _make_positive((2, 6), "cpu", 41).double()
    t = _make_positive((2, 6), "cpu", 43).double()
    a, b, cost = solver.data_preprocess(s, t)
    from torchmorph_amd.optimal_transport import _hist_solve

    lu, lv = _hist_solve(a, b, cost, 1.5, 500, 0.0, True).log_scalings()
    olu, olv = oracle_mod.sinkhorn_log_hist(a.numpy(), b.numpy(), cost.numpy(), 1.5, 500)
    np.testing.assert_allclose(lu.numpy(), olu, atol=1e-10)
    np.testing.assert_allclose(lv.numpy(), olv, atol=1e-10)
