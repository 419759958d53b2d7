"""Warm-start network: BF16 MFMA forward/backward against the FP64 oracle of
the same network (bf16-rounded at the same points), training, fallback."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ws():
    from torchmorph_amd import warmstart

    return warmstart


def _pairs(rng, sizes, eps=(0.3, 1.1)):
    from torchmorph_amd.warmstart import gmm_clouds

    xs = gmm_clouds(rng, [n for n, _ in sizes])
    ys = gmm_clouds(rng, [m for _, m in sizes])
    e = [eps[k % len(eps)] for k in range(len(sizes))]
    return xs, ys, e


def _cos(a, b):
    return float((a * b).sum() / (np.linalg.norm(a) * np.linalg.norm(b)))


def test_param_count_and_layout(ws):
    from oracle import deepsets as ds

    assert ws.n_params() == ds.N_PARAMS == 182017
    p = ws.init_params(0, "cpu").double().numpy()
    q = ds.init_params(0)
    assert p.shape == q.shape


def test_forward_matches_oracle(ws):
    from oracle import deepsets as ds

    rng = np.random.default_rng(0)
    sizes = [(300, 200), (64, 513), (1000, 77)]
    xs, ys, eps = _pairs(rng, sizes)
    net = ws.WarmStartNet.create(seed=1)
    got = net.predict([torch.tensor(x, device="cuda") for x in xs],
                      [torch.tensor(y, device="cuda") for y in ys], eps).cpu().numpy()
    params = net.params.cpu().double().numpy()
    want = np.concatenate(ds.predict(params, list(zip(xs, ys, eps)), round_bf16=True))
    scale = np.abs(want).max()
    assert np.abs(got - want).max() <= 2e-2 * scale, np.abs(got - want).max() / scale
    assert _cos(got, want) > 0.9999


def test_gradient_matches_oracle(ws):
    from oracle import deepsets as ds

    rng = np.random.default_rng(1)
    sizes = [(257, 300), (128, 64), (700, 511)]
    xs, ys, eps = _pairs(rng, sizes)
    targets = [rng.normal(size=n).astype(np.float32) * 3 for n, _ in sizes]
    net = ws.WarmStartNet.create(seed=2)
    loss, grad = net.loss_and_grad([torch.tensor(x, device="cuda") for x in xs],
                                   [torch.tensor(y, device="cuda") for y in ys], eps,
                                   torch.tensor(np.concatenate(targets), device="cuda"),
                                   loss_alpha=0.5)
    params = net.params.cpu().double().numpy()
    oloss, ograd = ds.loss_and_grad(params, list(zip(xs, ys, eps)), targets, round_bf16=True,
                                    alpha=0.5)
    assert abs(float(loss) - oloss) <= 1e-3 * oloss
    g = grad.cpu().double().numpy()
    o = 0
    for name, shape in ds.LAYOUT:
        k = int(np.prod(shape))
        a, b = g[o:o + k], ograd[o:o + k]
        o += k
        if np.linalg.norm(b) < 1e-12:
            continue
        assert _cos(a, b) > 0.999, name
        assert np.linalg.norm(a - b) <= 3e-2 * np.linalg.norm(b), name


def test_training_reduces_loss_and_checkpoint_roundtrip(ws, tmp_path):
    rng = np.random.default_rng(2)
    sizes = [(int(n), int(m)) for n, m in rng.integers(64, 400, size=(24, 2))]
    xs, ys, eps = _pairs(rng, sizes, eps=(0.5,))
    X = [torch.tensor(x, device="cuda") for x in xs]
    Y = [torch.tensor(y, device="cuda") for y in ys]
    res = ws.solve(None, X, Y, eps, tol=1e-5)
    net = ws.WarmStartNet.create(seed=3)
    first = float(net.train_step(X, Y, eps, res.f, lr=1e-3))
    for _ in range(150):
        last = float(net.train_step(X, Y, eps, res.f, lr=1e-3))
    assert last < 0.3 * first, (first, last)
    path = tmp_path / "ws.pt"
    net.save(path)
    net2 = ws.WarmStartNet.load(path)
    assert torch.equal(net2.params, net.params) and net2.step == net.step
    assert torch.equal(net2.predict(X, Y, eps), net.predict(X, Y, eps))


def test_nonfinite_prediction_falls_back_to_cold_start(ws):
    rng = np.random.default_rng(3)
    xs, ys, eps = _pairs(rng, [(100, 90), (80, 120)], eps=(0.5,))
    X = [torch.tensor(x, device="cuda") for x in xs]
    Y = [torch.tensor(y, device="cuda") for y in ys]
    net = ws.WarmStartNet.create(seed=4)
    net.params[-1] = float("nan")  # b4 -> every prediction NaN
    warm = ws.solve(net, X, Y, eps, tol=1e-5)
    cold = ws.solve(None, X, Y, eps, tol=1e-5)
    assert torch.equal(warm.f, cold.f) and torch.equal(warm.iterations, cold.iterations)
