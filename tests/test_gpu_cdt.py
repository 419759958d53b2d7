"""GPU parity of the chamfer DT against SciPy golden vectors and the oracle
(values and feature indices bit-exact, both metrics)."""
import numpy as np
import pytest
import torch

from tests import _golden

pytestmark = pytest.mark.gpu

CDT = _golden.load("cdt")


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


@pytest.mark.parametrize("case", CDT, ids=_golden.ids(CDT))
def test_cdt_matches_golden(tm, case):
    meta, arr = case
    x = torch.as_tensor(arr["input"], device="cuda")
    d, i = tm.chamfer_distance_transform(x, metric=meta["metric"], return_indices=True)
    assert torch.equal(d.cpu(), torch.from_numpy(arr["dist"].astype(np.float32)))
    assert np.array_equal(i.cpu().numpy(), arr["idx"])


def _vs_oracle(tm, oracle_mod, mask, metric, idx_dtype=torch.int32):
    x = torch.as_tensor(mask, device="cuda")
    idx = torch.empty((x.ndim - 2, *x.shape), dtype=idx_dtype, device="cuda")
    d = tm.chamfer_distance_transform(x, metric=metric, indices=idx)
    od, oi = oracle_mod.cdt(mask, metric)
    assert torch.equal(d.cpu(), torch.from_numpy(od.astype(np.float32)))
    assert np.array_equal(idx.cpu().numpy(), oi)


@pytest.mark.parametrize("metric", ["chessboard", "taxicab"])
@pytest.mark.parametrize("shape", [(3, 1, 130, 77), (1, 2, 64, 200), (2, 1, 1, 300),
                                   (1, 1, 300, 1), (2, 1, 17, 19, 23), (1, 1, 40, 3, 70)])
def test_cdt_random_vs_oracle(tm, oracle_mod, metric, shape):
    rng = np.random.default_rng(sum(shape))
    mask = rng.random(shape) < 0.9
    _vs_oracle(tm, oracle_mod, mask, metric)


@pytest.mark.parametrize("metric", ["chessboard", "taxicab"])
def test_cdt_sparse_long_diagonals(tm, oracle_mod, metric):
    """Long diagonal chains the reference's Jacobi passes cannot resolve."""
    mask = np.ones((2, 1, 150, 130), bool)
    mask[0, 0, 0, 0] = False
    mask[1, 0, 149, 7] = False
    _vs_oracle(tm, oracle_mod, mask, metric, torch.int64)


def test_cdt_4d(tm, oracle_mod):
    rng = np.random.default_rng(4)
    mask = rng.random((1, 2, 5, 4, 6, 7)) < 0.85
    _vs_oracle(tm, oracle_mod, mask, "chessboard")
    _vs_oracle(tm, oracle_mod, mask, "taxicab")


def test_cdt_metric_aliases(tm):
    meta, arr = CDT[2]
    x = torch.as_tensor(arr["input"], device="cuda")
    for alias in ("cityblock", "manhattan"):
        assert torch.equal(tm.chamfer_distance_transform(x, metric=alias),
                           tm.chamfer_distance_transform(x, metric="taxicab"))
    with pytest.raises(ValueError, match="metric must be"):
        tm.chamfer_distance_transform(x, metric="invalid")


def test_cdt_is_deterministic(tm):
    g = torch.Generator().manual_seed(3)
    x = (torch.rand((2, 1, 33, 40, 50), generator=g) < 0.9).cuda()
    a = tm.chamfer_distance_transform(x, return_indices=True)
    b = tm.chamfer_distance_transform(x, return_indices=True)
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


@pytest.mark.parametrize("metric", ["chessboard", "taxicab"])
@pytest.mark.parametrize("shape,p", [((3, 1, 130, 77), 0.9), ((1, 2, 64, 200), 0.5),
                                     ((2, 1, 300), 0.97), ((1, 1, 300, 1), 0.9),
                                     ((2, 1, 17, 19, 23), 0.9), ((1, 1, 40, 3, 70), 0.995),
                                     ((1, 2, 5, 4, 6, 7), 0.85), ((1, 1, 9, 11), 1.0),
                                     ((1, 1, 9, 11), 0.0), ((2, 1, 1, 1), 1.0)])
def test_cdt_distances_only_vs_oracle(tm, oracle_mod, metric, shape, p):
    """Without indices the values come from the separable per-axis passes;
    they must equal SciPy's raster result exactly (-1 on all-foreground)."""
    rng = np.random.default_rng(sum(shape) + int(p * 100))
    mask = rng.random(shape) < p
    for dtype in (torch.bool, torch.float32):
        x = torch.as_tensor(mask, device="cuda").to(dtype)
        d = tm.chamfer_distance_transform(x, metric=metric)
        od, _ = oracle_mod.cdt(mask, metric)
        assert torch.equal(d.cpu(), torch.from_numpy(od.astype(np.float32)))


@pytest.mark.parametrize("metric", ["chessboard", "taxicab"])
def test_cdt_distances_only_single_site(tm, oracle_mod, metric):
    mask = np.ones((2, 1, 150, 130), bool)
    mask[0, 0, 0, 0] = False
    mask[1, 0, 149, 7] = False
    d = tm.chamfer_distance_transform(torch.as_tensor(mask, device="cuda"), metric=metric)
    od, _ = oracle_mod.cdt(mask, metric)
    assert torch.equal(d.cpu(), torch.from_numpy(od.astype(np.float32)))
