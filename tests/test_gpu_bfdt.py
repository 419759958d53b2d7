"""GPU parity of the brute-force distance transform (reference
distance_transform.py:141-175, bfdt_kernel.cu) against the CPU oracle:
float32 distances equal to the oracle's FP64 values rounded, indices equal
(first background in row-major order at the smallest distance)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

METRICS = ["euclidean", "taxicab", "chessboard"]


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _case(seed, nd, lead=(2, 1)):
    rng = np.random.default_rng(seed)
    shape = tuple(int(v) for v in rng.integers(1, 9 if nd < 4 else 5, size=nd))
    mask = rng.random(lead + shape) < rng.uniform(0.5, 0.95)
    mask.reshape(-1, *shape)[0].flat[0] = False
    return mask, rng.uniform(0.5, 2.0, size=nd)


@pytest.mark.parametrize("metric", METRICS)
@pytest.mark.parametrize("nd", [1, 2, 3, 4])
@pytest.mark.parametrize("sampled", [False, True])
def test_bfdt_matches_oracle(tm, oracle_mod, metric, nd, sampled):
    mask, samp = _case(10 * nd + METRICS.index(metric), nd)
    samp = samp if sampled else None
    d, i = tm.brute_force_distance_transform(torch.as_tensor(mask, device="cuda"), metric=metric,
                                             sampling=samp, return_indices=True)
    want_d, want_i = oracle_mod.bfdt(mask, metric, samp)
    assert d.dtype == torch.float32 and i.dtype == torch.int32
    assert np.array_equal(d.cpu().numpy(), want_d.astype(np.float32))
    assert np.array_equal(i.cpu().numpy(), want_i)


def test_bfdt_larger_item_crosses_tiles(tm, oracle_mod):
    # several 256-voxel tiles per item and items of the same batch that are
    # all foreground / all background
    rng = np.random.default_rng(7)
    mask = rng.random((3, 1, 23, 37)) < 0.97
    mask[1] = True
    mask[2] = False
    for metric in METRICS:
        d, i = tm.brute_force_distance_transform(torch.as_tensor(mask, device="cuda"),
                                                 metric=metric, sampling=[0.6, 1.7],
                                                 return_indices=True)
        want_d, want_i = oracle_mod.bfdt(mask, metric, [0.6, 1.7])
        assert np.array_equal(d.cpu().numpy(), want_d.astype(np.float32)), metric
        assert np.array_equal(i.cpu().numpy(), want_i), metric
    assert torch.isinf(tm.brute_force_distance_transform(
        torch.as_tensor(mask, device="cuda"))[1]).all()


def test_bfdt_flags_outputs_and_dtypes(tm, oracle_mod):
    mask, _ = _case(3, 2)
    x = torch.as_tensor(mask, device="cuda")
    want_d, want_i = oracle_mod.bfdt(mask, "taxicab")
    only_i = tm.brute_force_distance_transform(x, "taxicab", return_distances=False,
                                               return_indices=True)
    assert np.array_equal(only_i.cpu().numpy(), want_i)
    dist = torch.empty(x.shape, device="cuda")
    idx = torch.empty((2, *x.shape), dtype=torch.int64, device="cuda")
    assert tm.brute_force_distance_transform(x, "taxicab", distances=dist, indices=idx) is None
    assert np.array_equal(dist.cpu().numpy(), want_d.astype(np.float32))
    assert np.array_equal(idx.cpu().numpy(), want_i)
    # float input: nonzero (NaN included) is foreground, as in the reference
    xf = x.float() * 3.0
    xf.view(-1)[5] = float("nan")
    mf = mask.copy()
    mf.reshape(-1)[5] = True
    d = tm.brute_force_distance_transform(xf, "chessboard")
    assert np.array_equal(d.cpu().numpy(), oracle_mod.bfdt(mf, "chessboard")[0].astype(np.float32))
    with pytest.raises(ValueError, match="metric must be"):
        tm.brute_force_distance_transform(x, "invalid")


def test_bfdt_reference_binding_and_edt_agreement(tm, oracle_mod):
    from torchmorph_amd import _C

    mask, _ = _case(11, 3, lead=(1, 2))
    x = torch.as_tensor(mask, device="cuda")
    d, i = _C.bfdt_cuda(x.float(), "euclidean", [1.0, 1.0, 1.0], True, True)
    want_d, want_i = oracle_mod.bfdt(mask)
    assert np.array_equal(d.cpu().numpy(), want_d.astype(np.float32))
    assert np.array_equal(i.cpu().numpy(), want_i)
    # unit sampling: the same float32 distances as the exact EDT
    assert torch.equal(d, tm.euclidean_distance_transform(x))
