"""GPU parity of the exact EDT against the SciPy golden vectors and the CPU
oracle (bit-exact float32 distances, exact int32/int64 indices)."""
import numpy as np
import pytest
import torch

from tests import _golden

pytestmark = pytest.mark.gpu

EDT = _golden.load("edt")


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _f32(a):
    return torch.from_numpy(np.asarray(a, np.float64).astype(np.float32))


@pytest.mark.parametrize("case", EDT, ids=_golden.ids(EDT))
@pytest.mark.parametrize("idx_dtype", [torch.int32, torch.int64])
def test_edt_matches_golden(tm, case, idx_dtype):
    meta, arr = case
    x = torch.as_tensor(arr["input"], device="cuda")
    nd = x.ndim - 2
    indices = torch.empty((nd, *x.shape), dtype=idx_dtype, device="cuda")
    d = tm.euclidean_distance_transform(x, sampling=meta["sampling"], indices=indices)
    assert torch.equal(d.cpu(), _f32(arr["dist"]))
    assert np.array_equal(indices.cpu().numpy(), arr["idx"])


@pytest.mark.parametrize("dtype", [torch.bool, torch.float32, torch.uint8, torch.int64,
                                   torch.float64])
def test_edt_input_dtypes(tm, dtype):
    meta, arr = EDT[3]
    x = torch.as_tensor(arr["input"]).to(dtype).cuda()
    d, i = tm.euclidean_distance_transform(x, return_indices=True)
    assert torch.equal(d.cpu(), _f32(arr["dist"]))
    assert i.dtype == torch.int32
    assert np.array_equal(i.cpu().numpy(), arr["idx"])


def _check_vs_oracle(tm, oracle_mod, mask, sampling=None, idx_dtype=torch.int32):
    x = torch.as_tensor(mask, device="cuda")
    nd = x.ndim - 2
    idx = torch.empty((nd, *x.shape), dtype=idx_dtype, device="cuda")
    d = tm.euclidean_distance_transform(x, sampling=sampling, indices=idx)
    od, oi = oracle_mod.edt(mask, sampling)
    assert torch.equal(d.cpu(), _f32(od)), "distances differ"
    assert np.array_equal(idx.cpu().numpy(), oi), "indices differ"


def test_edt_config1_full_size(tm, oracle_mod):
    """B=8, C=1, 256x256, p_fg=0.9, int64 indices (the headline EDT workload)."""
    g = torch.Generator().manual_seed(0)
    mask = (torch.rand((8, 1, 256, 256), generator=g) < 0.9).numpy()
    _check_vs_oracle(tm, oracle_mod, mask, idx_dtype=torch.int64)


@pytest.mark.parametrize("shape", [(1, 1, 37, 1024), (1, 1, 9, 1025), (2, 1, 300, 5),
                                   (1, 2, 1, 700), (1, 1, 700, 1), (3, 1, 129, 131)])
def test_edt_2d_shapes_across_paths(tm, oracle_mod, shape):
    rng = np.random.default_rng(sum(shape))
    mask = rng.random(shape) < 0.93
    _check_vs_oracle(tm, oracle_mod, mask)


def test_edt_sparse_background(tm, oracle_mod):
    mask = np.ones((2, 1, 300, 500), bool)
    mask[0, 0, 123, 77] = False
    mask[1, 0, 0, 499] = False
    mask[1, 0, 299, 0] = False
    _check_vs_oracle(tm, oracle_mod, mask)


@pytest.mark.parametrize("kind", ["one_site", "left_column", "p_bg_1e-4"])
def test_edt_sparse_background_1024x4096(tm, oracle_mod, kind):
    """The largest 2-D-path shape with sparse background: row windows pass
    the step cap and rows are resolved through their envelopes."""
    rng = np.random.default_rng(3)
    mask = np.ones((1, 1, 1024, 4096), bool)
    if kind == "one_site":
        mask[0, 0, 500, 2000] = False
    elif kind == "left_column":
        mask[0, 0, :, 0] = False
    else:
        mask = rng.random((1, 1, 1024, 4096)) > 1e-4
    _check_vs_oracle(tm, oracle_mod, mask, idx_dtype=torch.int64)


@pytest.mark.parametrize("sampling", [None, [1.0, 1.0, 1.0], [1.3, 0.7, 1.9], [0.5, 2.0, 1.0]])
def test_edt_3d_vs_oracle(tm, oracle_mod, sampling):
    rng = np.random.default_rng(5)
    mask = rng.random((2, 1, 40, 33, 47)) < 0.9
    _check_vs_oracle(tm, oracle_mod, mask, sampling)


def test_edt_3d_long_lines_use_global_stack(tm, oracle_mod):
    rng = np.random.default_rng(6)
    mask = rng.random((1, 1, 3, 300, 200)) < 0.97
    _check_vs_oracle(tm, oracle_mod, mask, [0.8, 1.1, 0.9])


def test_edt_2d_anisotropic(tm, oracle_mod):
    rng = np.random.default_rng(8)
    mask = rng.random((4, 1, 120, 90)) < 0.9
    _check_vs_oracle(tm, oracle_mod, mask, [0.63, 1.71])


def test_edt_1d(tm, oracle_mod):
    rng = np.random.default_rng(9)
    mask = rng.random((3, 2, 1000)) < 0.95
    _check_vs_oracle(tm, oracle_mod, mask, [1.25])


def test_edt_is_deterministic(tm):
    g = torch.Generator().manual_seed(1)
    x = (torch.rand((2, 1, 64, 96, 80), generator=g) < 0.95).cuda()
    a = tm.euclidean_distance_transform(x, return_indices=True, sampling=[1.1, 0.9, 1.3])
    b = tm.euclidean_distance_transform(x, return_indices=True, sampling=[1.1, 0.9, 1.3])
    assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1])


def test_edt_return_flags_and_preallocated(tm):
    meta, arr = EDT[1]
    x = torch.as_tensor(arr["input"], device="cuda")
    d = tm.euclidean_distance_transform(x)
    i = tm.euclidean_distance_transform(x, return_distances=False, return_indices=True)
    assert d.shape == x.shape and i.shape == (2, *x.shape)
    dist64 = torch.empty_like(x, dtype=torch.float64)
    idx = torch.empty((2, *x.shape), dtype=torch.int32, device="cuda")
    r = tm.euclidean_distance_transform(x, return_distances=False, return_indices=False,
                                        distances=dist64, indices=idx)
    assert r is None
    torch.testing.assert_close(dist64, d.double())
    torch.testing.assert_close(idx, i)
    with pytest.raises(ValueError, match="At least one"):
        tm.euclidean_distance_transform(x, return_distances=False, return_indices=False)


def test_edt_uses_current_stream(tm):
    """Reference test_distance_transform.py:357-394 pattern."""
    x = torch.ones((1, 1, 16, 16), device="cuda")
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        torch.cuda._sleep(5_000_000)
        x.zero_()
        d = tm.euclidean_distance_transform(x)
    stream.synchronize()
    assert torch.equal(d, torch.zeros_like(d))


def test_edt_noncontiguous_input(tm, oracle_mod):
    rng = np.random.default_rng(11)
    base = torch.as_tensor(rng.random((2, 2, 30, 20)) < 0.85, device="cuda")
    x = base.transpose(-1, -2)
    assert not x.is_contiguous()
    _, want = None, oracle_mod.edt(x.cpu().numpy())
    d, i = tm.euclidean_distance_transform(x, return_indices=True)
    assert torch.equal(d.cpu(), _f32(want[0]))
    assert np.array_equal(i.cpu().numpy(), want[1])
