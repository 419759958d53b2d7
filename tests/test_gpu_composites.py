"""Composite grey operators (reference grey.py:200-289): the fused kernels
(tm_grey_ex) and the fallback composition both equal the oracle's erosion /
dilation combined with the reference's float32 tensor expressions, bit for
bit."""
import numpy as np
import pytest
import torch

from tests._golden import assert_bits_equal

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _oracle_composite(oracle_mod, name, x, st, fp, se_shape, origin, mode, cval):
    def ero(v):
        return oracle_mod.grey(v, st, fp, se_shape, origin, mode, cval, False)

    def dil(v):
        return oracle_mod.grey(v, st, fp, se_shape, origin, mode, cval, True)

    x = x.astype(np.float32)
    if name == "morphological_gradient":
        return dil(x) - ero(x)
    if name == "morphological_laplace":
        return (dil(x) + ero(x)) - 2 * x
    if name == "white_tophat":
        return x - dil(ero(x))
    return ero(dil(x)) - x


def _element(kind, nd, rng):
    """(kwargs for the API, structure, footprint, se_shape, origin) of the oracle."""
    if kind == "box":
        shp = (3, 5, 3)[-nd:]
        return dict(size=shp), None, None, shp, (0,) * nd
    if kind == "box_even_origin":
        shp = (4, 3, 2)[-nd:]
        org = (1, -1, 0)[-nd:]
        return dict(size=shp, origin=org), None, None, shp, org
    if kind == "footprint":
        fp = rng.random((3,) * nd) < 0.6
        fp.flat[0] = True
        return dict(footprint=torch.as_tensor(fp)), None, fp, fp.shape, (0,) * nd
    st = rng.integers(0, 4, size=(3,) * nd).astype(np.float32)
    return dict(structure=torch.as_tensor(st)), st, None, st.shape, (0,) * nd


OPS = ["morphological_gradient", "morphological_laplace", "white_tophat", "black_tophat"]


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("kind", ["box", "box_even_origin", "footprint", "heights"])
@pytest.mark.parametrize("shape", [(2, 1, 40), (2, 2, 23, 41), (1, 2, 9, 17, 70)])
@pytest.mark.parametrize("mode", ["reflect", "constant", "wrap"])
def test_grey_composite_matches_oracle(tm, oracle_mod, op, kind, shape, mode):
    rng = np.random.default_rng(len(shape) * 7 + OPS.index(op))
    nd = len(shape) - 2
    x = (np.floor(rng.random(shape) * 32) / 8).astype(np.float32)
    kw, st, fp, se_shape, origin = _element(kind, nd, rng)
    got = getattr(tm, op)(torch.as_tensor(x, device="cuda"), mode=mode, cval=0.75, **kw)
    want = _oracle_composite(oracle_mod, op, x, st, fp, se_shape, origin, mode, 0.75)
    assert got.dtype == torch.float32
    assert_bits_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize("op", OPS)
def test_grey_composite_fallback_paths(tm, oracle_mod, op):
    # 4 spatial axes and an element beyond the table kernel: the plain
    # composition runs, with the same results
    rng = np.random.default_rng(3)
    x = (np.floor(rng.random((1, 1, 5, 6, 7, 8)) * 16) / 4).astype(np.float32)
    got = getattr(tm, op)(torch.as_tensor(x, device="cuda"), size=2)
    want = _oracle_composite(oracle_mod, op, x, None, None, (2, 2, 2, 2), (0, 0, 0, 0),
                             "reflect", 0.0)
    assert_bits_equal(got.cpu().numpy(), want)
    st = rng.integers(0, 3, size=(1, 20, 15)).astype(np.float32)  # 300 elements
    x3 = (np.floor(rng.random((1, 1, 4, 30, 30)) * 16) / 4).astype(np.float32)
    got = getattr(tm, op)(torch.as_tensor(x3, device="cuda"), structure=torch.as_tensor(st))
    want = _oracle_composite(oracle_mod, op, x3, st, None, st.shape, (0, 0, 0), "reflect", 0.0)
    assert_bits_equal(got.cpu().numpy(), want)


def test_grey_composite_output_argument_and_dtype(tm, oracle_mod):
    rng = np.random.default_rng(5)
    x = rng.integers(0, 9, size=(2, 1, 16, 20)).astype(np.int64)
    out = torch.empty((2, 1, 16, 20), dtype=torch.float32, device="cuda")
    r = tm.morphological_gradient(torch.as_tensor(x, device="cuda"), size=3, output=out)
    assert r is out
    want = _oracle_composite(oracle_mod, "morphological_gradient", x.astype(np.float32), None,
                             None, (3, 3), (0, 0), "reflect", 0.0)
    assert_bits_equal(out.cpu().numpy(), want)
