"""Hand-written grey erosion/dilation backward: routing rule against the CPU
oracle, and FP64 finite differences at tie-free points."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _grad(tm, x, dy, dilation, **kw):
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    fn = tm.grey_dilation if dilation else tm.grey_erosion
    y = fn(xt, differentiable=True, **kw)
    assert y.requires_grad
    y.backward(torch.tensor(dy, device="cuda"))
    return y.detach().cpu().numpy(), xt.grad.cpu().numpy()


CASES = [
    dict(size=3, mode="reflect"),
    dict(size=(3, 5), mode="constant", cval=0.3),
    dict(size=(2, 3), mode="nearest", origin=(0, 1)),
    dict(size=3, mode="mirror"),
    dict(size=(3, 3), mode="wrap"),
]


@pytest.mark.parametrize("kw", CASES, ids=lambda k: f"{k['mode']}")
@pytest.mark.parametrize("dilation", [False, True])
def test_grey_backward_matches_oracle_with_ties(tm, oracle_mod, kw, dilation):
    """Quantised input (many ties): exact routing rule vs the oracle."""
    rng = np.random.default_rng(0)
    x = (np.floor(rng.random((2, 1, 23, 31)) * 8) / 8).astype(np.float32)
    dy = rng.normal(size=x.shape).astype(np.float32)
    y, dx = _grad(tm, x, dy, dilation, **kw)
    size = kw["size"]
    se = (size, size) if isinstance(size, int) else size
    origin = kw.get("origin", (0, 0))
    for b in range(2):
        want = oracle_mod.grey_backward(x[b, 0], None, None, se, origin, kw["mode"],
                                        kw.get("cval", 0.0), dilation, dy[b, 0])
        np.testing.assert_allclose(dx[b, 0], want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dilation", [False, True])
def test_grey_backward_nonflat_3d_vs_oracle(tm, oracle_mod, dilation):
    rng = np.random.default_rng(1)
    x = rng.random((1, 1, 9, 10, 12)).astype(np.float32)
    zz, yy, xx = np.mgrid[-1:2, -1:2, -1:2]
    fp = (zz ** 2 + yy ** 2 + xx ** 2) <= 2
    st = rng.integers(0, 4, size=fp.shape).astype(np.float32)
    dy = rng.normal(size=x.shape).astype(np.float32)
    _, dx = _grad(tm, x, dy, dilation, footprint=torch.tensor(fp),
                  structure=torch.tensor(st), mode="reflect")
    want = oracle_mod.grey_backward(x[0, 0], st, fp, (3, 3, 3), (0, 0, 0), "reflect", 0.0,
                                    dilation, dy[0, 0])
    np.testing.assert_allclose(dx[0, 0], want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("dilation", [False, True])
def test_grey_backward_fp64_finite_differences(tm, oracle_mod, dilation):
    """Tie-free input: the routed gradient is the exact derivative; check
    directional derivatives against central differences of an FP64 forward."""
    rng = np.random.default_rng(2)
    x = rng.permutation(20 * 24).reshape(1, 1, 20, 24).astype(np.float32) / 7.0
    dy = rng.normal(size=x.shape)
    _, dx = _grad(tm, x, dy.astype(np.float32), dilation, size=3, mode="reflect")
    for k in range(3):
        v = rng.normal(size=x.shape)
        h = 1e-6
        fp = oracle_mod.grey_fp64(x[0, 0].astype(np.float64) + h * v[0, 0], None, None, (3, 3),
                                  (0, 0), "reflect", 0.0, dilation)
        fm = oracle_mod.grey_fp64(x[0, 0].astype(np.float64) - h * v[0, 0], None, None, (3, 3),
                                  (0, 0), "reflect", 0.0, dilation)
        fd = ((fp - fm) / (2 * h) * dy[0, 0]).sum()
        an = (dx.astype(np.float64) * v).sum()
        assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (fd, an)


def test_grey_backward_config4_shape(tm, oracle_mod):
    """16x1x512x512 float32, flat size 3: full-size run, one slice checked."""
    g = torch.Generator().manual_seed(4)
    x = torch.rand((16, 1, 512, 512), generator=g)
    dy = torch.randn((16, 1, 512, 512), generator=g)
    xt = x.cuda().requires_grad_(True)
    tm.grey_dilation(xt, size=3, differentiable=True).backward(dy.cuda())
    want = oracle_mod.grey_backward(x[3, 0, :64, :64].numpy(), None, None, (3, 3), (0, 0),
                                    "reflect", 0.0, True, dy[3, 0, :64, :64].numpy())
    # interior of the crop must match (outputs on the crop's own border see
    # a reflected neighbourhood, so their routing reaches two voxels in)
    np.testing.assert_allclose(xt.grad[3, 0, 2:62, 2:62].cpu().numpy(), want[2:62, 2:62],
                               rtol=1e-6, atol=1e-6)
    # total routed gradient equals the total upstream gradient (no cval reads)
    assert abs(float(xt.grad.double().sum()) - float(dy.double().sum())) < 1e-2


@pytest.mark.parametrize("mode", ["reflect", "mirror", "nearest", "constant", "wrap"])
@pytest.mark.parametrize("dilation", [False, True])
@pytest.mark.parametrize("shape", [(1, 1, 3, 2), (2, 1, 1, 5), (1, 1, 5, 1), (1, 1, 70, 130)])
def test_grey_backward_tiny_items_and_large_elements(tm, oracle_mod, mode, dilation, shape):
    """Items smaller than the element (many reflections of a border voxel)
    and a 7 x 5 non-flat element with an origin, against the oracle."""
    rng = np.random.default_rng(sum(shape))
    x = (np.floor(rng.random(shape) * 6) / 6).astype(np.float32)
    dy = rng.normal(size=shape).astype(np.float32)
    fp = rng.random((7, 5)) < 0.7
    fp[3, 2] = True
    st = rng.integers(0, 3, size=(7, 5)).astype(np.float32)
    kw = dict(footprint=torch.tensor(fp), structure=torch.tensor(st), mode=mode, cval=0.5,
              origin=(1, -2))
    _, dx = _grad(tm, x, dy, dilation, **kw)
    for b in range(shape[0]):
        want = oracle_mod.grey_backward(x[b, 0], st, fp, (7, 5), (1, -2), mode, 0.5, dilation,
                                        dy[b, 0])
        np.testing.assert_allclose(dx[b, 0], want, rtol=1e-6, atol=1e-6)


@pytest.mark.parametrize("mode", ["reflect", "mirror", "nearest", "constant"])
@pytest.mark.parametrize("dilation", [False, True])
@pytest.mark.parametrize("shape", [(2, 1, 23, 131), (1, 1, 70, 64), (1, 1, 3, 2), (2, 1, 1, 5),
                                   (1, 1, 5, 1), (1, 1, 2, 63), (1, 1, 17, 62)])
def test_grey_backward_flat3_every_variant(tm, oracle_mod, mode, dilation, shape):
    """The two launch variants of the flat 3 x 3 backward (0: LDS tiles, 1:
    register streaming) route and sum identically (bit for bit: both sum in
    element order) and match the oracle (which sums in its own order, hence
    the tolerance), on quantised input (many ties), across strip and tile
    edges and items smaller than the element; signed zeros and infinities in
    the upstream gradient."""
    from torchmorph_amd import autotune
    from tests._golden import assert_bits_equal

    rng = np.random.default_rng(sum(shape) + 3)
    x = (np.floor(rng.random(shape) * 5) / 5).astype(np.float32)
    dy = rng.normal(size=shape).astype(np.float32)
    flat = dy.reshape(-1)
    if flat.size > 8:
        flat[:4] = [-0.0, 0.0, np.inf, -np.inf]
    p2 = lambda v: 1 << (int(v) - 1).bit_length()
    key = f"greybw {int(dilation)} {p2(shape[2])} {p2(shape[3])} {p2(shape[0] * shape[1])}"
    got = []
    try:
        for v in (0, 1):
            autotune.clear()
            autotune.load({key: v})
            got.append(_grad(tm, x, dy, dilation, size=3, mode=mode, cval=0.25)[1])
    finally:
        autotune.clear()
    assert_bits_equal(got[1], got[0], "register variant vs LDS variant")
    for b in range(shape[0]):
        want = oracle_mod.grey_backward(x[b, 0], None, None, (3, 3), (0, 0), mode, 0.25,
                                        dilation, dy[b, 0])
        np.testing.assert_allclose(got[1][b, 0], want, rtol=1e-6, atol=1e-6)
