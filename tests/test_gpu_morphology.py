"""GPU parity of grey and binary morphology against SciPy golden vectors and
the CPU oracle (bit-exact)."""
import numpy as np
import pytest
import torch

from tests import _golden
from tests.test_oracle import _binary_kwargs, _grey_kwargs

pytestmark = pytest.mark.gpu

GREY = _golden.load("grey")
BINARY = _golden.load("binary")


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _cuda(a, dtype=None):
    return None if a is None else torch.as_tensor(np.asarray(a), device="cuda", dtype=dtype)


@pytest.mark.parametrize("case", GREY, ids=_golden.ids(GREY))
@pytest.mark.parametrize("op", ["erosion", "dilation", "opening", "closing"])
def test_grey_matches_golden(tm, case, op):
    meta, arr = case
    x = torch.as_tensor(arr["input"], device="cuda")
    fn = getattr(tm, f"grey_{op}")
    kw = {}
    if "size" in meta:
        kw["size"] = tuple(meta["size"])
    if meta.get("footprint") is not None:
        kw["footprint"] = _cuda(meta["footprint"], torch.bool)
    if meta.get("structure") is not None:
        kw["structure"] = _cuda(meta["structure"], torch.float32)
    origin = meta.get("origin", 0)
    kw["origin"] = origin if isinstance(origin, int) else tuple(origin)
    kw["mode"] = meta.get("mode", "reflect")
    kw["cval"] = meta.get("cval", 0.0)
    got = fn(x, **kw)
    assert got.dtype == torch.float32
    _golden.assert_bits_equal(got.cpu().numpy(), arr[op])


def _grey_vs_oracle(tm, oracle_mod, x, dilation, **kw):
    fn = tm.grey_dilation if dilation else tm.grey_erosion
    t = {}
    if kw.get("size") is not None:
        t["size"] = kw["size"]
    if kw.get("footprint") is not None:
        t["footprint"] = torch.as_tensor(kw["footprint"])
    if kw.get("structure") is not None:
        t["structure"] = torch.as_tensor(kw["structure"])
    got = fn(torch.as_tensor(x, device="cuda"), mode=kw.get("mode", "reflect"),
             cval=kw.get("cval", 0.0), origin=kw.get("origin", 0), **t)
    nd = x.ndim - 2
    st = kw.get("structure")
    fp = kw.get("footprint")
    se = st.shape if st is not None else (fp.shape if fp is not None else kw["size"])
    origin = kw.get("origin", 0)
    origin = (origin,) * nd if isinstance(origin, int) else origin
    want = oracle_mod.grey(x, st, fp, se, origin, kw.get("mode", "reflect"), kw.get("cval", 0.0),
                           dilation)
    _golden.assert_bits_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize("size", [3, 7, 15])
@pytest.mark.parametrize("dilation", [False, True])
def test_grey_flat_3d_config_sizes(tm, oracle_mod, size, dilation):
    """Config-2 element sizes on a reduced volume (quantised values, ties)."""
    rng = np.random.default_rng(size)
    x = (np.floor(rng.random((2, 2, 40, 37, 70)) * 256) / 256).astype(np.float32)
    _grey_vs_oracle(tm, oracle_mod, x, dilation, size=(size,) * 3)


@pytest.mark.parametrize("mode", ["reflect", "constant", "nearest", "mirror", "wrap"])
@pytest.mark.parametrize("dilation", [False, True])
def test_grey_ball_heights_3d(tm, oracle_mod, mode, dilation):
    rng = np.random.default_rng(7)
    x = (np.floor(rng.random((1, 2, 30, 21, 67)) * 256) / 256).astype(np.float32)
    zz, yy, xx = np.mgrid[-3:4, -3:4, -3:4]
    ball = (zz ** 2 + yy ** 2 + xx ** 2) <= 9
    assert ball.sum() == 123
    heights = rng.integers(0, 4, size=ball.shape).astype(np.float32)
    _grey_vs_oracle(tm, oracle_mod, x, dilation, footprint=ball, structure=heights, mode=mode,
                    cval=0.5)


@pytest.mark.parametrize("dilation", [False, True])
def test_grey_2d_1d_and_4d(tm, oracle_mod, dilation):
    rng = np.random.default_rng(1)
    x2 = rng.random((3, 1, 45, 130)).astype(np.float32)
    _grey_vs_oracle(tm, oracle_mod, x2, dilation, size=(5, 4), origin=(1, -1), mode="mirror")
    x1 = rng.random((2, 2, 300)).astype(np.float32)
    _grey_vs_oracle(tm, oracle_mod, x1, dilation, size=(9,), mode="wrap")
    x4 = rng.random((1, 2, 5, 6, 4, 7)).astype(np.float32)
    _grey_vs_oracle(tm, oracle_mod, x4, dilation, size=(3, 2, 3, 3), mode="nearest")
    fp = rng.random((3, 3, 3, 3)) < 0.5
    fp[1, 1, 1, 1] = True
    _grey_vs_oracle(tm, oracle_mod, x4, dilation, footprint=fp, mode="constant", cval=-2.0)


def test_grey_large_box_uses_fallback(tm, oracle_mod):
    rng = np.random.default_rng(2)
    x = rng.random((1, 1, 20, 30, 90)).astype(np.float32)
    _grey_vs_oracle(tm, oracle_mod, x, True, size=(21, 3, 41))
    _grey_vs_oracle(tm, oracle_mod, x, False, size=(3, 25, 5))


@pytest.mark.parametrize("mode", ["reflect", "constant", "wrap"])
def test_grey_halo_tiles_beyond_register_staging(tm, oracle_mod, mode):
    """Halo tiles of more than 16 x 256 values: the elements past the lanes'
    registers are staged directly (flat box and sparse table, 2-D and 3-D)."""
    rng = np.random.default_rng(4)
    x = rng.random((1, 2, 5, 50, 140)).astype(np.float32)
    _grey_vs_oracle(tm, oracle_mod, x, True, size=(1, 33, 65), mode=mode, cval=0.25)
    _grey_vs_oracle(tm, oracle_mod, x, False, size=(2, 30, 80), origin=(0, 3, -5), mode=mode,
                    cval=0.25)
    fp = np.zeros((1, 31, 101), bool)
    fp[0, rng.integers(0, 31, 120), rng.integers(0, 101, 120)] = True
    _grey_vs_oracle(tm, oracle_mod, x, True, footprint=fp, mode=mode, cval=0.25)


def test_grey_signed_zero_follows_reference(tm, oracle_mod):
    x = np.zeros((1, 1, 6, 6), np.float32)
    x[0, 0, ::2] = -0.0
    _grey_vs_oracle(tm, oracle_mod, x, True, size=(3, 3))
    got = tm.grey_dilation(torch.as_tensor(x, device="cuda"), size=3).cpu().numpy()
    assert not np.signbit(got).any()  # the reference's "+ 0" turns -0 into +0


def test_grey_not_differentiable_by_default(tm):
    x = torch.rand((1, 1, 8, 8), device="cuda", requires_grad=True)
    assert not tm.grey_dilation(x, size=3).requires_grad


@pytest.mark.parametrize("case", BINARY, ids=_golden.ids(BINARY))
@pytest.mark.parametrize("op", ["erosion", "dilation", "opening", "closing"])
def test_binary_matches_golden(tm, oracle_mod, case, op):
    meta, arr = case
    x = torch.as_tensor(arr["input"], device="cuda", dtype=torch.float32)
    kw = {}
    if meta.get("structure") is not None:
        kw["structure"] = _cuda(meta["structure"], torch.bool)
    if meta.get("mask") is not None:
        kw["mask"] = _cuda(meta["mask"], torch.bool).expand(x.shape)
    for key in ("iterations", "border_value"):
        if key in meta:
            kw[key] = meta[key]
    if "origin" in meta:
        o = meta["origin"]
        kw["origin"] = o if isinstance(o, int) else tuple(o)
    got = getattr(tm, f"binary_{op}")(x, **kw)
    assert got.dtype == torch.bool
    if op in ("opening", "closing") and meta.get("mask") is not None:
        # SciPy's opening / closing take no mask: chain the oracle's masked
        # passes, each restoring its own input outside the mask (the
        # reference's two _binary_morphology calls, binary.py:231-290)
        nd = x.ndim - 2
        st = np.asarray(meta["structure"]) if meta.get("structure") is not None else \
            np.asarray(tm.generate_binary_structure(nd, 1))
        org = kw.get("origin", 0)
        org = (org,) * nd if isinstance(org, int) else org
        its = kw.get("iterations", 1)
        bv = kw.get("border_value", False)
        m = np.asarray(meta["mask"])
        first = op == "closing"
        mid = oracle_mod.binary(arr["input"], st, org, bv, first, iterations=its, mask=m)
        want = oracle_mod.binary(mid, st, org, bv, not first, iterations=its, mask=m)
        np.testing.assert_array_equal(got.cpu().numpy(), want.astype(bool))
        return
    np.testing.assert_array_equal(got.cpu().numpy(), arr[op])


@pytest.mark.parametrize("shape", [(4, 2, 32, 33, 130), (2, 1, 70, 200), (3, 1, 500)])
@pytest.mark.parametrize("iterations", [1, 4, -1])
@pytest.mark.parametrize("dilation", [False, True])
def test_binary_vs_oracle(tm, oracle_mod, shape, iterations, dilation):
    rng = np.random.default_rng(len(shape) * 10 + iterations)
    x = rng.random(shape) < 0.7
    nd = len(shape) - 2
    st = np.asarray(tm.generate_binary_structure(nd, 1))
    fn = tm.binary_dilation if dilation else tm.binary_erosion
    got = fn(torch.as_tensor(x, device="cuda"), iterations=iterations)
    want = oracle_mod.binary(x, st, (0,) * nd, False, dilation, iterations=iterations)
    np.testing.assert_array_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize("shape,iterations", [((2, 1, 37, 21, 130), 4), ((1, 2, 19, 40, 64), 9),
                                              ((1, 1, 70, 131), 20), ((3, 1, 300), 3),
                                              ((1, 1, 5, 6, 7, 70), 2)])
@pytest.mark.parametrize("border_value", [False, True])
def test_binary_fused_steps_vs_oracle(tm, oracle_mod, shape, iterations, border_value):
    # several steps per launch in LDS tiles: tile and item edges, rows of
    # more than one word, reach 2 along every axis, origins, step counts above
    # one launch's share, and (4 axes) the per-step kernel
    rng = np.random.default_rng(sum(shape) + iterations)
    nd = len(shape) - 2
    x = rng.random(shape) < 0.75
    st = np.asarray(tm.iterate_structure(tm.generate_binary_structure(nd, 1), 2))
    st = st & (rng.random(st.shape) < 0.8)
    st.flat[st.size // 2] = True
    origin = tuple(int(v) for v in rng.integers(-1, 2, size=nd))
    for dilation in (False, True):
        fn = tm.binary_dilation if dilation else tm.binary_erosion
        got = fn(torch.as_tensor(x, device="cuda"), structure=torch.as_tensor(st),
                 iterations=iterations, border_value=border_value, origin=origin)
        want = oracle_mod.binary(x, st, origin, border_value, dilation, iterations=iterations)
        np.testing.assert_array_equal(got.cpu().numpy(), want, err_msg=f"dilation={dilation}")


@pytest.mark.parametrize("shape,iterations", [((2, 1, 20, 18, 128), 4), ((1, 1, 40, 192), 12),
                                              ((3, 1, 64), 1)])
@pytest.mark.parametrize("dtype", ["bool", "float32"])
def test_binary_fused_byte_io_vs_oracle(tm, oracle_mod, shape, iterations, dtype):
    # rows of whole 64-voxel words: the first tiled launch packs the bool
    # input itself and the last writes the bool output (12 steps: two
    # launches, one of each); float input keeps the pack pass
    rng = np.random.default_rng(shape[-1] + iterations)
    nd = len(shape) - 2
    x = rng.random(shape) < 0.8
    st = np.asarray(tm.generate_binary_structure(nd, nd))
    xt = torch.as_tensor(x, device="cuda")
    if dtype == "float32":
        xt = xt.float() * 2.5
    for dilation in (False, True):
        for border_value in (False, True):
            fn = tm.binary_dilation if dilation else tm.binary_erosion
            got = fn(xt, structure=torch.as_tensor(st), iterations=iterations,
                     border_value=border_value)
            want = oracle_mod.binary(x, st, (0,) * nd, border_value, dilation,
                                     iterations=iterations)
            np.testing.assert_array_equal(got.cpu().numpy(), want,
                                          err_msg=f"dilation={dilation} border={border_value}")


@pytest.mark.parametrize("shape,iterations", [((2, 1, 24, 30, 128), 3), ((1, 2, 70, 64), 10)])
@pytest.mark.parametrize("op", ["opening", "closing"])
def test_binary_opening_closing_whole_words_vs_oracle(tm, oracle_mod, shape, iterations, op):
    # the second pass reads the first pass's output buffer: with rows of whole
    # 64-voxel words it must still pack that input before writing over it
    rng = np.random.default_rng(shape[-2] + iterations)
    nd = len(shape) - 2
    x = rng.random(shape) < 0.6
    st = np.asarray(tm.generate_binary_structure(nd, 1))
    first = op == "closing"  # closing dilates first
    got = getattr(tm, f"binary_{op}")(torch.as_tensor(x, device="cuda"), iterations=iterations)
    mid = oracle_mod.binary(x, st, (0,) * nd, False, first, iterations=iterations)
    want = oracle_mod.binary(mid, st, (0,) * nd, False, not first, iterations=iterations)
    np.testing.assert_array_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize("shape,se_len,iterations", [((1, 1, 3, 4000), 3, 3), ((2, 1, 2, 4100), 3, 2),
                                                     ((1, 1, 4000), 127, 2), ((1, 1, 4000), 131, 2)])
def test_binary_long_rows_and_wide_structures(tm, oracle_mod, shape, se_len, iterations):
    # rows of 63 and 65 words (the fused steps take up to 64) and x reaches of
    # 63 and 65 voxels (the fused steps take up to 63): both sides of each limit
    rng = np.random.default_rng(se_len + shape[-1])
    nd = len(shape) - 2
    x = rng.random(shape) < 0.97
    st = np.zeros((1,) * (nd - 1) + (se_len,), bool)
    st[..., rng.random(se_len) < 0.3] = True
    st[..., 0] = st[..., -1] = st[..., se_len // 2] = True
    for dilation in (False, True):
        fn = tm.binary_dilation if dilation else tm.binary_erosion
        got = fn(torch.as_tensor(x, device="cuda"), structure=torch.as_tensor(st),
                 iterations=iterations)
        want = oracle_mod.binary(x, st, (0,) * nd, False, dilation, iterations=iterations)
        np.testing.assert_array_equal(got.cpu().numpy(), want, err_msg=f"dilation={dilation}")


def test_binary_full_structure_border_and_mask(tm, oracle_mod):
    rng = np.random.default_rng(5)
    x = rng.random((2, 1, 20, 25, 70)) < 0.6
    m = rng.random((2, 1, 20, 25, 70)) < 0.8
    st = np.ones((3, 3, 3), bool)
    got = tm.binary_erosion(torch.as_tensor(x, device="cuda"), structure=torch.as_tensor(st),
                            iterations=3, border_value=True, mask=torch.as_tensor(m, device="cuda"))
    want = oracle_mod.binary(x, st, (0, 0, 0), True, False, iterations=3, mask=m)
    np.testing.assert_array_equal(got.cpu().numpy(), want)


def test_binary_fill_holes_and_propagation(tm):
    holes = np.zeros((1, 1, 9, 9), bool)
    holes[0, 0, 1:8, 1:8] = True
    holes[0, 0, 3:5, 3:6] = False
    got = tm.binary_fill_holes(torch.as_tensor(holes, device="cuda"))
    assert got.cpu().numpy()[0, 0, 1:8, 1:8].all()
    assert not got.cpu().numpy()[0, 0, 0].any()


def test_binary_empty_structure(tm):
    x = torch.ones((1, 1, 3, 3), device="cuda")
    z = torch.zeros((3, 3), dtype=torch.bool)
    assert tm.binary_erosion(x, structure=z).all()
    assert not tm.binary_dilation(x, structure=z).any()


@pytest.mark.parametrize("mode", ["reflect", "constant", "nearest", "mirror", "wrap"])
@pytest.mark.parametrize("size,origin", [((11, 6, 13), (0, 1, -2)), ((8, 12, 10), (-1, 0, 3)),
                                         ((14, 4, 6), (2, -1, 0)), ((6, 9, 3), (0, 0, 1))])
def test_grey_box_block_scans_match_oracle(tm, oracle_mod, mode, size, origin):
    # axes of 10 or more reduce by block prefix / suffix scans; sizes, origins
    # and z depths that cut the scan blocks unevenly, with ties, infinities,
    # signed zeros and NaN in the input
    rng = np.random.default_rng(sum(size) + len(mode))
    x = (np.floor(rng.random((1, 2, 19, 21, 70)) * 8) / 4 - 1).astype(np.float32)
    flat = x.reshape(-1)
    idx = rng.choice(flat.size, size=40, replace=False)
    flat[idx[:10]] = -0.0
    flat[idx[10:20]] = np.inf
    flat[idx[20:30]] = -np.inf
    flat[idx[30:]] = np.nan
    for dilation in (False, True):
        _grey_vs_oracle(tm, oracle_mod, x, dilation, size=size, origin=origin, mode=mode,
                        cval=0.5)
