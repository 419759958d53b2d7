"""Every launch variant the autotuner can choose gives bit-identical results:
each variant is forced through the cache and compared with the oracle."""
import numpy as np
import pytest
import torch

from tests._golden import assert_bits_equal

pytestmark = pytest.mark.gpu


def _p2(v):
    p = 1
    while p < v:
        p <<= 1
    return p


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


@pytest.mark.parametrize("shape", [(8, 1, 256, 256), (2, 1, 100, 300), (1, 2, 17, 33),
                                   (3, 1, 256, 512), (1, 1, 5, 512), (2, 1, 300, 700),
                                   (1, 1, 1000, 61), (2, 1, 130, 1)])
@pytest.mark.parametrize("idx", ["i64", "i32", "none"])
def test_edt2d_every_variant_matches_oracle(tm, oracle_mod, shape, idx):
    from torchmorph_amd import autotune

    rng = np.random.default_rng(sum(shape))
    mask = rng.random(shape) < 0.9
    mask[0, 0, : shape[2] // 3, :] = True  # long foreground runs along columns
    od, oi = oracle_mod.edt(mask)
    want_d = torch.from_numpy(od.astype(np.float32))
    H, W = shape[2:]
    bits = {"i64": 64, "i32": 32, "none": 0}[idx]
    key = f"edt2d {_p2(H)} {_p2(W)} {bits} 1"
    x = torch.as_tensor(mask, device="cuda")
    try:
        for v in range(18):
            autotune.clear()
            autotune.load({key: v})
            if idx == "none":
                d = tm.euclidean_distance_transform(x)
                assert torch.equal(d.cpu(), want_d), v
            else:
                dt = torch.int64 if idx == "i64" else torch.int32
                ind = torch.empty((2, *shape), dtype=dt, device="cuda")
                d = tm.euclidean_distance_transform(x, indices=ind)
                assert torch.equal(d.cpu(), want_d), v
                assert np.array_equal(ind.cpu().numpy(), oi), v
    finally:
        autotune.clear()


def test_edt2d_variants_all_foreground_and_background(tm, oracle_mod):
    from torchmorph_amd import autotune

    try:
        for v in range(18):
            autotune.clear()
            autotune.load({"edt2d 64 128 64 1": v})
            for fill in (True, False):
                mask = np.full((2, 1, 40, 100), fill)
                mask[1, 0, 39, 99] = not fill
                x = torch.as_tensor(mask, device="cuda")
                ind = torch.empty((2, 2, 1, 40, 100), dtype=torch.int64, device="cuda")
                d = tm.euclidean_distance_transform(x, indices=ind)
                od, oi = oracle_mod.edt(mask)
                assert torch.equal(d.cpu(), torch.from_numpy(od.astype(np.float32))), (v, fill)
                assert np.array_equal(ind.cpu().numpy(), oi), (v, fill)
    finally:
        autotune.clear()


def _sparse_masks():
    """Masks whose row windows pass the step cap, so rows (or parts of rows)
    are resolved through the row envelope: one site, an all-background
    column far from most pixels, equidistant sites (ties at the midpoints),
    rows whose pixels split between the window and the envelope."""
    rng = np.random.default_rng(5)
    out = []
    m = np.ones((1, 1, 64, 1024), bool)
    m[0, 0, 30, 500] = False
    out.append(m)
    m = np.ones((1, 1, 40, 512), bool)
    m[0, 0, :, 0] = False
    out.append(m)
    m = np.ones((2, 1, 48, 800), bool)
    m[0, 0, [5, 5, 40], [100, 300, 700]] = False
    m[1, 0, [0, 47, 47], [399, 401, 600]] = False
    out.append(m)
    out.append(rng.random((2, 1, 100, 600)) > 0.002)
    m = rng.random((1, 1, 256, 304)) > 0.2
    m[0, 0, :, :200] = True  # columns without background next to dense ones
    out.append(m)
    m = np.ones((1, 1, 9, 4096), bool)
    m[0, 0, 4, [0, 4095]] = False
    out.append(m)
    m = np.ones((2, 1, 64, 256), bool)  # rows too short for the fallback: long windows
    m[0, 0, 10, 0] = False
    m[1, 0, [3, 60], [17, 255]] = False
    out.append(m)
    return out


@pytest.mark.parametrize("case", range(7))
@pytest.mark.parametrize("idx", ["i64", "none"])
def test_edt2d_sparse_rows_every_variant(tm, oracle_mod, case, idx):
    from torchmorph_amd import autotune

    mask = _sparse_masks()[case]
    od, oi = oracle_mod.edt(mask)
    want_d = torch.from_numpy(od.astype(np.float32))
    H, W = mask.shape[2:]
    bits = {"i64": 64, "none": 0}[idx]
    key = f"edt2d {_p2(H)} {_p2(W)} {bits} 1"
    x = torch.as_tensor(mask, device="cuda")
    try:
        for v in range(18):
            autotune.clear()
            autotune.load({key: v})
            if idx == "none":
                d = tm.euclidean_distance_transform(x)
                assert torch.equal(d.cpu(), want_d), v
            else:
                ind = torch.empty((2, *mask.shape), dtype=torch.int64, device="cuda")
                d = tm.euclidean_distance_transform(x, indices=ind)
                assert torch.equal(d.cpu(), want_d), v
                assert np.array_equal(ind.cpu().numpy(), oi), v
    finally:
        autotune.clear()


def test_inapplicable_cached_variant_falls_back(tm, oracle_mod):
    """A cached choice that cannot run here (variant 7 = 8-row bands needs
    more LDS than a 4096-wide row allows) falls back to the default."""
    from torchmorph_amd import autotune

    rng = np.random.default_rng(9)
    mask = rng.random((1, 1, 16, 4096)) < 0.9
    try:
        autotune.clear()
        autotune.load({"edt2d 16 4096 0 1": 7})
        d = tm.euclidean_distance_transform(torch.as_tensor(mask, device="cuda"))
        od, _ = oracle_mod.edt(mask)
        assert torch.equal(d.cpu(), torch.from_numpy(od.astype(np.float32)))
    finally:
        autotune.clear()


@pytest.mark.parametrize("se", ["cube3", "cube7", "ball2"])
@pytest.mark.parametrize("mode", ["reflect", "constant", "wrap"])
def test_grey3_every_zchunk_variant_matches_oracle(tm, oracle_mod, se, mode):
    """The streaming 3-D grey kernels give the same bits for every z-chunk and,
    for the 3 x 3 x 3 box, for every register-streaming tile."""
    from torchmorph_amd import autotune

    rng = np.random.default_rng(len(se) + len(mode))
    x = (np.floor(rng.random((2, 1, 37, 21, 70)) * 16) / 16).astype(np.float32)
    if se.startswith("cube"):
        k = int(se[4:])
        kw = dict(size=k)
        st, fp, shape = None, None, (k, k, k)
        key_kind, kk = 1, (k, k, k)
    else:
        r = np.sqrt(((np.indices((5, 5, 5)) - 2) ** 2).sum(0))
        fp = r <= 2.0
        st = np.where(fp, np.floor(2.0 - r), 0.0).astype(np.float32)
        kw = dict(structure=torch.from_numpy(st), footprint=torch.from_numpy(fp))
        shape = (5, 5, 5)
        key_kind, kk = 2, (5, 5, 5)
    want = oracle_mod.grey(x, st, fp, shape, (0, 0, 0), mode=mode, cval=0.25, dilation=True)
    xt = torch.as_tensor(x, device="cuda")
    key = "grey3 {} {} {} {} 64 32 128 2".format(key_kind, *kk)
    try:
        for v in range(8):  # 4..7: register-streaming boxes (the ball falls back)
            autotune.clear()
            autotune.load({key: v})
            y = tm.grey_dilation(xt, mode=mode, cval=0.25, **kw)
            assert np.array_equal(y.cpu().numpy(), want), v
    finally:
        autotune.clear()


def _boxr_cases():
    """(input, size, origin): 3-D and 2-D items whose K x K (x K) flat box, K =
    3, 5, 7, the register-streaming grey kernel runs: ragged tiles in x, y and
    z, items smaller than a tile or the window, shifted origins, signed zeros,
    infinities and NaN."""
    rng = np.random.default_rng(21)
    q = lambda shape: (np.floor(rng.random(shape) * 16) / 16 - 0.5).astype(np.float32)
    out = [(q((2, 1, 37, 21, 70)), 3, 0),
           (q((1, 2, 9, 40, 130)), 3, (1, -1, 0)),
           (q((1, 1, 2, 5, 3)), 3, 0),
           (q((3, 1, 50, 200)), 3, 0),
           (q((1, 1, 19, 65)), 3, (-1, 1)),
           (q((1, 1, 1, 1)), 3, 0)]
    x = q((1, 1, 6, 17, 66))
    flat = x.reshape(-1)
    idx = rng.choice(flat.size, 40, replace=False)
    flat[idx[:10]] = -0.0
    flat[idx[10:20]] = 0.0
    flat[idx[20:25]] = np.inf
    flat[idx[25:30]] = -np.inf
    flat[idx[30:]] = np.nan
    out.append((x, 3, 0))
    out += [(q((1, 1, 20, 30, 80)), 5, 0),
            (q((2, 1, 11, 9, 70)), 7, (1, -2, 3)),
            (q((2, 1, 40, 150)), 7, 0),
            (q((1, 1, 6, 5)), 5, 0)]
    return out


def _p2s(v):
    return str(_p2(v))


@pytest.mark.parametrize("case", range(11))
@pytest.mark.parametrize("mode", ["reflect", "constant", "nearest", "mirror", "wrap"])
def test_grey_register_box_every_tile_matches_oracle(tm, oracle_mod, case, mode):
    """Launch variants 4..7 (register-streaming K x K x K / K x K boxes) give
    the oracle's bits for dilation, erosion, the white tophat's fused
    subtraction and (zero origin) the one-pass gradient and laplace, for every
    border mode (NaNs compared as one canonical NaN: the payload of a NaN
    that arithmetic creates is the platform's)."""
    from torchmorph_amd import autotune

    x, size, origin = _boxr_cases()[case]
    sp = x.shape[2:]
    nd = len(sp)
    s3 = (1,) * (3 - nd) + tuple(sp)
    k3 = (1,) * (3 - nd) + (size,) * nd
    cls = [*map(str, k3), *map(_p2s, s3), _p2s(x.shape[0] * x.shape[1])]
    key, key_dual = " ".join(["grey3", "1", *cls]), " ".join(["grey3", "3", *cls])
    o = (origin,) * nd if isinstance(origin, int) else tuple(origin)
    shape = (size,) * nd
    want_d = oracle_mod.grey(x, None, None, shape, o, mode=mode, cval=0.25, dilation=True)
    want_e = oracle_mod.grey(x, None, None, shape, o, mode=mode, cval=0.25, dilation=False)
    opened = oracle_mod.grey(want_e, None, None, shape, o, mode=mode, cval=0.25, dilation=True)
    with np.errstate(invalid="ignore"):  # inf - inf in the special-values case
        want_wt = x - opened
        want_g = want_d - want_e
        want_l = (want_d + want_e) - 2 * x
    xt = torch.as_tensor(x, device="cuda")
    try:
        for v in range(4, 8):
            autotune.clear()
            autotune.load({key: v, key_dual: v})
            d = tm.grey_dilation(xt, size=size, mode=mode, cval=0.25, origin=origin)
            e = tm.grey_erosion(xt, size=size, mode=mode, cval=0.25, origin=origin)
            wt = tm.white_tophat(xt, size=size, mode=mode, cval=0.25, origin=origin)
            assert_bits_equal(d.cpu(), want_d, f"dilation variant {v}")
            assert_bits_equal(e.cpu(), want_e, f"erosion variant {v}")
            assert_bits_equal(wt.cpu(), want_wt, f"white_tophat variant {v}")
            if origin == 0:
                g = tm.morphological_gradient(xt, size=size, mode=mode, cval=0.25)
                lp = tm.morphological_laplace(xt, size=size, mode=mode, cval=0.25)
                assert_bits_equal(g.cpu(), want_g, f"gradient variant {v}")
                assert_bits_equal(lp.cpu(), want_l, f"laplace variant {v}")
    finally:
        autotune.clear()


def _ball3():
    r = np.sqrt(((np.indices((7, 7, 7)) - 3) ** 2).sum(0))
    fp = r <= 3.0
    return np.where(fp, np.floor(3.0 - r), 0.0).astype(np.float32), fp


@pytest.mark.parametrize("shape", [(2, 1, 20, 30, 70), (1, 2, 13, 7, 130), (1, 1, 5, 9, 3),
                                   (1, 1, 2, 1, 66)])
@pytest.mark.parametrize("mode", ["reflect", "constant", "nearest", "mirror", "wrap"])
def test_grey_ball3_every_variant_matches_oracle(tm, oracle_mod, shape, mode):
    """The ball of radius 3 with heights floor(3 - r) (the 3-D morphology
    benchmark's non-flat element): launch variants 4..7 (superlevel-ball
    register kernel) and the element-table kernel give the oracle's bits for
    dilation, erosion and the tophats' fused stores; NaN, infinities and
    signed zeros in the input."""
    from torchmorph_amd import autotune

    rng = np.random.default_rng(sum(shape) + len(mode))
    x = (np.floor(rng.random(shape) * 16) / 16 - 0.5).astype(np.float32)
    flat = x.reshape(-1)
    if flat.size > 40:
        idx = rng.choice(flat.size, 20, replace=False)
        flat[idx[:5]] = -0.0
        flat[idx[5:10]] = np.inf
        flat[idx[10:15]] = -np.inf
        flat[idx[15:]] = np.nan
    st, fp = _ball3()
    kw = dict(structure=torch.from_numpy(st), footprint=torch.from_numpy(fp), mode=mode,
              cval=0.25)
    want_d = oracle_mod.grey(x, st, fp, (7, 7, 7), (0, 0, 0), mode=mode, cval=0.25, dilation=True)
    want_e = oracle_mod.grey(x, st, fp, (7, 7, 7), (0, 0, 0), mode=mode, cval=0.25, dilation=False)
    with np.errstate(invalid="ignore"):
        want_wt = x - oracle_mod.grey(want_e, st, fp, (7, 7, 7), (0, 0, 0), mode=mode, cval=0.25,
                                      dilation=True)
        want_bt = oracle_mod.grey(want_d, st, fp, (7, 7, 7), (0, 0, 0), mode=mode, cval=0.25,
                                  dilation=False) - x
    key = " ".join(["grey3", "2", "7", "7", "7", *map(_p2s, shape[2:]),
                    _p2s(shape[0] * shape[1])])
    xt = torch.as_tensor(x, device="cuda")
    try:
        for v in (0, 4, 5, 6, 7):
            autotune.clear()
            autotune.load({key: v})
            assert_bits_equal(tm.grey_dilation(xt, **kw).cpu(), want_d, f"dilation {v}")
            assert_bits_equal(tm.grey_erosion(xt, **kw).cpu(), want_e, f"erosion {v}")
            assert_bits_equal(tm.white_tophat(xt, **kw).cpu(), want_wt, f"white tophat {v}")
            assert_bits_equal(tm.black_tophat(xt, **kw).cpu(), want_bt, f"black tophat {v}")
    finally:
        autotune.clear()


def _edtnd_key(shape, idx):
    """Cache key of the N-D EDT on (B, C, *spatial)."""
    sp = shape[2:]
    return " ".join(["edtnd", str(len(sp)), str(_p2(shape[0] * shape[1])), str(int(idx))]
                    + [str(_p2(s)) for s in sp])


@pytest.mark.parametrize("shape", [(2, 1, 20, 33, 40), (1, 1, 7, 130, 300), (3, 1, 50, 64, 2),
                                   (2, 1, 100, 150), (1, 1, 3, 4, 5, 6), (1, 1, 9, 1, 70)])
@pytest.mark.parametrize("idx", ["i64", "none"])
def test_edtnd_every_variant_matches_oracle(tm, oracle_mod, shape, idx):
    from torchmorph_amd import autotune

    rng = np.random.default_rng(sum(shape) + 7)
    mask = rng.random(shape) < 0.93
    samp = [1.3, 0.7, 1.9, 0.55][: len(shape) - 2]
    od, oi = oracle_mod.edt(mask, samp)
    want_d = torch.from_numpy(od.astype(np.float32))
    x = torch.as_tensor(mask, device="cuda")
    key = _edtnd_key(shape, idx != "none")
    try:
        for v in range(16):
            autotune.clear()
            autotune.load({key: v})
            if idx == "none":
                d = tm.euclidean_distance_transform(x, sampling=samp)
                assert torch.equal(d.cpu(), want_d), v
            else:
                ind = torch.empty((len(shape) - 2, *shape), dtype=torch.int64, device="cuda")
                d = tm.euclidean_distance_transform(x, sampling=samp, indices=ind)
                assert torch.equal(d.cpu(), want_d), v
                assert np.array_equal(ind.cpu().numpy(), oi), v
    finally:
        autotune.clear()


@pytest.mark.parametrize("shape", [(2, 1, 20, 33, 40), (1, 1, 7, 130, 300), (3, 1, 50, 64, 2),
                                   (2, 1, 100, 150), (1, 1, 3, 4, 5, 6), (1, 1, 9, 1, 70),
                                   (4, 1, 300), (1, 1, 600, 7)])
@pytest.mark.parametrize("metric", ["chessboard", "taxicab"])
def test_cdt_separable_every_variant_matches_oracle(tm, oracle_mod, shape, metric):
    from torchmorph_amd import autotune

    rng = np.random.default_rng(sum(shape) + 11)
    mask = rng.random(shape) < 0.9
    mask[0, 0] = True  # an item without background: -1 everywhere
    mask[-1, 0].flat[0] = False
    want = torch.from_numpy(oracle_mod.cdt(mask, metric)[0].astype(np.float32))
    sp = shape[2:]
    key = " ".join(["cdtsep", "0" if metric == "chessboard" else "1", str(len(sp)),
                    str(_p2(shape[0] * shape[1]))] + [str(_p2(s)) for s in sp])
    x = torch.as_tensor(mask, device="cuda")
    try:
        for v in range(2):
            autotune.clear()
            autotune.load({key: v})
            d = tm.chamfer_distance_transform(x, metric=metric)
            assert torch.equal(d.cpu(), want), v
    finally:
        autotune.clear()
