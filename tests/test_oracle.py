"""Pins the CPU oracle (oracle/) before it is trusted as the GPU checker.

Each oracle function is checked against (1) the SciPy golden vectors the
reference's own tests would compare with, (2) closed forms or brute force on
small inputs chosen so that a plausible slip (swapped coefficient, wrong tie
direction, dropped term) fails, and (3) SciPy itself on fresh random inputs
when SciPy is importable.
"""
import itertools

import numpy as np
import pytest

from tests import _golden

EDT = _golden.load("edt")
CDT = _golden.load("cdt")
GREY = _golden.load("grey")
BINARY = _golden.load("binary")


@pytest.mark.parametrize("case", EDT, ids=_golden.ids(EDT))
def test_edt_oracle_matches_golden(oracle_mod, case):
    meta, arr = case
    dist, idx = oracle_mod.edt(arr["input"], meta["sampling"])
    assert np.array_equal(dist, arr["dist"])
    assert np.array_equal(idx, arr["idx"])


def _brute_edt(mask, sampling):
    """Minimum distance to any background voxel (or the virtual (-1,0..))."""
    nd = mask.ndim
    s = np.asarray(sampling, np.float64)
    coords = np.stack(np.meshgrid(*[np.arange(n) for n in mask.shape], indexing="ij"), -1)
    bg = coords[~mask]
    if len(bg) == 0:
        virt = np.zeros(nd, np.int64)
        virt[0] = -1
        bg = virt[None]
    d = (((coords[..., None, :] - bg) * s) ** 2).sum(-1)
    return np.sqrt(d.min(-1)), d


@pytest.mark.parametrize("seed", range(40))
def test_edt_oracle_distance_is_brute_force_minimum(oracle_mod, seed):
    rng = np.random.default_rng(seed)
    nd = int(rng.integers(1, 4))
    shape = tuple(int(v) for v in rng.integers(1, 7, size=nd))
    mask = rng.random(shape) < rng.uniform(0.4, 0.95)
    samp = rng.uniform(0.5, 2.0, size=nd) if seed % 2 else np.ones(nd)
    dist, idx = oracle_mod.edt(mask[None, None], samp)
    want, d2 = _brute_edt(mask, samp)
    np.testing.assert_allclose(dist[0, 0], want, rtol=1e-12, atol=1e-12)
    # every index is a background voxel at exactly the reported distance
    for p in itertools.product(*[range(n) for n in shape]):
        f = idx[(slice(None), 0, 0) + p]
        if mask.all():
            assert f[0] == -1 and all(v == 0 for v in f[1:])
        else:
            assert not mask[tuple(f)]


def test_edt_oracle_closed_forms(oracle_mod):
    # 1-D single background: |x - x0| * s
    m = np.ones((1, 1, 17), bool)
    m[0, 0, 5] = False
    d, i = oracle_mod.edt(m, [0.75])
    np.testing.assert_array_equal(d[0, 0], np.abs(np.arange(17) - 5) * 0.75)
    assert (i == 5).all()
    # all background: zero distance, own indices
    m = np.zeros((1, 1, 3, 4), bool)
    d, i = oracle_mod.edt(m)
    assert (d == 0).all()
    np.testing.assert_array_equal(i[:, 0, 0], np.indices((3, 4)))
    # equidistant pair: the lower index wins (SciPy tie rule), both axes
    m = np.ones((1, 1, 1, 5), bool)
    m[0, 0, 0, [1, 3]] = False
    d, i = oracle_mod.edt(m)
    assert i[1, 0, 0, 0, 2] == 1
    m = np.ones((1, 1, 5, 1), bool)
    m[0, 0, [1, 3], 0] = False
    d, i = oracle_mod.edt(m)
    assert i[0, 0, 0, 2, 0] == 1


def test_edt_oracle_detects_planted_removal_bug(oracle_mod):
    """The golden case that caught the swapped Maurer coefficients: a
    brute-force-smaller site must win over an equal-height pair."""
    m = np.array([[0, 1, 0, 1], [1, 1, 1, 1], [1, 1, 1, 0]], bool)
    d, i = oracle_mod.edt(m[None, None])
    assert d[0, 0, 2, 1] == 2.0
    assert (i[0, 0, 0, 2, 1], i[1, 0, 0, 2, 1]) == (2, 3)


@pytest.mark.parametrize("case", CDT, ids=_golden.ids(CDT))
def test_cdt_oracle_matches_golden(oracle_mod, case):
    meta, arr = case
    dist, idx = oracle_mod.cdt(arr["input"], meta["metric"])
    assert np.array_equal(dist, arr["dist"])
    assert np.array_equal(idx, arr["idx"])


def test_cdt_oracle_closed_forms(oracle_mod):
    m = np.ones((1, 1, 64, 64), bool)
    m[0, 0, 0, 0] = False
    yy, xx = np.indices((64, 64))
    d, _ = oracle_mod.cdt(m, "chessboard")
    np.testing.assert_array_equal(d[0, 0], np.maximum(yy, xx))
    assert d[0, 0, 40, 40] == 40
    d, _ = oracle_mod.cdt(m, "taxicab")
    np.testing.assert_array_equal(d[0, 0], yy + xx)
    d, i = oracle_mod.cdt(np.ones((1, 1, 3, 4), bool))
    assert (d == -1).all()
    np.testing.assert_array_equal(i[:, 0, 0], np.indices((3, 4)))


@pytest.mark.parametrize("seed", range(4))
def test_cdt_oracle_values_are_exact_metric_distances(oracle_mod, seed):
    """SciPy's two raster passes give exactly the taxicab / chessboard
    distance to the nearest background voxel (-1 with no background).  The
    GPU's distances-only path is built on this fact, so it is pinned here by
    brute force on 1-D..3-D masks of every density."""
    rng = np.random.default_rng(100 + seed)
    for trial in range(40):
        nd = int(rng.integers(1, 4))
        shape = tuple(int(s) for s in rng.integers(1, 9 if nd == 3 else 14, size=nd))
        m = rng.random(shape) < rng.choice([0.3, 0.7, 0.95, 1.0])
        bg = np.argwhere(~m)
        pts = np.indices(shape).reshape(nd, -1).T
        for metric in ("chessboard", "taxicab"):
            d, _ = oracle_mod.cdt(m[None, None], metric)
            if len(bg) == 0:
                want = np.full(shape, -1)
            else:
                diff = np.abs(pts[:, None, :] - bg[None, :, :])
                dd = diff.sum(-1) if metric == "taxicab" else diff.max(-1)
                want = dd.min(1).reshape(shape)
            np.testing.assert_array_equal(d[0, 0], want, err_msg=f"{shape} {metric}")


def _grey_kwargs(meta, x):
    nd = x.ndim - 2
    st = meta.get("structure")
    fp = meta.get("footprint")
    if st is not None:
        se_shape = st.shape
    elif fp is not None:
        se_shape = fp.shape
    else:
        se_shape = tuple(meta["size"])
    origin = meta.get("origin", 0)
    origin = (origin,) * nd if isinstance(origin, int) else tuple(origin)
    return dict(structure=st, footprint=fp, se_shape=se_shape, origin=origin,
                mode=meta.get("mode", "reflect"), cval=meta.get("cval", 0.0))


def _grey_oracle(oracle_mod, x, kw, op):
    """Reference grey.py:108-197: opening = dilation(erosion), closing =
    erosion(dilation).  SciPy reflects structure+origin for the second step
    of opening/closing the same way its grey_dilation does."""
    if op in ("erosion", "dilation"):
        return oracle_mod.grey(x, dilation=(op == "dilation"), **kw)
    first, second = (False, True) if op == "opening" else (True, False)
    y = oracle_mod.grey(x, dilation=first, **kw)
    return oracle_mod.grey(y, dilation=second, **kw)


@pytest.mark.parametrize("case", GREY, ids=_golden.ids(GREY))
@pytest.mark.parametrize("op", ["erosion", "dilation", "opening", "closing"])
def test_grey_oracle_matches_golden(oracle_mod, case, op):
    meta, arr = case
    x = arr["input"]
    got = _grey_oracle(oracle_mod, x, _grey_kwargs(meta, x), op)
    np.testing.assert_array_equal(got, arr[op])


def _binary_kwargs(meta, x):
    nd = x.ndim - 2
    st = meta.get("structure")
    if st is None:
        from itertools import product
        st = np.zeros((3,) * nd, bool)
        for p in product(range(3), repeat=nd):
            st[p] = sum(abs(v - 1) for v in p) <= 1
    origin = meta.get("origin", 0)
    origin = (origin,) * nd if isinstance(origin, int) else tuple(origin)
    return dict(structure=st, origin=origin, border_value=bool(meta.get("border_value", False)),
                iterations=int(meta.get("iterations", 1)), mask=meta.get("mask"))


@pytest.mark.parametrize("case", BINARY, ids=_golden.ids(BINARY))
@pytest.mark.parametrize("op", ["erosion", "dilation", "opening", "closing"])
def test_binary_oracle_matches_golden(oracle_mod, case, op):
    meta, arr = case
    x = arr["input"]
    kw = _binary_kwargs(meta, x)
    if op in ("opening", "closing") and kw["mask"] is not None:
        pytest.skip("SciPy's opening/closing ignore mask; the reference applies it per step")
    if op in ("erosion", "dilation"):
        got = oracle_mod.binary(x, dilation=(op == "dilation"), **kw)
    else:
        first, second = (False, True) if op == "opening" else (True, False)
        got = oracle_mod.binary(oracle_mod.binary(x, dilation=first, **kw), dilation=second, **kw)
    np.testing.assert_array_equal(got, arr[op])


def test_sinkhorn_oracle_invariants(oracle_mod):
    """FP64 point-cloud solver: marginals to tolerance, the g-step makes the
    column marginal exact, symmetric under swapping the clouds, and the dual
    value equals the primal entropic value at convergence."""
    rng = np.random.default_rng(0)
    x = rng.normal(size=(17, 3))
    y = rng.normal(size=(13, 3)) + 0.5
    a = rng.random(17) + 0.2
    a /= a.sum()
    b = rng.random(13) + 0.2
    b /= b.sum()
    eps = 0.3
    f, g, it, err = oracle_mod.sinkhorn_points(x, y, a, b, eps, tol=1e-12)
    assert err <= 1e-12
    P = oracle_mod.plan_points(x, y, a, b, f, g, eps)
    np.testing.assert_allclose(P.sum(0), b, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(P.sum(1), a, rtol=0, atol=1e-11)
    C = oracle_mod.sqeuclid(x, y)
    primal = (P * C).sum() + eps * (P * np.log(P / (a[:, None] * b[None, :]))).sum()
    dual = (f * a).sum() + (g * b).sum()
    assert abs(primal - dual) < 1e-9
    g2, f2, _, _ = oracle_mod.sinkhorn_points(y, x, b, a, eps, tol=1e-12)
    np.testing.assert_allclose(f2 - f2.mean(), f - f.mean(), atol=1e-7)
    np.testing.assert_allclose(g2 - g2.mean(), g - g.mean(), atol=1e-7)


def test_sinkhorn_oracle_small_eps_approaches_exact_ot(oracle_mod):
    """As eps -> 0 the entropic cost tends to the exact OT cost, which for
    two uniform 1-D clouds of equal size is the sorted matching."""
    rng = np.random.default_rng(3)
    x = np.sort(rng.normal(size=(8, 1)), 0)
    y = np.sort(rng.normal(size=(8, 1)) + 1.0, 0)
    a = np.full(8, 1 / 8)
    exact = ((x[:, 0] - y[:, 0]) ** 2).mean()
    f, g, _, _ = oracle_mod.sinkhorn_points(x, y, a, a, 1e-3, tol=1e-10)
    P = oracle_mod.plan_points(x, y, a, a, f, g, 1e-3)
    cost = (P * oracle_mod.sqeuclid(x, y)).sum()
    assert abs(cost - exact) < 1e-3


def test_sinkhorn_hist_forms_agree(oracle_mod):
    """Scaling and log-domain histogram iterations (reference
    optimal_transport.py:71-125) reach the same potentials."""
    rng = np.random.default_rng(1)
    a = rng.random((3, 6)) + 0.1
    a /= a.sum(1, keepdims=True)
    b = rng.random((3, 6)) + 0.1
    b /= b.sum(1, keepdims=True)
    cost = np.abs(np.arange(6)[:, None] - np.arange(6)[None, :]).astype(np.float64)
    lu1, lv1 = oracle_mod.sinkhorn_log_hist(a, b, cost, 1.5, 2000)
    lu2, lv2 = oracle_mod.sinkhorn_scaling_hist(a, b, cost, 1.5, 2000)
    np.testing.assert_allclose(lu1 - lu1.mean(1, keepdims=True), lu2 - lu2.mean(1, keepdims=True),
                               atol=1e-9)
    P = np.exp(lu1[:, :, None] - cost[None] / 1.5 + lv1[:, None, :])
    np.testing.assert_allclose(P.sum(1), b, atol=1e-12)
    np.testing.assert_allclose(P.sum(2), a, atol=1e-9)


def test_oracle_matches_scipy_on_fresh_random_inputs(oracle_mod):
    ndi = pytest.importorskip("scipy.ndimage")
    rng = np.random.default_rng(99)
    for trial in range(60):
        nd = int(rng.integers(1, 4))
        shape = tuple(int(v) for v in rng.integers(1, 9, size=nd))
        m = rng.random(shape) < rng.uniform(0.3, 0.97)
        samp = rng.uniform(0.5, 2.0, size=nd) if trial % 2 else None
        d, i = oracle_mod.edt(m[None, None], samp)
        sd, si = ndi.distance_transform_edt(m, sampling=samp, return_indices=True)
        assert np.array_equal(d[0, 0], sd) and np.array_equal(i[:, 0, 0], si)
        for metric in ("chessboard", "taxicab"):
            d, i = oracle_mod.cdt(m[None, None], metric)
            sd, si = ndi.distance_transform_cdt(m, metric=metric, return_indices=True)
            assert np.array_equal(d[0, 0], sd) and np.array_equal(i[:, 0, 0], si)


def test_grey_backward_oracle_fd_and_mass(oracle_mod):
    """The oracle's routing rule is the derivative at tie-free points and
    conserves the upstream gradient when no cval is read."""
    rng = np.random.default_rng(5)
    x = rng.permutation(11 * 13).reshape(11, 13).astype(np.float64) / 5.0
    dy = rng.normal(size=x.shape)
    for dil in (False, True):
        dx = oracle_mod.grey_backward(x.astype(np.float32), None, None, (3, 3), (0, 0),
                                      "reflect", 0.0, dil, dy)
        assert abs(dx.sum() - dy.sum()) < 1e-9
        v = rng.normal(size=x.shape)
        h = 1e-7
        fp = oracle_mod.grey_fp64(x + h * v, None, None, (3, 3), (0, 0), "reflect", 0.0, dil)
        fm = oracle_mod.grey_fp64(x - h * v, None, None, (3, 3), (0, 0), "reflect", 0.0, dil)
        assert abs(((fp - fm) / (2 * h) * dy).sum() - (dx * v).sum()) < 1e-6
    # first element in element order wins a tie: flat 1-D size 3 dilation of a
    # constant line routes everything to the element reading p+1 (offset -d
    # of the first element d=-1)
    dx = oracle_mod.grey_backward(np.zeros((1, 5), np.float32), None, None, (1, 3), (0, 0),
                                  "nearest", 0.0, True, np.ones((1, 5)))
    np.testing.assert_array_equal(dx, [[0, 1, 1, 1, 2]])
    # the float32 forward of the oracle agrees with the float64 one on exact data
    xq = np.floor(rng.random((9, 9)) * 16).astype(np.float32)
    np.testing.assert_array_equal(
        oracle_mod.grey(xq[None, None], None, None, (3, 3), (0, 0), "mirror", 0.0, True)[0, 0],
        oracle_mod.grey_fp64(xq, None, None, (3, 3), (0, 0), "mirror", 0.0, True))


def test_deepsets_oracle_gradient_matches_central_differences():
    from oracle import deepsets as ds

    rng = np.random.default_rng(7)
    params = ds.init_params(3) * 0.7
    pairs, targets = [], []
    for n, m in [(9, 7), (5, 11)]:
        x = rng.normal(size=(n, 3))
        y = rng.normal(size=(m, 3)) + 0.5
        pairs.append((x, y, 0.3))
        targets.append(rng.normal(size=n))
    loss, grad = ds.loss_and_grad(params, pairs, targets)
    assert loss > 0
    h = 1e-6
    idx = rng.choice(ds.N_PARAMS, size=40, replace=False)
    # include biases and pair-level weights explicitly
    idx = np.concatenate([idx, [ds.N_PARAMS - 1, 4 * 128 + 3, ds.N_PARAMS - 300]])
    for i in idx:
        e = np.zeros_like(params)
        e[i] = h
        lp, _ = ds.loss_and_grad(params + e, pairs, targets)
        lm, _ = ds.loss_and_grad(params - e, pairs, targets)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, grad[i])
    # loss is invariant to a constant shift of the targets (dual gauge)
    l2, _ = ds.loss_and_grad(params, pairs, [t + 3.0 for t in targets])
    assert abs(l2 - loss) < 1e-12


# ----------------------------------------------------------- brute-force DT --
BF_METRICS = ["euclidean", "taxicab", "chessboard"]


@pytest.mark.parametrize("metric", BF_METRICS)
@pytest.mark.parametrize("seed", range(12))
def test_bfdt_oracle_matches_scipy_bf(oracle_mod, metric, seed):
    # scipy.ndimage.distance_transform_bf is the reference test's comparison
    # (reference test/test_distance_transform.py:46-66); it applies sampling
    # to the euclidean metric only, so the other metrics run unit-sampled
    ndi = pytest.importorskip("scipy.ndimage")
    rng = np.random.default_rng(300 + seed)
    nd = int(rng.integers(1, 4))
    shape = tuple(int(v) for v in rng.integers(1, 8, size=nd))
    mask = rng.random(shape) < rng.uniform(0.3, 0.95)
    mask.flat[int(rng.integers(mask.size))] = False  # at least one background voxel
    samp = rng.uniform(0.5, 2.0, size=nd) if (metric == "euclidean" and seed % 2) else None
    dist, idx = oracle_mod.bfdt(mask[None, None], metric, samp)
    want = ndi.distance_transform_bf(mask, metric=metric, sampling=samp)
    assert np.array_equal(dist[0, 0], want.astype(np.float64))
    # every index is a background voxel at the reported distance, and the
    # first such voxel in row-major order (np.argmin keeps the first minimum)
    coords = np.stack(np.meshgrid(*[np.arange(n) for n in shape], indexing="ij"), -1)
    s = np.ones(nd) if samp is None else samp
    bg = coords[~mask]
    delta = np.abs(coords[..., None, :] - bg) * s
    if metric == "euclidean":
        all_d = np.sqrt((delta ** 2).sum(-1))
    elif metric == "taxicab":
        all_d = delta.sum(-1)
    else:
        all_d = delta.max(-1)
    first = bg[np.argmin(all_d, axis=-1)]
    for p in itertools.product(*[range(n) for n in shape]):
        f = tuple(idx[(slice(None), 0, 0) + p])
        if mask[p]:
            assert f == tuple(first[p]), (p, f)
        else:
            assert f == p


def test_bfdt_oracle_special_items(oracle_mod):
    # all background: zeros and own indices; all foreground: +inf /
    # 4294967295 and index 0 (reference bfdt_kernel.cu:203-220); a single
    # background voxel: |x - x0| * s per metric, in closed form
    zeros = np.zeros((1, 1, 3, 4), bool)
    d, i = oracle_mod.bfdt(zeros, "taxicab")
    assert (d == 0).all()
    assert np.array_equal(i[:, 0, 0], np.indices((3, 4)))
    ones = np.ones((1, 1, 3, 4), bool)
    assert np.isinf(oracle_mod.bfdt(ones, "euclidean")[0]).all()
    d, i = oracle_mod.bfdt(ones, "chessboard")
    assert (d == 4294967295.0).all() and (i == 0).all()
    m = np.ones((1, 1, 5, 7), bool)
    m[0, 0, 1, 5] = False
    s = np.array([0.75, 1.5])
    yy, xx = np.indices((5, 7))
    dy, dx = np.abs(yy - 1) * s[0], np.abs(xx - 5) * s[1]
    assert np.array_equal(oracle_mod.bfdt(m, "euclidean", s)[0][0, 0], np.sqrt(dy ** 2 + dx ** 2))
    assert np.array_equal(oracle_mod.bfdt(m, "taxicab", s)[0][0, 0], dy + dx)
    assert np.array_equal(oracle_mod.bfdt(m, "chessboard", s)[0][0, 0], np.maximum(dy, dx))


@pytest.mark.parametrize("seed", range(6))
def test_bfdt_oracle_equals_edt_oracle_distances(oracle_mod, seed):
    # unit sampling: squared distances are exact integers, so the envelope
    # and the brute-force minimum give the same double
    rng = np.random.default_rng(400 + seed)
    shape = tuple(int(v) for v in rng.integers(2, 12, size=2 + seed % 2))
    mask = rng.random(shape) < 0.85
    mask.flat[0] = False
    bf, _ = oracle_mod.bfdt(mask[None, None])
    edt, _ = oracle_mod.edt(mask[None, None])
    assert np.array_equal(bf, edt)


def test_grey_signed_zero_order_independent(oracle_mod):
    """A window holding both zeros erodes to -0 and dilates to +0 whatever the
    scan order (the GPU's min / max instructions order -0 below +0); a plain
    `a < b ? a : b` scan would depend on which zero comes first."""
    pos, neg = np.float32(0.0), np.float32(-0.0)
    for row in ([pos, neg, pos, pos, neg, pos, pos], [neg, pos, pos, neg, neg, pos, neg]):
        x = np.array(row, np.float32)[None, None]
        ero = oracle_mod.grey(x, None, None, (3,), (0,), "reflect", 0.0, False)[0, 0]
        dil = oracle_mod.grey(x, None, None, (3,), (0,), "reflect", 0.0, True)[0, 0]
        padded = np.concatenate([x[0, 0, :1], x[0, 0], x[0, 0, -1:]])  # reflect: edge repeated
        for i in range(x.shape[-1]):
            window = padded[i:i + 3]
            assert ero[i] == 0 and np.signbit(ero[i]) == bool(np.signbit(window).any()), (row, i)
            assert dil[i] == 0 and not np.signbit(dil[i]), (row, i)
