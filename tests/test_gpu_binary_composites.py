"""Binary composites and large structures on the GPU.

* binary_propagation / binary_fill_holes / binary_hit_or_miss against
  SciPy's golden vectors (reference test/test_binary.py:190-367 cases plus
  seeded random blobs with holes);
* the same composites at larger sizes against the oracle (compositions of
  oracle.binary, itself pinned to SciPy);
* structures beyond the fused step tables (a full 7^3 element: 343 active
  elements; a 5-D full-connectivity element: 243; x reaches beyond +-127),
  which run as chunks of 64 elements per step;
* masked opening / closing: both passes chained through the oracle with the
  mask re-applied against each pass's own input.
"""
import numpy as np
import pytest
import torch

from tests import _golden

pytestmark = pytest.mark.gpu

COMPOSITES = _golden.load("composite")


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _origin(v):
    return v if isinstance(v, int) else tuple(v)


# The reference's hit-or-miss erodes the complement with border_value False
# (reference morphology/binary.py:216-220); SciPy's erodes it as if the
# outside were background (border 1 for the complement).  The two differ only
# where structure2 reaches past the item; in these cases the drop-in follows
# the reference, checked against the oracle, and the difference from SciPy is
# checked to lie in that border band.
REFERENCE_BORDER_CASES = {"hm_rnd_2d_pair", "hm_rnd_2d_origin1_only", "hm_rnd_2d_empty_hit"}


def _reference_hit_or_miss(tm, oracle_mod, x, meta):
    nd = x.ndim - 2
    s1 = meta.get("structure1")
    s1 = np.asarray(tm.generate_binary_structure(nd, 1)) if s1 is None else np.asarray(s1)
    s2 = ~s1 if meta.get("structure2") is None else np.asarray(meta["structure2"])
    o1 = meta.get("origin1", 0)
    o1 = (o1,) * nd if isinstance(o1, int) else tuple(o1)
    o2 = meta.get("origin2")
    o2 = o1 if o2 is None else ((o2,) * nd if isinstance(o2, int) else tuple(o2))
    hit = oracle_mod.binary(x, s1, o1, False, False).astype(bool) if s1.any() else \
        np.ones(x.shape, bool)
    miss = oracle_mod.binary(~x, s2, o2, False, False).astype(bool) if s2.any() else \
        np.ones(x.shape, bool)
    reach = max(max(s1.shape), max(s2.shape)) // 2 + 1
    return hit & miss, reach


@pytest.mark.parametrize("case", COMPOSITES, ids=[c[0]["name"] for c in COMPOSITES])
def test_composites_match_golden(tm, oracle_mod, case):
    meta, arr = case
    x = torch.as_tensor(arr["input"], dtype=torch.float32, device="cuda")
    kw = {}
    for key in ("structure", "structure1", "structure2"):
        if meta.get(key) is not None:
            kw[key] = torch.as_tensor(meta[key], dtype=torch.bool, device="cuda")
    if meta.get("mask") is not None:
        kw["mask"] = torch.as_tensor(meta["mask"], dtype=torch.bool,
                                     device="cuda").expand(x.shape)
    for key in ("origin", "origin1", "origin2"):
        if key in meta:
            kw[key] = _origin(meta[key])
    if "border_value" in meta:
        kw["border_value"] = meta["border_value"]
    got = getattr(tm, f"binary_{meta['op']}")(x, **kw)
    assert got.dtype == torch.bool and got.shape == x.shape
    if meta["name"] in REFERENCE_BORDER_CASES:
        want, reach = _reference_hit_or_miss(tm, oracle_mod, arr["input"], meta)
        np.testing.assert_array_equal(got.cpu().numpy(), want)
        diff = want != arr["expected"]
        assert diff.any()  # the case does exercise the border rule
        inner = np.zeros_like(diff)
        inner[(slice(None), slice(None)) + (slice(reach, -reach),) * (x.ndim - 2)] = True
        assert not (diff & inner).any()
    else:
        np.testing.assert_array_equal(got.cpu().numpy(), arr["expected"])
    out = torch.empty_like(got)
    assert getattr(tm, f"binary_{meta['op']}")(x, output=out, **kw) is out
    assert torch.equal(out, got)


def _fill_holes_oracle(oracle_mod, x, st, origin):
    seed = np.zeros_like(x, dtype=bool)
    reached = oracle_mod.binary(seed, st, origin, True, True, iterations=-1, mask=(x == 0))
    return ~reached.astype(bool)


@pytest.mark.parametrize("shape,conn", [((2, 1, 190, 257), 1), ((1, 2, 40, 52, 130), 1),
                                        ((1, 1, 33, 40, 64), 3)])
def test_fill_holes_large_vs_oracle(tm, oracle_mod, shape, conn):
    rng = np.random.default_rng(sum(shape) + conn)
    x = rng.random(shape) < 0.6
    nd = len(shape) - 2
    st = np.asarray(tm.generate_binary_structure(nd, conn))
    got = tm.binary_fill_holes(torch.as_tensor(x, device="cuda"),
                               structure=torch.as_tensor(st))
    want = _fill_holes_oracle(oracle_mod, x, st, (0,) * nd)
    np.testing.assert_array_equal(got.cpu().numpy(), want)


@pytest.mark.parametrize("shape", [(2, 1, 130, 67), (1, 2, 20, 33, 129)])
def test_hit_or_miss_large_vs_oracle(tm, oracle_mod, shape):
    rng = np.random.default_rng(shape[-1])
    x = rng.random(shape) < 0.5
    nd = len(shape) - 2
    s1 = rng.random((3,) * nd) < 0.4
    s2 = (rng.random((3,) * nd) < 0.4) & ~s1
    s1.flat[s1.size // 2] = True
    o1 = tuple(int(v) for v in rng.integers(-1, 2, size=nd))
    o2 = tuple(int(v) for v in rng.integers(-1, 2, size=nd))
    got = tm.binary_hit_or_miss(torch.as_tensor(x, device="cuda"),
                                structure1=torch.as_tensor(s1), structure2=torch.as_tensor(s2),
                                origin1=o1, origin2=o2)
    hit = oracle_mod.binary(x, s1, o1, False, False)
    miss = oracle_mod.binary(~x, s2, o2, False, False)
    np.testing.assert_array_equal(got.cpu().numpy(), hit.astype(bool) & miss.astype(bool))


@pytest.mark.parametrize("case", ["full7_3d", "full_5d", "long_reach_1d", "full5_3d_until"])
def test_structures_beyond_fused_tables_vs_oracle(tm, oracle_mod, case):
    rng = np.random.default_rng(len(case))
    mask = None
    iterations = 2
    if case == "full7_3d":          # 343 active elements
        x = rng.random((2, 1, 20, 23, 70)) < 0.97
        st = np.ones((7, 7, 7), bool)
    elif case == "full_5d":         # 3^5 = 243 active elements
        x = rng.random((1, 2, 5, 6, 4, 7, 9)) < 0.9
        st = np.asarray(tm.generate_binary_structure(5, 5))
    elif case == "long_reach_1d":   # offsets up to +-150
        x = rng.random((2, 1, 700)) < 0.99
        st = np.zeros(301, bool)
        st[[0, 37, 150, 151, 299, 300]] = True
    else:                            # 150 elements, with a mask, until stable
        x = rng.random((1, 1, 16, 18, 40)) < 0.05
        st = np.ones((5, 5, 6), bool)    # centre (2, 2, 3) is active: stable in finite steps
        mask = rng.random(x.shape) < 0.5
        iterations = -1
    nd = x.ndim - 2
    assert st.sum() > 128 or st.shape[-1] > 255  # beyond the fused step tables
    xt = torch.as_tensor(x, device="cuda")
    mt = None if mask is None else torch.as_tensor(mask, device="cuda")
    for dilation in (False, True):
        fn = tm.binary_dilation if dilation else tm.binary_erosion
        for border_value in (False, True):
            got = fn(xt, structure=torch.as_tensor(st), iterations=iterations,
                     border_value=border_value, mask=mt)
            want = oracle_mod.binary(x, st, (0,) * nd, border_value, dilation,
                                     iterations=iterations, mask=mask)
            np.testing.assert_array_equal(got.cpu().numpy(), want,
                                          err_msg=f"{case} dilation={dilation} "
                                                  f"border={border_value}")


def test_wide_structure_opening_and_hit_or_miss(tm, oracle_mod):
    rng = np.random.default_rng(77)
    x = rng.random((1, 1, 30, 40, 66)) < 0.9
    st = np.ones((3, 7, 9), bool)  # 189 elements
    got = tm.binary_opening(torch.as_tensor(x, device="cuda"), structure=torch.as_tensor(st))
    mid = oracle_mod.binary(x, st, (0, 0, 0), False, False)
    want = oracle_mod.binary(mid, st, (0, 0, 0), False, True)
    np.testing.assert_array_equal(got.cpu().numpy(), want)
    s2 = np.zeros((9, 9, 9), bool)
    s2[::2, ::2, ::2] = True  # 125 elements, not the complement
    s2[4, 4, 4] = False
    hm = tm.binary_hit_or_miss(torch.as_tensor(x, device="cuda"), structure1=torch.as_tensor(st),
                               structure2=torch.as_tensor(s2))
    hit = oracle_mod.binary(x, st, (0, 0, 0), False, False)
    miss = oracle_mod.binary(~x, s2, (0, 0, 0), False, False)
    np.testing.assert_array_equal(hm.cpu().numpy(), hit.astype(bool) & miss.astype(bool))


@pytest.mark.parametrize("op", ["opening", "closing"])
@pytest.mark.parametrize("shape,iterations", [((2, 1, 19, 23, 70), 2), ((1, 2, 64, 128), 3),
                                              ((1, 1, 3, 3), 1)])
def test_masked_opening_closing_vs_chained_oracle(tm, oracle_mod, op, shape, iterations):
    """Each pass restores, outside the mask, its own input (the first pass's
    output for the second pass), as the reference's two _binary_morphology
    calls do (binary.py:231-290)."""
    rng = np.random.default_rng(shape[-1] + iterations + len(op))
    x = rng.random(shape) < 0.6
    m = rng.random(shape) < 0.7
    nd = len(shape) - 2
    st = np.asarray(tm.generate_binary_structure(nd, 1))
    first = op == "closing"  # closing dilates first
    got = getattr(tm, f"binary_{op}")(torch.as_tensor(x, device="cuda"), iterations=iterations,
                                      mask=torch.as_tensor(m, device="cuda"))
    mid = oracle_mod.binary(x, st, (0,) * nd, False, first, iterations=iterations, mask=m)
    want = oracle_mod.binary(mid, st, (0,) * nd, False, not first, iterations=iterations, mask=m)
    np.testing.assert_array_equal(got.cpu().numpy(), want)
