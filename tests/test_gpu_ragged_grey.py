"""Ragged grey morphology: every item of one launch equals the dense
grey_erosion / grey_dilation / opening / closing of that item alone (bitwise)
and the oracle."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SHAPES = {
    1: [(1,), (17,), (300,), (64,), (65,)],
    2: [(37, 53), (1, 300), (300, 1), (64, 64), (130, 17), (1, 1), (16, 64), (17, 65)],
    3: [(20, 33, 70), (1, 1, 1), (17, 16, 64), (33, 5, 129), (2, 40, 3), (16, 17, 65),
        (40, 20, 20)],
}


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _items(rng, shapes):
    # quantised values: many ties, exact in float32
    return [(np.floor(rng.random(sh) * 64) / 64).astype(np.float32) for sh in shapes]


def _ball(nd, r):
    g = np.indices((2 * r + 1,) * nd) - r
    return (g ** 2).sum(0) <= r * r


def _element(name, nd, rng):
    if name == "box3":
        return dict(size=3)
    if name == "box_uneven_origin":
        return dict(size=(3, 4, 5)[-nd:], origin=(1, -2, 2)[-nd:])
    if name == "ball_footprint":
        return dict(footprint=torch.as_tensor(_ball(nd, 2)))
    if name == "heights":
        st = rng.integers(0, 4, size=(3,) * nd).astype(np.float32)
        return dict(structure=torch.as_tensor(st), footprint=torch.as_tensor(_ball(nd, 1)))
    if name == "tall_box":  # more halo rows than the small register ring holds
        return dict(size=(5, 15, 9)[-nd:])
    raise AssertionError(name)


def _same_bits(a, b):
    return a.shape == b.shape and torch.equal(a.contiguous().view(torch.int32),
                                              b.contiguous().view(torch.int32))


ELEMENTS = ["box3", "box_uneven_origin", "ball_footprint", "heights", "tall_box"]


@pytest.mark.parametrize("operation", ["erosion", "dilation", "opening", "closing"])
@pytest.mark.parametrize("elem", ELEMENTS)
@pytest.mark.parametrize("nd", [1, 2, 3])
def test_ragged_grey_equals_dense_per_item(tm, nd, elem, operation):
    rng = np.random.default_rng(100 * nd + ELEMENTS.index(elem))
    items = _items(rng, SHAPES[nd])
    kw = _element(elem, nd, rng)
    mode = "constant" if operation in ("dilation", "closing") else "reflect"
    res = getattr(tm, f"grey_{operation}_ragged")(
        [torch.as_tensor(x, device="cuda") for x in items], mode=mode, cval=0.375, **kw)
    assert len(res) == len(items)
    for k, x in enumerate(items):
        dense = getattr(tm, f"grey_{operation}")(torch.as_tensor(x[None, None], device="cuda"),
                                                  mode=mode, cval=0.375, **kw)[0, 0]
        assert _same_bits(res[k], dense), (k, x.shape)


@pytest.mark.parametrize("mode", ["constant", "reflect", "nearest", "mirror", "wrap"])
@pytest.mark.parametrize("dilation", [False, True])
@pytest.mark.parametrize("nd", [2, 3])
def test_ragged_grey_matches_oracle_every_mode(tm, oracle_mod, nd, mode, dilation):
    from torchmorph_amd.ragged import RaggedGreyMorphology

    rng = np.random.default_rng(7 + nd)
    items = _items(rng, SHAPES[nd])
    st = rng.integers(0, 3, size=(3, 4, 5)[-nd:]).astype(np.float32)
    fp = rng.random(st.shape) < 0.7
    fp.flat[0] = True
    origin = (1, -2, 2)[-nd:]
    plan = RaggedGreyMorphology([x.shape for x in items], "dilation" if dilation else "erosion",
                                structure=torch.as_tensor(st), footprint=torch.as_tensor(fp),
                                mode=mode, cval=-0.5, origin=origin, device="cuda")
    res = plan([torch.as_tensor(x, device="cuda") for x in items])
    for k, x in enumerate(items):
        want = oracle_mod.grey(x[None, None], st, fp, st.shape, origin, mode, -0.5,
                               dilation)[0, 0]
        assert np.array_equal(res[k].cpu().numpy().view(np.int32), want.view(np.int32)), k


def test_ragged_grey_plan_reuse_and_packed_input(tm):
    from torchmorph_amd.ragged import RaggedGreyMorphology

    rng = np.random.default_rng(3)
    shapes = SHAPES[3]
    plan = RaggedGreyMorphology(shapes, "closing", size=(3, 5, 3), device="cuda")
    for rep in range(3):  # descriptors are uploaded once and reused
        items = [torch.as_tensor(x, device="cuda") for x in _items(rng, shapes)]
        flat = plan.pack(items)
        res = plan(flat)
        for k, x in enumerate(items):
            dense = tm.grey_closing(x[None, None], size=(3, 5, 3))[0, 0]
            assert _same_bits(res[k], dense), (rep, k)
    # a second plan in between re-stages its own descriptors
    other = RaggedGreyMorphology(SHAPES[2], "erosion", size=3, device="cuda")
    o = other([torch.ones(sh, device="cuda") for sh in SHAPES[2]])
    assert all(torch.equal(v, torch.ones_like(v)) for v in o)
    res = plan(flat)
    assert _same_bits(res[0], tm.grey_closing(items[0][None, None], size=(3, 5, 3))[0, 0])


def test_ragged_grey_special_values(tm):
    # signed zeros, infinities and NaN travel as the dense kernels carry them
    rng = np.random.default_rng(5)
    items = _items(rng, SHAPES[2])
    for x in items:
        flat = x.reshape(-1)
        idx = rng.choice(flat.size, size=min(flat.size, 7), replace=False)
        vals = np.array([-0.0, 0.0, np.inf, -np.inf, np.nan, -0.0, 1.0], np.float32)
        flat[idx] = vals[: len(idx)]
    for op in ("erosion", "dilation"):
        res = getattr(tm, f"grey_{op}_ragged")([torch.as_tensor(x, device="cuda") for x in items],
                                               size=3)
        for k, x in enumerate(items):
            dense = getattr(tm, f"grey_{op}")(torch.as_tensor(x[None, None], device="cuda"),
                                              size=3)[0, 0]
            assert _same_bits(res[k], dense), (op, k)


def test_ragged_grey_many_small_items(tm):
    # more items than tiles of any one item: block -> item table at scale
    rng = np.random.default_rng(9)
    shapes = [tuple(int(v) for v in rng.integers(1, 40, size=3)) for _ in range(300)]
    items = _items(rng, shapes)
    res = tm.grey_dilation_ragged([torch.as_tensor(x, device="cuda") for x in items],
                                  footprint=torch.as_tensor(_ball(3, 1)))
    for k in range(0, len(items), 13):
        dense = tm.grey_dilation(torch.as_tensor(items[k][None, None], device="cuda"),
                                 footprint=torch.as_tensor(_ball(3, 1)))[0, 0]
        assert _same_bits(res[k], dense), k


def test_ragged_grey_rejects_oversized_element(tm):
    from torchmorph_amd.ragged import RaggedGreyMorphology

    # 300 active elements exceed the table kernel; a non-flat element is no box
    st = torch.zeros((1, 20, 15))
    st[0, 0, 0] = 1.0
    plan = RaggedGreyMorphology([(4, 30, 30)], "erosion", structure=st, device="cuda")
    with pytest.raises(RuntimeError, match="fits neither streaming kernel"):
        plan([torch.zeros((4, 30, 30), device="cuda")])


@pytest.mark.parametrize("operation", ["erosion", "dilation", "opening", "closing"])
@pytest.mark.parametrize("nd", [1, 2, 3])
@pytest.mark.parametrize("iterations", [1, 3])
def test_ragged_binary_equals_dense_per_item(tm, nd, operation, iterations):
    rng = np.random.default_rng(10 * nd + iterations)
    items = [rng.random(sh) < 0.7 for sh in SHAPES[nd]]
    full = torch.ones((3,) * nd, dtype=torch.bool)
    for structure, border, origin in ((None, False, 0), (full, True, 0),
                                      (full, False, (1, -1, 0)[-nd:])):
        res = getattr(tm, f"binary_{operation}_ragged")(
            [torch.as_tensor(x, device="cuda") for x in items], structure=structure,
            iterations=iterations, border_value=border, origin=origin)
        for k, x in enumerate(items):
            dense = getattr(tm, f"binary_{operation}")(
                torch.as_tensor(x[None, None], device="cuda"), structure=structure,
                iterations=iterations, border_value=border, origin=origin)[0, 0]
            assert res[k].dtype == torch.bool
            assert torch.equal(res[k], dense), (k, x.shape, border, origin)


def test_ragged_binary_empty_structure_and_limits(tm):
    from torchmorph_amd.ragged import RaggedBinaryMorphology

    items = [torch.rand((5, 7), device="cuda") < 0.5, torch.rand((3, 9), device="cuda") < 0.5]
    empty = torch.zeros((3, 3), dtype=torch.bool)
    for op in ("erosion", "dilation", "opening", "closing"):
        res = getattr(tm, f"binary_{op}_ragged")(items, structure=empty)
        for k, x in enumerate(items):
            dense = getattr(tm, f"binary_{op}")(x[None, None], structure=empty)[0, 0]
            assert torch.equal(res[k], dense), (op, k)
    with pytest.raises(ValueError, match="iterations >= 1"):
        RaggedBinaryMorphology([(4, 4)], iterations=0, device="cuda")
