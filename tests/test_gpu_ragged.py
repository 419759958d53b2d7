"""Ragged work-list distance transforms: every job equals the dense entry
point on its mask alone (bitwise) and the oracle (SciPy semantics)."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tm():
    import torchmorph_amd

    return torchmorph_amd


def _f32(a):
    return torch.from_numpy(np.asarray(a, np.float64).astype(np.float32))


def _masks(rng, shapes, p=0.9):
    return [rng.random(sh) < p for sh in shapes]


SHAPES = [(37, 53), (1, 300), (300, 1), (64, 64), (9, 10, 11), (5, 1, 7), (130, 17),
          (1,), (17,), (3, 4, 5, 6), (1, 1), (40, 3, 70)]


@pytest.mark.parametrize("schedule", ["persistent", "per_level", "per_pass"])
def test_ragged_edt_matches_oracle_and_dense(tm, oracle_mod, schedule):
    from torchmorph_amd.ragged import RaggedDistanceTransform

    rng = np.random.default_rng(1)
    masks = _masks(rng, SHAPES)
    masks[6][:] = True  # all foreground: virtual background indices
    masks[10][:] = False
    samp = [list(rng.uniform(0.5, 2.0, size=len(sh))) for sh in SHAPES]
    plan = RaggedDistanceTransform(SHAPES, ("edt",), samp, return_indices=True,
                                   schedule=schedule, device="cuda")
    (res,) = plan([torch.as_tensor(m, device="cuda") for m in masks])
    for k, m in enumerate(masks):
        d, i = res[k]
        od, oi = oracle_mod.edt(m[None, None], samp[k])
        assert torch.equal(d.cpu(), _f32(od[0, 0])), k
        assert np.array_equal(i.cpu().numpy(), oi[:, 0, 0]), k
        dd, di = tm.euclidean_distance_transform(torch.as_tensor(m[None, None], device="cuda"),
                                                 sampling=samp[k], return_indices=True)
        assert torch.equal(d, dd[0, 0]) and torch.equal(i, di[:, 0, 0].to(i.dtype)), k


@pytest.mark.parametrize("schedule", ["persistent", "per_level", "per_pass"])
@pytest.mark.parametrize("indices", [False, True])
def test_ragged_cdt_matches_oracle(tm, oracle_mod, schedule, indices):
    from torchmorph_amd.ragged import RaggedDistanceTransform

    rng = np.random.default_rng(2)
    masks = _masks(rng, SHAPES, 0.93)
    masks[4][:] = True
    plan = RaggedDistanceTransform(SHAPES, ("cdt_taxicab", "cdt_chessboard"),
                                   return_indices=indices, index_dtype=torch.int32,
                                   schedule=schedule, device="cuda")
    taxi, cheb = plan([torch.as_tensor(m, device="cuda") for m in masks])
    for metric, res in (("taxicab", taxi), ("chessboard", cheb)):
        for k, m in enumerate(masks):
            od, oi = oracle_mod.cdt(m[None, None], metric)
            d, i = res[k] if indices else (res[k], None)
            assert torch.equal(d.cpu(), torch.from_numpy(od[0, 0].astype(np.float32))), (metric, k)
            if indices:
                assert np.array_equal(i.cpu().numpy(), oi[:, 0, 0]), (metric, k)


def test_ragged_mixed_ops_float_input_and_reuse(tm, oracle_mod):
    """EDT + both chamfer metrics in one launch on float32 masks; a second run
    with new masks reuses the uploaded work-list."""
    from torchmorph_amd.ragged import RaggedDistanceTransform

    rng = np.random.default_rng(3)
    shapes = [tuple(int(s) for s in rng.integers(20, 90, size=int(rng.integers(2, 4))))
              for _ in range(24)]
    plan = RaggedDistanceTransform(shapes, ("edt", "cdt_taxicab", "cdt_chessboard"),
                                   sampling=0.75, device="cuda")
    for rep in range(2):
        masks = _masks(rng, shapes, 0.8 + 0.1 * rep)
        edt, taxi, cheb = plan([torch.as_tensor(m, device="cuda", dtype=torch.float32)
                                for m in masks])
        for k, m in enumerate(masks):
            od, _ = oracle_mod.edt(m[None, None], [0.75] * m.ndim)
            assert torch.equal(edt[k].cpu(), _f32(od[0, 0])), (rep, k)
            for metric, res in (("taxicab", taxi), ("chessboard", cheb)):
                cd, _ = oracle_mod.cdt(m[None, None], metric)
                assert torch.equal(res[k].cpu(), torch.from_numpy(cd[0, 0].astype(np.float32)))


def test_ragged_large_items_against_dense(tm):
    """C3-like sizes (lines up to 1024, a 3-D volume) against the dense path."""
    from torchmorph_amd.ragged import RaggedDistanceTransform

    g = torch.Generator(device="cuda").manual_seed(5)
    shapes = [(1024, 700), (33, 1024), (96, 80, 64), (256, 31), (48, 48, 48)]
    masks = [torch.rand(sh, generator=g, device="cuda") < 0.97 for sh in shapes]
    samp = [[1.5, 0.6], [0.9, 1.1], [1.0, 2.0, 0.5], [1.0, 1.0], [0.7, 0.7, 1.3]]
    plan = RaggedDistanceTransform(shapes, ("edt", "cdt_taxicab", "cdt_chessboard"), samp,
                                   device="cuda")
    edt, taxi, cheb = plan(masks)
    for k, m in enumerate(masks):
        x = m[None, None]
        assert torch.equal(edt[k], tm.euclidean_distance_transform(x, sampling=samp[k])[0, 0])
        assert torch.equal(taxi[k], tm.chamfer_distance_transform(x, metric="taxicab")[0, 0])
        assert torch.equal(cheb[k], tm.chamfer_distance_transform(x, metric="chessboard")[0, 0])


def test_ragged_convenience_functions(tm, oracle_mod):
    from torchmorph_amd.ragged import (chamfer_distance_transform_ragged,
                                       euclidean_distance_transform_ragged)

    rng = np.random.default_rng(6)
    masks = _masks(rng, [(20, 30), (7, 8, 9), (50,)], 0.85)
    xs = [torch.as_tensor(m, device="cuda") for m in masks]
    d = euclidean_distance_transform_ragged(xs, sampling=[[1.0, 2.0], [1, 1, 1], [0.5]])
    c = chamfer_distance_transform_ragged(xs, metric="cityblock")
    for k, m in enumerate(masks):
        s = [[1.0, 2.0], [1, 1, 1], [0.5]][k]
        assert torch.equal(d[k].cpu(), _f32(oracle_mod.edt(m[None, None], s)[0][0, 0]))
        assert torch.equal(c[k].cpu(), torch.from_numpy(
            oracle_mod.cdt(m[None, None], "taxicab")[0][0, 0].astype(np.float32)))


def test_ragged_argument_errors(tm):
    from torchmorph_amd.ragged import RaggedDistanceTransform

    with pytest.raises(ValueError):
        RaggedDistanceTransform([(4, 70000)], device="cuda")
    with pytest.raises(ValueError):
        RaggedDistanceTransform([(4, 4)], ("sobel",), device="cuda")
    with pytest.raises(ValueError):
        RaggedDistanceTransform([(4, 4)], sampling=[[1.0, -1.0]], device="cuda")
    plan = RaggedDistanceTransform([(4, 4), (3, 3)], device="cuda")
    with pytest.raises(ValueError):
        plan([torch.zeros(4, 4, device="cuda")])
    with pytest.raises(ValueError):
        plan.run(torch.zeros(10, dtype=torch.uint8, device="cuda"))


def _mixed_case(seed):
    rng = np.random.default_rng(seed)
    shapes = [tuple(int(s) for s in rng.integers(20, 300, size=2)) for _ in range(20)]
    shapes += [tuple(int(s) for s in rng.integers(8, 60, size=3)) for _ in range(6)]
    masks = _masks(rng, shapes, 0.85)
    samp = [list(rng.uniform(0.5, 2.0, size=len(sh))) for sh in shapes]
    return shapes, masks, samp


def _check_mixed(oracle_mod, masks, samp, edt, taxi, cheb, items):
    for k in items:
        m = masks[k][None, None]
        od, _ = oracle_mod.edt(m, samp[k])
        assert torch.equal(edt[k].cpu(), _f32(od[0, 0])), ("edt", k)
        for metric, res in (("taxicab", taxi), ("chessboard", cheb)):
            oc, _ = oracle_mod.cdt(m, metric)
            assert torch.equal(res[k].cpu(), _f32(oc[0, 0])), (metric, k)


def test_ragged_graph_replays_match_eager(tm, oracle_mod):
    """A per-pass plan captured as a HIP graph: every replay, with eager work
    and synchronisations between replays, equals the eager run bit for bit;
    new masks copied into the plan's input buffer replay the same graph."""
    from torchmorph_amd.ragged import RaggedDistanceTransform

    shapes, masks, samp = _mixed_case(11)
    ops = ("edt", "cdt_taxicab", "cdt_chessboard")
    eager = RaggedDistanceTransform(shapes, ops, samp, device="cuda")
    want = [[t.clone() for t in r] for r in eager([torch.as_tensor(m, device="cuda") for m in masks])]
    _check_mixed(oracle_mod, masks, samp, *want, items=range(len(shapes)))
    plan = RaggedDistanceTransform(shapes, ops, samp, graph=True, device="cuda")
    dm = [torch.as_tensor(m, device="cuda") for m in masks]
    for rep in range(4):
        plan.distances.fill_(-7.0)  # stale outputs must be overwritten by the replay
        got = plan(dm)
        for o in range(3):
            for k in range(len(shapes)):
                assert torch.equal(got[o][k], want[o][k]), (rep, o, k)
        # eager work between replays
        tm.euclidean_distance_transform(dm[0][None, None])
        torch.cuda.synchronize()
    assert len(plan._graphs) == 1
    # new masks through the same graph
    rng = np.random.default_rng(12)
    masks2 = _masks(rng, shapes, 0.8)
    got2 = plan([torch.as_tensor(m, device="cuda") for m in masks2])
    assert len(plan._graphs) == 1
    _check_mixed(oracle_mod, masks2, samp, *got2, items=(0, 5, 19, 20, 25))


def test_ragged_every_row_chunk_variant(tm, oracle_mod):
    """Each chunk depth the tuner may pick for the row-staged passes gives the
    same bits (forced through the cache)."""
    from torchmorph_amd import autotune
    from torchmorph_amd.ragged import RaggedDistanceTransform

    shapes, masks, samp = _mixed_case(13)
    ops = ("edt", "cdt_taxicab", "cdt_chessboard")
    dm = [torch.as_tensor(m, device="cuda") for m in masks]
    try:
        autotune.clear()
        plan = RaggedDistanceTransform(shapes, ops, samp, device="cuda")
        ref = [[t.clone() for t in r] for r in plan(dm)]
        _check_mixed(oracle_mod, masks, samp, *ref, items=range(len(shapes)))
        keys = [k for k in autotune.export() if k.startswith("ragged_rows")]
        assert len(keys) == 1, keys
        for v in range(5):
            autotune.clear()
            autotune.load({keys[0]: v})
            plan.distances.fill_(-7.0)
            got = plan(dm)
            for o in range(3):
                for k in range(len(shapes)):
                    assert torch.equal(got[o][k], ref[o][k]), (v, o, k)
    finally:
        autotune.clear()
