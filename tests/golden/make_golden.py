"""Writes the SciPy golden vectors under tests/golden/ (run in a container that
has SciPy; the fixtures are committed so that the GPU tests need no SciPy).

SciPy's scipy.ndimage is the oracle the reference tests use
(reference test/test_distance_transform.py:4, test/test_grey.py:4-12,
test/test_binary.py:5-13).  Inputs are the reference tests' own cases
(checkerboards test_distance_transform.py:13-24, CASE_2D/CASE_3D
test_grey.py:42-47, binary cases test_binary.py:60-88, seed-42 high-dim
inputs test_distance_transform.py:302-320) plus seeded random cases that
exercise ties, anisotropic sampling, border modes and origins.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os

import numpy as np
from scipy import ndimage as ndi

HERE = os.path.dirname(os.path.abspath(__file__))


def make_case(shape):
    values = np.ones(shape, dtype=np.float32)
    slices = [slice(None, None, 2) for _ in shape[2:]]
    values[(slice(None), slice(None), *slices)] = 0
    return values


def per_item(x, fn):
    lead = x.shape[:2]
    items = x.reshape((-1,) + x.shape[2:])
    return [fn(it) for it in items], lead


def edt_cases():
    rng = np.random.default_rng(20260101)
    cases = []
    for nd, shp in [(1, (2, 1, 7)), (2, (2, 1, 5, 6)), (3, (1, 2, 4, 5, 6))]:
        cases.append((f"checkerboard_{nd}d", make_case(shp), None))
    cases.append(("random_2d", (rng.random((2, 2, 33, 41)) < 0.8).astype(np.float32), None))
    cases.append(("random_2d_sparse", (rng.random((1, 2, 40, 37)) < 0.995).astype(np.float32), None))
    cases.append(("random_2d_aniso", (rng.random((2, 1, 30, 26)) < 0.85).astype(np.float32),
                  [0.5, 2.0]))
    cases.append(("random_2d_aniso_odd", (rng.random((1, 2, 19, 23)) < 0.9).astype(np.float32),
                  [1.37, 0.61]))
    cases.append(("random_3d", (rng.random((1, 2, 12, 10, 14)) < 0.85).astype(np.float32), None))
    cases.append(("random_3d_aniso", (rng.random((2, 1, 9, 13, 11)) < 0.8).astype(np.float32),
                  [1.3, 0.7, 1.9]))
    cases.append(("random_1d", (rng.random((3, 1, 50)) < 0.8).astype(np.float32), None))
    lattice = np.ones((1, 1, 24, 24), np.float32)
    lattice[0, 0, ::4, ::4] = 0
    cases.append(("lattice_ties_2d", lattice, None))
    lat3 = np.ones((1, 1, 9, 9, 9), np.float32)
    lat3[0, 0, ::4, ::4, ::4] = 0
    cases.append(("lattice_ties_3d", lat3, None))
    for v in (0.0, 1.0):
        cases.append((f"uniform_{int(v)}", np.full((2, 2, 3, 4), v, np.float32), None))
    for nd in (1, 2, 4):
        cases.append((f"all_fg_sampling_{nd}d", np.ones((2, 2, *([3] * nd)), np.float32),
                      [0.5 + d for d in range(nd)]))
    r42 = np.random.default_rng(42)
    for nd in (4, 5, 6, 7, 8):
        cases.append((f"highdim_{nd}d", r42.integers(0, 2, size=(2, 2, *([2] * nd))).astype(
            np.float32), None))
    one = np.ones((1, 1, 64, 64), np.float32)
    one[0, 0, 0, 0] = 0
    cases.append(("single_bg_corner_64", one, None))
    cases.append(("c1_like_256", (np.random.default_rng(7).random((1, 1, 256, 256)) < 0.9).astype(
        np.float32), None))
    out = {}
    meta = []
    for i, (name, x, samp) in enumerate(cases):
        res, lead = per_item(x, lambda it: ndi.distance_transform_edt(it, sampling=samp,
                                                                      return_indices=True))
        dist = np.stack([r[0] for r in res]).reshape(x.shape)
        idx = np.stack([r[1] for r in res], axis=1).reshape((x.ndim - 2,) + x.shape)
        out[f"{i}_input"] = x.astype(bool)
        out[f"{i}_dist"] = dist
        out[f"{i}_idx"] = idx.astype(np.int32)
        meta.append({"name": name, "sampling": samp})
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, "edt_golden.npz"), **out)


def cdt_cases():
    rng = np.random.default_rng(20260102)
    cases = []
    for nd, shp in [(1, (2, 1, 7)), (2, (2, 1, 5, 6)), (3, (1, 2, 4, 5, 6))]:
        cases.append((f"checkerboard_{nd}d", make_case(shp)))
    cases.append(("random_2d", (rng.random((2, 2, 31, 29)) < 0.85).astype(np.float32)))
    cases.append(("random_2d_sparse", (rng.random((1, 1, 48, 52)) < 0.998).astype(np.float32)))
    cases.append(("random_3d", (rng.random((1, 2, 11, 9, 13)) < 0.9).astype(np.float32)))
    cases.append(("random_1d", (rng.random((2, 1, 40)) < 0.8).astype(np.float32)))
    for v in (0.0, 1.0):
        cases.append((f"uniform_{int(v)}", np.full((2, 2, 3, 4), v, np.float32)))
    r42 = np.random.default_rng(42)
    for nd in (4, 5):
        cases.append((f"highdim_{nd}d", r42.integers(0, 2, size=(2, 2, *([2] * nd))).astype(
            np.float32)))
    one = np.ones((1, 1, 64, 64), np.float32)
    one[0, 0, 0, 0] = 0
    cases.append(("single_bg_corner_64", one))
    out = {}
    meta = []
    k = 0
    for name, x in cases:
        for metric in ("chessboard", "taxicab"):
            res, lead = per_item(x, lambda it: ndi.distance_transform_cdt(it, metric=metric,
                                                                          return_indices=True))
            dist = np.stack([r[0] for r in res]).reshape(x.shape)
            idx = np.stack([r[1] for r in res], axis=1).reshape((x.ndim - 2,) + x.shape)
            out[f"{k}_input"] = x.astype(bool)
            out[f"{k}_dist"] = dist.astype(np.int32)
            out[f"{k}_idx"] = idx.astype(np.int32)
            meta.append({"name": name, "metric": metric})
            k += 1
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, "cdt_golden.npz"), **out)


def grey_cases():
    rng = np.random.default_rng(20260103)
    case2d = np.arange(1, 10, dtype=np.float32).reshape(1, 1, 3, 3)
    case3d = np.arange(2 * 1 * 4 * 4 * 4, dtype=np.float32).reshape(2, 1, 4, 4, 4)
    cross = ndi.generate_binary_structure(2, 1)
    asym = np.arange(9, dtype=np.float32).reshape(3, 3)
    q2 = (np.floor(rng.random((2, 1, 17, 23)) * 16) / 16).astype(np.float32)
    q3 = (np.floor(rng.random((1, 2, 9, 11, 10)) * 256) / 256).astype(np.float32)
    zz, yy, xx = np.mgrid[-2:3, -2:3, -2:3]
    ball2 = (zz ** 2 + yy ** 2 + xx ** 2) <= 4
    heights = rng.integers(0, 4, size=ball2.shape).astype(np.float32)
    cases = []
    for mode, cval in [("reflect", 0.0), ("constant", 0.0), ("constant", 99.0), ("constant", -99.0),
                       ("nearest", 0.0), ("mirror", 0.0), ("wrap", 0.0)]:
        cases.append((f"case2d_size3_{mode}_{cval}", case2d, dict(size=(3, 3), mode=mode, cval=cval)))
    cases.append(("case2d_even", case2d, dict(size=(2, 3))))
    cases.append(("case2d_footprint", case2d, dict(footprint=cross)))
    cases.append(("case2d_asym_origin1", case2d, dict(structure=asym, origin=1)))
    cases.append(("case2d_struct_fp", case2d, dict(structure=cross.astype(np.float32),
                                                   footprint=cross)))
    cases.append(("case3d_size3", case3d, dict(size=(3, 3, 3))))
    for s in (3, 5, 7):
        cases.append((f"q2_size{s}", q2, dict(size=(s, s))))
    cases.append(("q2_size_4x6_origin", q2, dict(size=(4, 6), origin=(1, -2))))
    for mode in ("reflect", "constant", "nearest", "mirror", "wrap"):
        cases.append((f"q3_size3_{mode}", q3, dict(size=(3, 3, 3), mode=mode, cval=0.25)))
    cases.append(("q3_size5", q3, dict(size=(5, 5, 5))))
    cases.append(("q3_ball2_heights", q3, dict(footprint=ball2, structure=heights)))
    cases.append(("q3_ball2_flat_origin", q3, dict(footprint=ball2, origin=(1, 0, -1))))
    cases.append(("inf_erosion", np.full((1, 1, 5, 5), np.inf, np.float32), dict(size=(3, 3))))
    cases.append(("ninf_dilation", np.full((1, 1, 5, 5), -np.inf, np.float32), dict(size=(3, 3))))
    out = {}
    meta = []
    for i, (name, x, kw) in enumerate(cases):
        out[f"{i}_input"] = x
        for op in ("erosion", "dilation", "opening", "closing"):
            fn = getattr(ndi, f"grey_{op}")
            res, lead = per_item(x, lambda it: fn(it, **kw))
            out[f"{i}_{op}"] = np.stack(res).reshape(x.shape).astype(np.float32)
        m = {"name": name}
        for key, val in kw.items():
            if isinstance(val, np.ndarray):
                out[f"{i}_{key}"] = val
                m[key] = f"{i}_{key}"
            else:
                m[key] = list(val) if isinstance(val, tuple) else val
        meta.append(m)
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, "grey_golden.npz"), **out)


def binary_cases():
    rng = np.random.default_rng(20260104)
    case2d = np.array([[[[0, 1, 0], [1, 1, 1], [0, 1, 0]]]], dtype=bool)
    case3d = np.zeros((2, 1, 5, 5, 5), dtype=bool)
    case3d[0, 0, 2:3, 2:4, 1:2] = True
    case4d = np.zeros((2, 1, 4, 4, 4, 4), dtype=bool)
    case4d[0, 0, 1:3, 1:2, 2:3] = True
    r2 = rng.random((2, 1, 21, 19)) < 0.6
    r3 = rng.random((1, 2, 10, 12, 11)) < 0.7
    s2 = ndi.generate_binary_structure(2, 1)
    s3_2 = ndi.generate_binary_structure(3, 2)
    s3_3 = ndi.generate_binary_structure(3, 3)
    s4 = ndi.generate_binary_structure(4, 4)
    mask2 = np.array([[1, 1, 0], [1, 1, 1], [1, 1, 0]], dtype=bool)
    cases = [
        ("2d_basic", case2d, dict()),
        ("2d_structure", case2d, dict(structure=s2)),
        ("2d_iterations", case2d, dict(iterations=2)),
        ("2d_until_stable", case2d, dict(iterations=-1)),
        ("2d_origin_border", case2d, dict(origin=1, border_value=True)),
        ("2d_mask", case2d, dict(mask=mask2)),
        ("3d_basic", case3d, dict()),
        ("3d_structure_1", case3d, dict(structure=s3_2)),
        ("3d_origin_border", case3d, dict(structure=s3_3, origin=1, border_value=True)),
        ("4d_structure", case4d, dict(structure=s4)),
        ("4d_iterations", case4d, dict(iterations=2)),
        ("r2_iter3", r2, dict(iterations=3)),
        ("r2_border", r2, dict(border_value=True, iterations=2)),
        ("r3_iter4", r3, dict(iterations=4)),
        ("r3_full_iter2_border", r3, dict(structure=s3_3, iterations=2, border_value=True)),
        ("r3_origin", r3, dict(structure=s3_2, origin=(1, 0, -1))),
    ]
    out = {}
    meta = []
    for i, (name, x, kw) in enumerate(cases):
        out[f"{i}_input"] = x
        for op in ("erosion", "dilation", "opening", "closing"):
            fn = getattr(ndi, f"binary_{op}")
            res, lead = per_item(x, lambda it: fn(it, **kw))
            out[f"{i}_{op}"] = np.stack(res).reshape(x.shape).astype(bool)
        m = {"name": name}
        for key, val in kw.items():
            if isinstance(val, np.ndarray):
                out[f"{i}_{key}"] = val
                m[key] = f"{i}_{key}"
            else:
                m[key] = list(val) if isinstance(val, tuple) else val
        meta.append(m)
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, "binary_golden.npz"), **out)


def composite_cases():
    """binary_propagation / binary_fill_holes / binary_hit_or_miss
    (reference test/test_binary.py:190-367: the CASE_* inputs, masks,
    structures and origins there, plus seeded random volumes)."""
    rng = np.random.default_rng(20260105)
    case2d = np.array([[[[0, 1, 0], [1, 1, 1], [0, 1, 0]]]], dtype=bool)
    case3d = np.zeros((2, 1, 5, 5, 5), dtype=bool)
    case3d[0, 0, 2:3, 2:4, 1:2] = True
    case4d = np.zeros((2, 1, 4, 4, 4, 4), dtype=bool)
    case4d[0, 0, 1:3, 1:2, 2:3] = True
    holes2 = np.zeros((1, 1, 5, 5), dtype=bool)
    holes2[0, 0, 1:4, 1:4] = True
    holes2[0, 0, 2, 2] = False
    holes3 = np.zeros((1, 1, 5, 5, 5), dtype=bool)
    holes3[0, 0, 1:4, 1:4, 1:4] = True
    holes3[0, 0, 2, 2, 2] = False
    hit3 = ndi.generate_binary_structure(3, 1)[None, None]
    mask2 = np.array([[1, 1, 0], [1, 1, 1], [1, 1, 0]], dtype=bool)
    mask3 = np.zeros((5, 5, 5), dtype=bool)
    mask3[0:2, 0:1, 0:3] = True
    mask4 = np.zeros((4, 4, 4, 4), dtype=bool)
    mask4[1:2, 0:1, 1:3] = True
    s2 = ndi.generate_binary_structure(2, 1)
    s2f = ndi.generate_binary_structure(2, 2)
    s3_2 = ndi.generate_binary_structure(3, 2)
    s3_3 = ndi.generate_binary_structure(3, 3)
    s4 = ndi.generate_binary_structure(4, 4)
    # random blobs with enclosed holes: a thresholded smooth field
    field2 = ndi.gaussian_filter(rng.normal(size=(2, 1, 48, 70)), (0, 0, 2.0, 2.0))
    blobs2 = field2 > 0.02
    field3 = ndi.gaussian_filter(rng.normal(size=(1, 2, 20, 24, 30)), (0, 0, 1.5, 1.5, 1.5))
    blobs3 = field3 > 0.0
    rnd2 = rng.random((2, 1, 33, 65)) < 0.55
    rnd3 = rng.random((1, 2, 9, 14, 70)) < 0.6
    seeds2 = rng.random((1, 1, 40, 40)) < 0.01
    pmask2 = rng.random((40, 40)) < 0.62
    seeds3 = rng.random((1, 1, 12, 16, 20)) < 0.01
    pmask3 = rng.random((12, 16, 20)) < 0.7
    hm_s1 = np.array([[0, 1, 0], [0, 1, 1], [0, 0, 0]], dtype=bool)
    hm_s2 = np.array([[1, 0, 0], [0, 0, 0], [0, 0, 1]], dtype=bool)
    cases = [
        # propagation (reference :170-236)
        ("prop_2d_basic", "propagation", case2d, dict()),
        ("prop_2d_origin_border", "propagation", case2d, dict(structure=s2, origin=1,
                                                              border_value=True)),
        ("prop_3d_structure_1", "propagation", case3d, dict(structure=s3_2)),
        ("prop_3d_origin_border", "propagation", case3d, dict(structure=s3_3, origin=1,
                                                              border_value=True)),
        ("prop_4d_structure", "propagation", case4d, dict(structure=s4)),
        ("prop_2d_mask", "propagation", case2d, dict(mask=mask2)),
        ("prop_3d_mask", "propagation", case3d, dict(mask=mask3)),
        ("prop_4d_mask", "propagation", case4d, dict(mask=mask4)),
        ("prop_seeds_2d_mask", "propagation", seeds2, dict(mask=pmask2)),
        ("prop_seeds_2d_mask_full", "propagation", seeds2, dict(mask=pmask2, structure=s2f)),
        ("prop_seeds_3d_mask", "propagation", seeds3, dict(mask=pmask3)),
        ("prop_seeds_3d_mask_border", "propagation", seeds3, dict(mask=pmask3,
                                                                  border_value=True)),
        # fill holes (reference :265-300)
        ("fill_2d_default", "fill_holes", holes2, dict()),
        ("fill_2d_structure", "fill_holes", holes2, dict(structure=s2)),
        ("fill_3d_default", "fill_holes", holes3, dict()),
        ("fill_blobs_2d", "fill_holes", blobs2, dict()),
        ("fill_blobs_2d_full", "fill_holes", blobs2, dict(structure=s2f)),
        ("fill_blobs_2d_origin", "fill_holes", blobs2, dict(structure=s2f, origin=(1, -1))),
        ("fill_blobs_3d", "fill_holes", blobs3, dict()),
        ("fill_blobs_3d_26", "fill_holes", blobs3, dict(structure=s3_3)),
        ("fill_rnd_2d", "fill_holes", rnd2, dict()),
        # hit or miss (reference :305-367)
        ("hm_2d_default", "hit_or_miss", case2d, dict()),
        ("hm_2d_structures", "hit_or_miss", case2d, dict(structure1=s2,
                                                         structure2=np.logical_not(s2))),
        ("hm_3d_default", "hit_or_miss", hit3, dict()),
        ("hm_rnd_2d_default", "hit_or_miss", rnd2, dict()),
        ("hm_rnd_2d_pair", "hit_or_miss", rnd2, dict(structure1=hm_s1, structure2=hm_s2)),
        ("hm_rnd_2d_origins", "hit_or_miss", rnd2, dict(structure1=hm_s1, structure2=hm_s2,
                                                        origin1=(1, 0), origin2=(0, -1))),
        ("hm_rnd_2d_origin1_only", "hit_or_miss", rnd2, dict(structure1=hm_s1, origin1=-1)),
        ("hm_rnd_2d_empty_miss", "hit_or_miss", rnd2, dict(structure1=np.ones((3, 3), bool),
                                                           structure2=np.zeros((3, 3), bool))),
        ("hm_rnd_2d_empty_hit", "hit_or_miss", rnd2, dict(structure1=np.zeros((3, 3), bool),
                                                          structure2=s2)),
        ("hm_rnd_3d_default", "hit_or_miss", rnd3, dict()),
        ("hm_rnd_3d_26", "hit_or_miss", rnd3, dict(structure1=s3_2,
                                                   structure2=np.logical_not(s3_3))),
    ]
    out = {}
    meta = []
    for i, (name, op, x, kw) in enumerate(cases):
        fn = getattr(ndi, f"binary_{op}")
        res, _ = per_item(x, lambda it: fn(it, **kw))
        out[f"{i}_input"] = x
        out[f"{i}_expected"] = np.stack(res).reshape(x.shape).astype(bool)
        m = {"name": name, "op": op}
        for key, val in kw.items():
            if isinstance(val, np.ndarray):
                out[f"{i}_{key}"] = val
                m[key] = f"{i}_{key}"
            else:
                m[key] = list(val) if isinstance(val, tuple) else val
        meta.append(m)
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, "composite_golden.npz"), **out)


def structure_cases():
    """generate_binary_structure / iterate_structure, with origins
    (reference test/test_structure.py:10-77, plus larger ranks and
    iteration counts)."""
    cases = []
    for rank, conn in [(1, 1), (2, 1), (2, 2), (3, 1), (3, 2), (3, 3), (4, 2), (4, 4), (5, 3),
                       (7, 2)]:
        cases.append(({"name": f"gen_{rank}d{conn}c", "kind": "generate", "rank": rank,
                       "connectivity": conn}, ndi.generate_binary_structure(rank, conn)))
    rng = np.random.default_rng(20260106)
    iters = [
        ("1d_cross_3", ndi.generate_binary_structure(1, 1), 3, None),
        ("2d_cross_1", ndi.generate_binary_structure(2, 1), 1, None),
        ("2d_cross_2", ndi.generate_binary_structure(2, 1), 2, None),
        ("2d_full_3", ndi.generate_binary_structure(2, 2), 3, None),
        ("3d_cross_2", ndi.generate_binary_structure(3, 1), 2, None),
        ("3d_18_3", ndi.generate_binary_structure(3, 2), 3, None),
        ("3d_full_2", ndi.generate_binary_structure(3, 3), 2, None),
        ("4d_c2_2", ndi.generate_binary_structure(4, 2), 2, None),
        ("asymmetric", np.array([[False, True], [True, True]]), 3, None),
        ("empty", np.zeros((3, 3), dtype=bool), 2, None),
        ("random_2x4x3", rng.random((2, 4, 3)) < 0.5, 3, None),
        ("origin_scalar", ndi.generate_binary_structure(2, 1), 3, 0),
        ("origin_positive", ndi.generate_binary_structure(2, 1), 3, (1, 0)),
        ("origin_mixed", ndi.generate_binary_structure(2, 1), 3, (-1, 1)),
        ("origin_3d", ndi.generate_binary_structure(3, 1), 2, (1, -1, 0)),
    ]
    for name, st, k, org in iters:
        if org is None:
            res = ndi.iterate_structure(st, k)
            m = {"name": f"iter_{name}", "kind": "iterate", "iterations": k}
        else:
            res, o = ndi.iterate_structure(st, k, origin=org)
            m = {"name": f"iter_{name}", "kind": "iterate", "iterations": k,
                 "origin": list(org) if isinstance(org, tuple) else org,
                 "expected_origin": [int(v) for v in o]}
        cases.append(((m, st), res))
    out = {}
    meta = []
    for i, (m, res) in enumerate(cases):
        if isinstance(m, tuple):
            m, st = m
            out[f"{i}_structure"] = st.astype(bool)
        out[f"{i}_expected"] = np.asarray(res, dtype=bool)
        meta.append(m)
    out["meta"] = np.array(json.dumps(meta))
    np.savez_compressed(os.path.join(HERE, "structure_golden.npz"), **out)


if __name__ == "__main__":
    edt_cases()
    cdt_cases()
    grey_cases()
    binary_cases()
    composite_cases()
    structure_cases()
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))
