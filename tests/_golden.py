"""Loader for the committed SciPy golden vectors (tests/golden/*.npz)."""
from __future__ import annotations

import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(kind: str):
    """Yields (meta dict, arrays dict) per case of tests/golden/<kind>_golden.npz."""
    z = np.load(os.path.join(HERE, f"{kind}_golden.npz"), allow_pickle=False)
    meta = json.loads(str(z["meta"]))
    cases = []
    for i, m in enumerate(meta):
        arrays = {k.split("_", 1)[1]: z[k] for k in z.files if k.startswith(f"{i}_")}
        for key, val in list(m.items()):
            if isinstance(val, str) and val in z.files:
                m[key] = z[val]
        cases.append((m, arrays))
    return cases


def ids(cases):
    return [c[0]["name"] + ("-" + c[0]["metric"] if "metric" in c[0] else "") for c in cases]


def assert_bits_equal(got, want, err_msg=""):
    """Bit-exact float comparison: the IEEE bit patterns must agree (so -0.0
    differs from +0.0), every NaN canonicalised to one quiet NaN first."""
    got = np.ascontiguousarray(np.asarray(got, dtype=np.float32))
    want = np.ascontiguousarray(np.asarray(want, dtype=np.float32))
    assert got.shape == want.shape, (got.shape, want.shape)
    gb = np.where(np.isnan(got), np.float32(np.nan), got).view(np.uint32)
    wb = np.where(np.isnan(want), np.float32(np.nan), want).view(np.uint32)
    bad = np.flatnonzero(gb != wb)
    assert bad.size == 0, (f"{err_msg} {bad.size} of {gb.size} values differ in their bits; "
                           f"first at flat {bad[0]}: got {got.flat[bad[0]]!r} "
                           f"want {want.flat[bad[0]]!r}")
