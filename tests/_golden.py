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
