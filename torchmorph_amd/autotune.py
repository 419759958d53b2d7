"""Autotuner control: the HIP library times launch variants per shape class on
first use (tm_tune.hip); this module exposes the cache as JSON so a tuned
configuration can be saved and restored across processes."""
from __future__ import annotations

import ctypes
import json

from . import _lib


def enable(flag: bool = True) -> None:
    _lib.lib().tm_autotune_enable(int(bool(flag)))


def clear() -> None:
    _lib.lib().tm_autotune_clear()


def export() -> dict[str, int]:
    L = _lib.lib()
    n = L.tm_autotune_export(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    L.tm_autotune_export(buf, n + 1)
    out = {}
    for line in buf.value.decode().splitlines():
        key, _, val = line.rpartition(" = ")
        if key:
            out[key] = int(val)
    return out


def load(table: dict[str, int]) -> int:
    text = "".join(f"{k} = {int(v)}\n" for k, v in table.items())
    return _lib.lib().tm_autotune_import(text.encode())


def save_json(path: str) -> None:
    with open(path, "w") as f:
        json.dump(export(), f, indent=1, sort_keys=True)


def load_json(path: str) -> int:
    with open(path) as f:
        return load(json.load(f))
