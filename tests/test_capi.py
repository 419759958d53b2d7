"""CPU checks of the C-ABI boundary: the HIP library loads next to torch and
exports every entry point declared in include/*.h (no compute calls)."""
import ctypes
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(tm_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "tm_edt" in names and "tm_last_error" in names


@pytest.mark.parametrize("name", declared_symbols())
def test_library_exports_declared_symbol(name):
    import torch  # noqa: F401  (shares torch's HIP runtime, as the package does)
    from torchmorph_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def test_python_bindings_cover_every_declared_symbol():
    from torchmorph_amd import _lib

    missing = [n for n in declared_symbols() if n not in _lib.SIGNATURES]
    assert not missing, missing


def test_version_and_error_calls_need_no_gpu():
    from torchmorph_amd import _lib

    lib = _lib.lib()
    assert b"gfx950" in lib.tm_version()
    # an argument error is reported without touching the device
    rc = lib.tm_edt(None, 0, 1, 2, _lib.i64_array([4, 4]), None, None, None, 2, None, 0, None)
    assert rc == -1
    assert "input is NULL" in _lib.last_error()
