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


def test_ragged_grey_workspace_query_needs_no_gpu():
    from torchmorph_amd import _lib
    from torchmorph_amd.ragged import TmRaggedJob

    lib = _lib.lib()
    jobs = (TmRaggedJob * 2)()
    for k, sh in enumerate([(20, 33, 70), (1, 1, 1)]):
        jobs[k].ndim = 3
        for a, v in enumerate(sh):
            jobs[k].shape[a] = v
        jobs[k].in_offset = jobs[k].out_offset = 0
    # two item descriptors (one 256-byte block) + a block -> item table of
    # 2 x 3 x 2 + 1 tiles of 64 x 16 x 16 (one 256-byte block)
    assert lib.tm_ragged_grey_workspace_bytes(jobs, 2) == 512
    jobs[1].ndim = 2
    assert lib.tm_ragged_grey_workspace_bytes(jobs, 2) == 0
    assert "job 1 has 2 axes" in _lib.last_error()
    jobs[1].ndim = 3
    jobs[1].out_offset = -1
    assert lib.tm_ragged_grey_workspace_bytes(jobs, 2) == 0
    assert "negative offset" in _lib.last_error()


def test_grey_element_tables_of_growing_rank_need_no_gpu():
    # tm_grey builds its element table on the host before it needs a device:
    # with more than 3 axes and no workspace it returns TM_ERR_WORKSPACE (-3).
    # The same element count with more axes needs a larger offset table.
    from torchmorph_amd import _lib

    lib = _lib.lib()
    fake_in, fake_out = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    for se in ([3, 3, 1, 1], [3, 3, 1, 1, 1], [3, 3, 1, 1, 1, 1, 1, 1], [4, 4, 1, 1],
               [5, 5, 1, 1, 1, 1, 1, 1]):
        nd = len(se)
        rc = lib.tm_grey(fake_in, 1, nd, _lib.i64_array([4] * nd), None, None,
                         _lib.i64_array(se), _lib.i64_array([0] * nd), 1, 0.0, 0, fake_out,
                         None, 0, None)
        assert rc == -3, _lib.last_error()


def test_ragged_grey_oversized_element_is_reported_on_the_host():
    from torchmorph_amd import _lib
    from torchmorph_amd.ragged import TmRaggedJob

    lib = _lib.lib()
    jobs = (TmRaggedJob * 1)()
    jobs[0].ndim = 2
    jobs[0].shape[0] = jobs[0].shape[1] = 30
    need = lib.tm_ragged_grey_workspace_bytes(jobs, 1)
    st = (ctypes.c_float * 300)(*([1.0] * 300))
    rc = lib.tm_ragged_grey(jobs, 1, ctypes.c_void_p(0x1000), st, None, _lib.i64_array([20, 15]),
                            _lib.i64_array([0, 0]), 1, 0.0, 0, ctypes.c_void_p(0x2000), 1,
                            ctypes.c_void_p(0x100000), need, None)
    assert rc == -1
    assert "fits neither streaming kernel" in _lib.last_error()


def test_grey_fused_query_and_epilogue_errors_need_no_gpu():
    from torchmorph_amd import _lib

    lib = _lib.lib()
    i64 = _lib.i64_array
    assert lib.tm_grey_fused_ok(3, None, None, i64([3, 3, 3]), i64([0, 0, 0]), 0) == 1
    assert lib.tm_grey_fused_ok(4, None, None, i64([3, 3, 3, 3]), i64([0] * 4), 0) == 0
    st = (ctypes.c_float * 300)(*([1.0] * 300))  # 300 non-flat elements: no streaming kernel
    assert lib.tm_grey_fused_ok(2, st, None, i64([20, 15]), i64([0, 0]), 1) == 0
    a, out = ctypes.c_void_p(0x1000), ctypes.c_void_p(0x2000)
    args = (ctypes.c_void_p(0x3000), 1, 2, i64([8, 8]), None, None, i64([3, 3]), i64([0, 0]),
            1, 0.0, 1)
    assert lib.tm_grey_ex(*args, 6, a, None, out, None) == -1
    assert "epilogue must be" in _lib.last_error()
    # one pass (4, 5): flat boxes of odd sizes with zero origin only
    assert lib.tm_grey_dual_ok(2, i64([3, 5]), i64([0, 0])) == 1
    assert lib.tm_grey_dual_ok(2, i64([3, 4]), i64([0, 0])) == 0
    assert lib.tm_grey_dual_ok(2, i64([3, 3]), i64([1, 0])) == 0
    assert lib.tm_grey_dual_ok(4, i64([3] * 4), i64([0] * 4)) == 0
    st9 = (ctypes.c_float * 9)(*([0.0] * 9))
    assert lib.tm_grey_ex(args[0], 1, 2, i64([8, 8]), st9, None, i64([3, 3]), i64([0, 0]), 1,
                          0.0, 1, 4, None, None, out, None) == -1
    assert "flat box" in _lib.last_error()
    assert lib.tm_grey_ex(args[0], 1, 2, i64([8, 8]), None, None, i64([4, 3]), i64([0, 0]), 1,
                          0.0, 1, 5, None, None, out, None) == -1
    assert "odd sizes" in _lib.last_error()
    assert lib.tm_grey_ex(*args, 1, out, None, out, None) == -1
    assert "must not alias" in _lib.last_error()
    assert lib.tm_grey_ex(*args, 3, a, None, out, None) == -1
    assert "operand is NULL" in _lib.last_error()
