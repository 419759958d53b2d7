"""tm_edt_noalloc (the allocation-free EDT entry point) against tm_edt.

Both are called through the C ABI with ctypes; outputs must be bit-identical
(distances as float32 bits, indices exactly) on two calls in a row, the
second with a smaller input, on the 2-D unit-sampling path and on the N-D
path; a problem beyond the scratch maximum is rejected before any launch."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _call(fn_name, mask, sampling, want_idx, index_dtype, workspace):
    from torchmorph_amd import _lib

    L = _lib.lib()
    x = mask.contiguous().view(torch.uint8)
    nd = x.ndim - 2
    batch = x.shape[0] * x.shape[1]
    shape = _lib.i64_array(x.shape[2:])
    samp = None if sampling is None else _lib.f64_array(sampling)
    dist = torch.full(x.shape, float("nan"), dtype=torch.float32, device=x.device)
    idx = None
    itype = _lib.TM_I64 if index_dtype == torch.int64 else _lib.TM_I32
    if want_idx:
        idx = torch.full((nd,) + tuple(x.shape), -7, dtype=index_dtype, device=x.device)
    stream = _lib.stream_of(x.device)
    if workspace:
        nbytes = L.tm_edt_workspace_bytes(batch, nd, shape, samp)
        ws = _lib.workspace(nbytes, x.device)
        rc = L.tm_edt(_lib.ptr(x), _lib.TM_U8, batch, nd, shape, samp, _lib.ptr(dist),
                      _lib.ptr(idx), itype, _lib.ptr(ws), nbytes, stream)
    else:
        rc = L.tm_edt_noalloc(_lib.ptr(x), _lib.TM_U8, batch, nd, shape, samp, _lib.ptr(dist),
                              _lib.ptr(idx), itype, stream)
    _lib.check(rc, fn_name)
    torch.cuda.synchronize()
    return dist, idx


def _same(a, b):
    return a.shape == b.shape and torch.equal(a.view(torch.int32), b.view(torch.int32))


@pytest.mark.parametrize("case", ["2d_unit", "3d_aniso", "2d_aniso_int32"])
def test_noalloc_matches_workspace_entry_twice(case):
    g = torch.Generator().manual_seed(11)
    if case == "2d_unit":
        shapes, samp, itype = [(8, 1, 256, 256), (3, 1, 100, 77)], None, torch.int64
    elif case == "3d_aniso":
        shapes, samp, itype = [(2, 1, 40, 52, 61), (1, 1, 17, 9, 30)], [1.5, 0.75, 1.0], \
            torch.int64
    else:
        shapes, samp, itype = [(4, 2, 130, 96), (1, 1, 33, 65)], [0.5, 2.0], torch.int32
    for shape in shapes:  # the second call is smaller and reuses the same scratch
        mask = (torch.rand(shape, generator=g) < 0.9).cuda()
        for want_idx in (True, False):
            d0, i0 = _call("tm_edt", mask, samp, want_idx, itype, workspace=True)
            d1, i1 = _call("tm_edt_noalloc", mask, samp, want_idx, itype, workspace=False)
            assert _same(d0, d1), (case, shape, want_idx)
            if want_idx:
                assert torch.equal(i0, i1), (case, shape)


def test_noalloc_rejects_problems_beyond_the_scratch():
    from torchmorph_amd import _lib

    L = _lib.lib()
    shape = _lib.i64_array([4096, 4096])
    rc = L.tm_edt_noalloc(ctypes.c_void_p(16), _lib.TM_U8, 8, 2, shape, None,
                          ctypes.c_void_p(16), None, _lib.TM_I64, None)
    assert rc == -1
    assert "above the" in _lib.last_error() and "maximum" in _lib.last_error()
