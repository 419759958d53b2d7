"""Differentiable grey erosion / dilation (opt-in ``differentiable=True``).

No reference counterpart: the reference's grey operators are not
differentiable (test_grey.py:218-222).  The backward is the HIP kernel pair
in tm_grey.hip (tm_grey_backward): the upstream gradient of every output goes
to the input voxel read by the first structuring element (in the reference's
element order) that attains the min/max -- a deterministic subgradient; no
gradient flows to voxels replaced by ``cval``.
"""
from __future__ import annotations

import torch

from .. import _lib


class _GreyMorph(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, struct, fp, se_shape, origin, mode_int, cval, dilation):
        from .grey import _launch

        x32 = x.detach().float().contiguous()
        y = torch.empty_like(x32)
        _launch(x32, struct, fp, se_shape, origin, mode_int, cval, dilation, y)
        ctx.save_for_backward(x32)
        ctx.args = (struct, fp, se_shape, origin, mode_int, cval, dilation)
        # host-side arguments of the backward call, formed once here
        st = None if struct is None else struct.numpy()
        fpn = None if fp is None else fp.to(torch.uint8).numpy()
        ctx.arrays = (_lib.i64_array(x32.shape[2:]), _lib.i64_array(se_shape),
                      _lib.i64_array(origin),
                      None if st is None else st.ctypes.data_as(_lib.ctypes.c_void_p),
                      None if fpn is None else fpn.ctypes.data_as(_lib.ctypes.c_void_p),
                      st, fpn)  # the last two keep the buffers alive
        ctx.in_dtype = x.dtype
        return y

    @staticmethod
    def backward(ctx, gy):
        (x32,) = ctx.saved_tensors
        struct, fp, se_shape, origin, mode_int, cval, dilation = ctx.args
        shp, se_arr, org_arr, st_p, fp_p = ctx.arrays[:5]
        dx = torch.empty_like(x32)
        gy = gy.float().contiguous()
        L = _lib.lib()
        nd = x32.ndim - 2
        batch = x32.shape[0] * x32.shape[1]

        def call(ws, nbytes):
            return L.tm_grey_backward(
                _lib.ptr(x32), batch, nd, shp, st_p, fp_p, se_arr, org_arr, mode_int,
                float(cval), int(dilation), _lib.ptr(gy), _lib.ptr(dx), _lib.ptr(ws), nbytes,
                _lib.stream_of(x32.device))

        with torch.cuda.device(x32.device):
            # 2-D items (modes other than wrap) usually run kernels that need
            # no workspace; the library asks for one when they do not apply
            rc = call(None, 0) if nd == 2 and mode_int != 4 else _lib.TM_ERR_WORKSPACE
            if rc == _lib.TM_ERR_WORKSPACE:
                nbytes = L.tm_grey_backward_workspace_bytes(batch, nd, shp)
                rc = call(_lib.workspace(nbytes, x32.device), nbytes)
            _lib.check(rc, "tm_grey_backward")
        return dx.to(ctx.in_dtype), None, None, None, None, None, None, None


def grey_autograd(x, struct, fp, se_shape, origin, mode_int, cval, dilation):
    return _GreyMorph.apply(x, struct, fp, se_shape, origin, mode_int, cval, dilation)
