"""Drop-in replacement for the reference's compiled extension ``torchmorph._C``.

Same nine function names, argument lists and return conventions as the
reference's pybind11 module (reference torchmorph/csrc/torchmorph.cpp:3-77,
79-155), implemented over the C ABI of libtorchmorph_hip.  With
``sys.modules["torchmorph._C"] = torchmorph_amd._C`` set before the
reference package imports its submodules, the reference's own Python
wrappers (grey.py, binary.py, distance_transform.py, optimal_transport.py)
run on the gfx950 kernels unchanged.
"""
from __future__ import annotations

import torch

from . import _lib

_MODES = {0, 1, 2, 3, 4}


def _grey(input, structure, footprint, origin, mode, cval, dilation):
    if not input.is_cuda or input.dtype != torch.float32:
        raise RuntimeError("input must be a float32 CUDA tensor")
    if mode not in _MODES:
        raise RuntimeError("mode must be in 0..4")
    from .morphology.grey import _launch

    x = input.contiguous()
    fp = None if footprint is None or footprint.numel() == 0 else footprint.to("cpu", torch.bool)
    st = structure.to("cpu", torch.float32).contiguous()
    out = torch.empty_like(x)
    _launch(x, st, fp, tuple(st.shape), tuple(int(o) for o in origin), int(mode), float(cval),
            dilation, out)
    return out


def grey_erosion_cuda(input, structure, footprint, origin, mode, cval):
    """Reference torchmorph.cpp:3-10."""
    return _grey(input, structure, footprint, origin, mode, cval, False)


def grey_dilation_cuda(input, structure, footprint, origin, mode, cval):
    """Reference torchmorph.cpp:12-19."""
    return _grey(input, structure, footprint, origin, mode, cval, True)


def _binary(input, structure, origin, border_value, dilation):
    from .morphology.binary import _run

    st = (structure != 0).to("cpu")
    return _run(input.contiguous(), st, tuple(int(o) for o in origin), bool(border_value),
                [(dilation, 1)], None)


def binary_erosion_cuda(input, structure, origin, border_value):
    """Reference torchmorph.cpp:21-26."""
    return _binary(input, structure, origin, border_value, False)


def binary_dilation_cuda(input, structure, origin, border_value):
    """Reference torchmorph.cpp:28-33."""
    return _binary(input, structure, origin, border_value, True)


def edt_cuda(input, sampling, return_distances, return_indices):
    """Reference torchmorph.cpp:35-40: (distances f32 or None, int32 indices or None)."""
    from .distance_transform import euclidean_distance_transform

    out = euclidean_distance_transform(input, sampling=list(sampling),
                                       return_distances=return_distances,
                                       return_indices=return_indices)
    if return_distances and return_indices:
        return out
    return (out, None) if return_distances else (None, out)


def cdt_cuda(input, metric, return_distances, return_indices):
    """Reference torchmorph.cpp:42-47."""
    from .distance_transform import chamfer_distance_transform

    out = chamfer_distance_transform(input, metric=metric, return_distances=return_distances,
                                     return_indices=return_indices)
    if return_distances and return_indices:
        return out
    return (out, None) if return_distances else (None, out)


def bfdt_cuda(input, metric, sampling, return_distances, return_indices):
    """Reference torchmorph.cpp:49-55."""
    from .distance_transform import brute_force_distance_transform

    out = brute_force_distance_transform(input, metric=metric, sampling=list(sampling),
                                         return_distances=return_distances,
                                         return_indices=return_indices)
    if return_distances and return_indices:
        return out
    return (out, None) if return_distances else (None, out)


def _hist(a, b, M, M_T, u, v, n_iter, epsilon, log_space):
    dtype = {torch.float32: _lib.TM_F32, torch.float64: _lib.TM_F64}.get(a.dtype)
    if dtype is None or not a.is_cuda:
        raise RuntimeError("Sinkhorn kernels take float32/float64 CUDA tensors")
    for t in (a, b, M, M_T, u, v):
        if not t.is_contiguous():
            raise RuntimeError("Sinkhorn kernels take contiguous tensors")
    n, d = a.shape
    fb = _lib.ctypes.c_int(0)
    with torch.cuda.device(a.device):
        _lib.check(_lib.lib().tm_sinkhorn_hist(dtype, int(log_space), _lib.ptr(a), _lib.ptr(b),
                                               _lib.ptr(M), _lib.ptr(M_T), _lib.ptr(u),
                                               _lib.ptr(v), n, d, int(n_iter), float(epsilon),
                                               _lib.ctypes.byref(fb),
                                               _lib.stream_of(a.device)), "tm_sinkhorn_hist")
    return u, v


def sinkhorn_fastiter(a, b, K, K_T, u, v, n_iter):
    """Reference torchmorph.cpp:58-66: n_iter scaling updates of u, v in place."""
    return _hist(a, b, K, K_T, u, v, n_iter, 1.0, False)


def sinkhorn_logiter(log_a, log_b, M, M_T, log_u, log_v, n_iter, epsilon):
    """Reference torchmorph.cpp:68-77: n_iter log-domain updates in place."""
    return _hist(log_a, log_b, M, M_T, log_u, log_v, n_iter, epsilon, True)
