"""Binary morphology on (B, C, spatial...) GPU tensors.

Drop-in for reference torchmorph/morphology/binary.py:9-290 (same names,
arguments, defaults, messages, iteration/mask/border semantics).  Every
erosion/dilation sequence (iterations, opening, closing) runs as one
bit-packed pipeline in libtorchmorph_hip (pack once, iterate on 64-voxel
words in cache, unpack once) instead of one launch plus a ``torch.where``
pass per iteration (binary.py:79-82).
"""
from __future__ import annotations

import functools

import torch
from torch import Tensor

from .. import _lib
from .._validation import validate_bcs_input, validate_output
from .structure import check_origin, generate_binary_structure, origin_tuple


def _normalize_structure(structure: Tensor, spatial_ndim: int, name: str = "structure") -> Tensor:
    if structure.ndim != spatial_ndim:
        raise ValueError(f"{name} dimension is not {spatial_ndim}, got {structure.ndim}")
    return (structure != 0).detach().to(device="cpu", dtype=torch.bool).contiguous()


def _run(x: Tensor, structure: Tensor, origin, border_value: bool, ops, mask: Tensor | None,
         into: Tensor | None = None):
    """Runs [(dilation?, iterations), ...] on x (bool/float, contiguous) into
    `into` (a bool contiguous tensor sharing no storage with x) or a new
    tensor."""
    L = _lib.lib()
    xin, in_dtype = _lib.mask_operand(x)
    nd = x.ndim - 2
    batch = x.shape[0] * x.shape[1]
    shp = _lib.i64_array(x.shape[2:])
    out = into if into is not None else torch.empty(x.shape, dtype=torch.bool, device=x.device)
    st = structure.to(torch.uint8).numpy()
    nops = len(ops)
    dil = (_lib.ctypes.c_int * nops)(*[int(d) for d, _ in ops])
    its = (_lib.ctypes.c_int * nops)(*[int(i) for _, i in ops])
    mk = None
    if mask is not None:
        mk = mask.to(device=x.device, dtype=torch.bool).expand(x.shape).contiguous().view(
            torch.uint8)
    nbytes = L.tm_binary_workspace_bytes(batch, nd, shp)
    ws = _lib.workspace(nbytes, x.device)
    with torch.cuda.device(x.device):
        _lib.check(L.tm_binary(_lib.ptr(xin), in_dtype, batch, nd, shp,
                               st.ctypes.data_as(_lib.ctypes.c_void_p),
                               _lib.i64_array(structure.shape), _lib.i64_array(origin),
                               int(bool(border_value)), nops, dil, its, _lib.ptr(mk),
                               _lib.ptr(out.view(torch.uint8)), _lib.ptr(ws), nbytes,
                               _lib.stream_of(x.device)), "tm_binary")
    return out


def _binary_morphology_cuda_step(input: Tensor, structure: Tensor, border_value: bool,
                                 origin: tuple[int, ...], *, mode: str) -> Tensor:
    """One erosion/dilation step (reference binary.py:15-25)."""
    return _run(input, structure, origin, border_value, [(mode == "dilation", 1)], None)


@functools.lru_cache(maxsize=None)
def _default_structure(rank: int) -> Tensor:
    """The cross of connectivity 1 (reference binary.py default), built once
    per rank; only read, never handed to the caller."""
    return generate_binary_structure(rank, 1)


def _prepare(input, structure, origin, mask, output):
    spatial_ndim = validate_bcs_input(input)
    validate_output(input, output)
    if structure is None:
        structure = _default_structure(spatial_ndim)
    structure = _normalize_structure(structure, spatial_ndim)
    origin = origin_tuple(origin, spatial_ndim)
    check_origin(origin, structure)
    if mask is not None and mask.shape != input.shape:
        raise ValueError(f"mask shape {mask.shape} must match input shape {input.shape}")
    return structure, origin


def _sequence(input, structure, origin, border_value, ops, mask, output):
    structure, origin = _prepare(input, structure, origin, mask, output)
    x = input.detach()
    if not bool(structure.any()):
        # empty structure: every step yields a constant (reference binary.py:54-58),
        # masked against that step's input
        cur = x != 0
        maskb = None if mask is None else mask.to(device=input.device, dtype=torch.bool)
        for dil, _ in ops:
            val = torch.full_like(cur, not dil)
            cur = val if maskb is None else torch.where(maskb, val, cur)
        result = cur
    else:
        x = x.contiguous()
        if (output is not None and output.dtype == torch.bool and output.is_contiguous()
                and output.untyped_storage().data_ptr() != x.untyped_storage().data_ptr()
                and (mask is None
                     or output.untyped_storage().data_ptr() != mask.untyped_storage().data_ptr())):
            return _run(x, structure, origin, border_value, ops, mask, into=output)
        result = _run(x, structure, origin, border_value, ops, mask)
    return _emit(result, output)


def binary_erosion(input: Tensor, structure: Tensor | None = None, iterations: int = 1,
                   mask: Tensor | None = None, output: Tensor | None = None,
                   border_value: bool = False, origin=0) -> Tensor:
    """N-D binary erosion (reference binary.py:91-114)."""
    return _sequence(input, structure, origin, border_value, [(False, iterations)], mask, output)


def binary_dilation(input: Tensor, structure: Tensor | None = None, iterations: int = 1,
                    mask: Tensor | None = None, output: Tensor | None = None,
                    border_value: bool = False, origin=0) -> Tensor:
    """N-D binary dilation (reference binary.py:117-136)."""
    return _sequence(input, structure, origin, border_value, [(True, iterations)], mask, output)


def binary_opening(input: Tensor, structure: Tensor | None = None, iterations: int = 1,
                   mask: Tensor | None = None, output: Tensor | None = None,
                   border_value: bool = False, origin=0) -> Tensor:
    """Dilation of the erosion, one fused pipeline (reference binary.py:231-259)."""
    return _sequence(input, structure, origin, border_value,
                     [(False, iterations), (True, iterations)], mask, output)


def binary_closing(input: Tensor, structure: Tensor | None = None, iterations: int = 1,
                   mask: Tensor | None = None, output: Tensor | None = None,
                   border_value: bool = False, origin=0) -> Tensor:
    """Erosion of the dilation, one fused pipeline (reference binary.py:262-290)."""
    return _sequence(input, structure, origin, border_value,
                     [(True, iterations), (False, iterations)], mask, output)


def binary_propagation(input: Tensor, structure: Tensor | None = None,
                       mask: Tensor | None = None, output: Tensor | None = None,
                       border_value: bool = False, origin=0) -> Tensor:
    """Dilation until stable (reference binary.py:139-157)."""
    return _sequence(input, structure, origin, border_value, [(True, -1)], mask, output)


def _emit(result: Tensor, output: Tensor | None) -> Tensor:
    if output is None:
        return result
    output.copy_(result)
    return output


def _packed_call(fn, input: Tensor, *args) -> Tensor:
    """Runs a packed-pipeline entry point `fn(input, dtype, batch, ndim,
    shape, *args, output, workspace, bytes, stream)` into a new bool tensor."""
    L = _lib.lib()
    x = input.detach().contiguous()
    xin, in_dtype = _lib.mask_operand(x)
    shape = _lib.i64_array(x.shape[2:])
    batch, nd = x.shape[0] * x.shape[1], x.ndim - 2
    out = torch.empty(x.shape, dtype=torch.bool, device=x.device)
    nbytes = L.tm_binary_workspace_bytes(batch, nd, shape)
    ws = _lib.workspace(nbytes, x.device)
    with torch.cuda.device(x.device):
        _lib.check(fn(_lib.ptr(xin), in_dtype, batch, nd, shape, *args,
                      _lib.ptr(out.view(torch.uint8)), _lib.ptr(ws), nbytes,
                      _lib.stream_of(x.device)), fn.__name__)
    return out


def _host_structure(structure: Tensor):
    """(uint8 host array, its shape as an int64 array) of a bool structure."""
    arr = structure.to(torch.uint8).numpy()
    return arr, arr.ctypes.data_as(_lib.ctypes.c_void_p), _lib.i64_array(structure.shape)


def binary_fill_holes(input: Tensor, structure: Tensor | None = None,
                      output: Tensor | None = None, origin=0) -> Tensor:
    """Sets every voxel that the outside of the item cannot reach through
    background voxels (reference binary.py:160-185, SciPy semantics: the
    reach is a dilation with ``structure`` from a border of ones, restricted
    to the background, repeated until stable).  One packed pipeline
    (``tm_binary_fill_holes``)."""
    spatial_ndim = validate_bcs_input(input)
    validate_output(input, output)
    if structure is None:
        structure = _default_structure(spatial_ndim)
    structure = _normalize_structure(structure, spatial_ndim)
    origin = origin_tuple(origin, spatial_ndim)
    check_origin(origin, structure)
    if not bool(structure.any()):
        # nothing propagates: no background is reached, everything is set
        return _emit(torch.ones(input.shape, dtype=torch.bool, device=input.device), output)
    keep, sptr, sshape = _host_structure(structure)
    result = _packed_call(_lib.lib().tm_binary_fill_holes, input, sptr, sshape,
                          _lib.i64_array(origin))
    del keep
    return _emit(result, output)


def binary_hit_or_miss(input: Tensor, structure1: Tensor | None = None,
                       structure2: Tensor | None = None, output: Tensor | None = None,
                       origin1=0, origin2=None) -> Tensor:
    """Voxels whose ``structure1`` neighbourhood is all set and whose
    ``structure2`` neighbourhood (default: the complement of structure1) is
    all clear (reference binary.py:188-228).  Both erosions and their AND run
    in one packed pipeline (``tm_binary_hit_or_miss``); an empty structure1
    hits everywhere, an empty structure2 misses nowhere."""
    spatial_ndim = validate_bcs_input(input)
    validate_output(input, output)
    origin1 = origin_tuple(origin1, spatial_ndim)
    origin2 = origin1 if origin2 is None else origin_tuple(origin2, spatial_ndim)
    s1 = _default_structure(spatial_ndim) if structure1 is None else structure1
    s1 = _normalize_structure(s1, spatial_ndim, "structure1")
    check_origin(origin1, s1, "origin1")
    s2 = (~s1).contiguous() if structure2 is None else \
        _normalize_structure(structure2, spatial_ndim, "structure2")
    check_origin(origin2, s2, "origin2")
    k1, p1, sh1 = _host_structure(s1)
    k2, p2, sh2 = _host_structure(s2)
    result = _packed_call(_lib.lib().tm_binary_hit_or_miss, input, p1, sh1,
                          _lib.i64_array(origin1), p2, sh2, _lib.i64_array(origin2))
    del k1, k2
    return _emit(result, output)
