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
from .structure import _normalize_origin, _validate_origin, generate_binary_structure


def _normalize_structure(structure: Tensor, spatial_ndim: int, name: str = "structure") -> Tensor:
    if structure.ndim != spatial_ndim:
        raise ValueError(f"{name} dimension is not {spatial_ndim}, got {structure.ndim}")
    return (structure != 0).detach().to(device="cpu", dtype=torch.bool).contiguous()


def _run(x: Tensor, structure: Tensor, origin, border_value: bool, ops, mask: Tensor | None):
    """Runs [(dilation?, iterations), ...] on x (bool/float, contiguous)."""
    L = _lib.lib()
    xin, in_dtype = _lib.mask_operand(x)
    nd = x.ndim - 2
    batch = x.shape[0] * x.shape[1]
    shp = _lib.i64_array(x.shape[2:])
    out = torch.empty(x.shape, dtype=torch.bool, device=x.device)
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
    origin = _normalize_origin(origin, spatial_ndim)
    _validate_origin(origin, structure)
    if mask is not None and mask.shape != input.shape:
        raise ValueError(f"mask shape {mask.shape} must match input shape {input.shape}")
    return structure, origin


def _sequence(input, structure, origin, border_value, ops, mask, output):
    structure, origin = _prepare(input, structure, origin, mask, output)
    x = input.detach()
    if not structure.any().item():
        # empty structure: every step yields a constant (reference binary.py:54-58),
        # masked against that step's input
        cur = x != 0
        maskb = None if mask is None else mask.to(device=input.device, dtype=torch.bool)
        for dil, _ in ops:
            val = torch.full_like(cur, not dil)
            cur = val if maskb is None else torch.where(maskb, val, cur)
        result = cur
    else:
        result = _run(x.contiguous(), structure, origin, border_value, ops, mask)
    if output is not None:
        output.copy_(result)
        return output
    return result


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


def binary_fill_holes(input: Tensor, structure: Tensor | None = None,
                      output: Tensor | None = None, origin=0) -> Tensor:
    """Fill holes (reference binary.py:160-185)."""
    validate_bcs_input(input)
    validate_output(input, output)
    mask = input == 0
    seed = torch.zeros_like(mask, dtype=torch.bool)
    background = binary_propagation(seed, structure=structure, mask=mask, border_value=True,
                                    origin=origin)
    result = torch.logical_not(background)
    if output is not None:
        output.copy_(result)
        return output
    return result


def binary_hit_or_miss(input: Tensor, structure1: Tensor | None = None,
                       structure2: Tensor | None = None, output: Tensor | None = None,
                       origin1=0, origin2=None) -> Tensor:
    """Hit-or-miss transform (reference binary.py:188-228)."""
    spatial_ndim = validate_bcs_input(input)
    validate_output(input, output)
    origin1 = _normalize_origin(origin1, spatial_ndim)
    origin2 = origin1 if origin2 is None else _normalize_origin(origin2, spatial_ndim)
    if structure1 is None:
        structure1 = generate_binary_structure(spatial_ndim, 1)
    structure1 = _normalize_structure(structure1, spatial_ndim, "structure1")
    _validate_origin(origin1, structure1, "origin1")
    if structure2 is None:
        structure2 = torch.logical_not(structure1).contiguous()
    else:
        structure2 = _normalize_structure(structure2, spatial_ndim, "structure2")
    _validate_origin(origin2, structure2, "origin2")
    input_bool = input != 0
    if structure1.any().item():
        hit = binary_erosion(input_bool, structure=structure1, origin=origin1)
    else:
        hit = torch.ones_like(input_bool, dtype=torch.bool)
    if structure2.any().item():
        miss = binary_erosion(input_bool == 0, structure=structure2, origin=origin2)
        result = torch.logical_and(hit, miss)
    else:
        result = hit
    if output is not None:
        output.copy_(result)
        return output
    return result
