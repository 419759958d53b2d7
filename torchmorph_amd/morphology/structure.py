"""Structuring-element helpers (CPU torch), same semantics and messages as
reference torchmorph/morphology/structure.py:5-97 (SciPy's
generate_binary_structure / iterate_structure, origin rules)."""
from __future__ import annotations

import itertools

import torch
from torch import Tensor


def generate_binary_structure(rank: int, connectivity: int) -> Tensor:
    """Elements within `connectivity` unit steps of the centre of a 3^rank cube."""
    if connectivity < 1 or connectivity > rank:
        raise ValueError(f"connectivity must be in [1, rank], got {connectivity}")
    if rank < 1:
        raise ValueError(f"rank must be >= 1, got {rank}")
    axes = [torch.tensor([-1, 0, 1]) for _ in range(rank)]
    offsets = torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=0)
    return (offsets != 0).sum(dim=0) <= connectivity


def _generated_structure_connectivity(structure: Tensor) -> int | None:
    if any(size != 3 for size in structure.shape):
        return None
    rank = structure.ndim
    for connectivity in range(1, rank + 1):
        if torch.equal(structure.cpu(), generate_binary_structure(rank, connectivity)):
            return connectivity
    return None


def _normalize_origin(origin, ndim: int) -> tuple[int, ...]:
    if isinstance(origin, int):
        return (origin,) * ndim
    origin_tuple = tuple(int(v) for v in origin)
    if len(origin_tuple) != ndim:
        raise ValueError(f"origin dimension is not {ndim}, got {len(origin_tuple)}")
    return origin_tuple


def _validate_origin(origin: tuple[int, ...], structure: Tensor, name: str = "origin") -> None:
    for value, size in zip(origin, structure.shape):
        if not -(size // 2) <= value <= (size - 1) // 2:
            raise ValueError(f"invalid {name}")


def iterate_structure(structure: Tensor, iterations: int, origin=None):
    """Dilate a binary structure with itself `iterations - 1` times
    (reference structure.py:60-97)."""
    structure = structure != 0
    if iterations < 2:
        result = structure.clone()
    else:
        conn = _generated_structure_connectivity(structure)
        shape = tuple(iterations * (s - 1) + 1 for s in structure.shape)
        if conn is not None:
            axes = [torch.arange(-iterations, iterations + 1, device=structure.device)
                    for _ in range(structure.ndim)]
            offsets = torch.stack(torch.meshgrid(*axes, indexing="ij"), dim=0)
            result = offsets.abs().sum(dim=0) <= conn * iterations
        else:
            coords = torch.nonzero(structure, as_tuple=False)
            result = torch.zeros(shape, dtype=torch.bool, device=structure.device)
            if coords.numel():
                cur = coords
                for _ in range(iterations - 1):
                    cur = torch.unique((cur[:, None, :] + coords[None, :, :]).reshape(
                        -1, structure.ndim), dim=0)
                result[tuple(cur.unbind(dim=1))] = True
    if origin is None:
        return result
    return result, [v * iterations for v in _normalize_origin(origin, structure.ndim)]


def _cross(ndim: int) -> Tensor:
    s = torch.zeros((3,) * ndim, dtype=torch.bool)
    for p in itertools.product(range(3), repeat=ndim):
        s[p] = sum(abs(v - 1) for v in p) <= 1
    return s
