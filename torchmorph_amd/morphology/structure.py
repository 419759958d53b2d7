"""Structuring-element helpers (CPU torch).

Public names, arguments and error messages are those of reference
torchmorph/morphology/structure.py:5-97; the definitions are SciPy's
(``scipy.ndimage.generate_binary_structure`` / ``iterate_structure``):

* the structure of a given connectivity holds the points of the 3^rank cube
  around the centre that are reached in at most ``connectivity`` unit steps,
  i.e. whose number of nonzero offsets is <= connectivity;
* iterating a structure k times is its k-fold Minkowski sum with itself (the
  structure dilated by itself k - 1 times), on an array of extent
  k (s - 1) + 1 per axis; the origin scales with k.

Origins follow SciPy's range rule: along an axis of extent s an origin must
lie in [-(s // 2), (s - 1) // 2].
"""
from __future__ import annotations

import torch
from torch import Tensor


def generate_binary_structure(rank: int, connectivity: int) -> Tensor:
    """(3,) * rank bool tensor: the cube points with at most ``connectivity``
    nonzero offsets from the centre."""
    if connectivity < 1 or connectivity > rank:
        raise ValueError(f"connectivity must be in [1, rank], got {connectivity}")
    if rank < 1:
        raise ValueError(f"rank must be >= 1, got {rank}")
    # the number of nonzero offsets of every cube point, one axis at a time
    steps = torch.zeros((), dtype=torch.int64)
    for _ in range(rank):
        steps = (steps[..., None] + torch.tensor([1, 0, 1])).contiguous()
    return steps <= connectivity


def origin_tuple(origin, ndim: int) -> tuple[int, ...]:
    """An int origin broadcast to ``ndim`` axes, or a sequence of exactly
    ``ndim`` ints."""
    if isinstance(origin, int):
        return (origin,) * ndim
    values = tuple(int(v) for v in origin)
    if len(values) != ndim:
        raise ValueError(f"origin dimension is not {ndim}, got {len(values)}")
    return values


def check_origin(origin: tuple[int, ...], structure: Tensor, name: str = "origin") -> None:
    """Raises ValueError("invalid <name>") when an origin value falls outside
    [-(s // 2), (s - 1) // 2] for its axis extent s."""
    for value, extent in zip(origin, structure.shape):
        if value < -(extent // 2) or value > (extent - 1) // 2:
            raise ValueError(f"invalid {name}")


def _minkowski_sum(a: Tensor, b: Tensor) -> Tensor:
    """Bool array of the sums p + q (p active in a, q active in b), on the
    full extent a.shape + b.shape - 1: one shifted OR of a per active q."""
    out = torch.zeros([x + y - 1 for x, y in zip(a.shape, b.shape)], dtype=torch.bool)
    for q in torch.nonzero(b).tolist():
        window = tuple(slice(lo, lo + n) for lo, n in zip(q, a.shape))
        out[window] |= a
    return out


def iterate_structure(structure: Tensor, iterations: int, origin=None):
    """``structure`` dilated by itself ``iterations - 1`` times (SciPy's
    iterate_structure); with ``origin`` also returns the origin scaled by
    ``iterations``."""
    base = (structure != 0).cpu()
    result = base.clone()
    for _ in range(iterations - 1):
        result = _minkowski_sum(result, base)
    result = result.to(structure.device)
    if origin is None:
        return result
    return result, [v * iterations for v in origin_tuple(origin, structure.ndim)]
