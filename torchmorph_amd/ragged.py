"""Ragged batches of distance transforms in one persistent launch.

No reference counterpart: the reference's distance transforms take dense
``(B, C, spatial...)`` tensors (reference distance_transform.py:10-138), so a
list of masks of different shapes costs one call per shape there.  Here every
(mask, op) pair is a job of one work-list (``tm_ragged_dt`` in
include/torchmorph_hip.h) and each job's result is bit-identical to
``euclidean_distance_transform`` / ``chamfer_distance_transform`` on that mask
alone (SciPy semantics, indices included).

    plan = RaggedDistanceTransform([m.shape for m in masks],
                                   ops=("edt", "cdt_taxicab", "cdt_chessboard"),
                                   sampling=per_mask_spacings)
    edt, taxi, cheb = plan(masks)          # one list per op, one tensor per mask
"""
from __future__ import annotations

import ctypes
import math
from collections.abc import Sequence

import torch
from torch import Tensor

from . import _lib

__all__ = [
    "RaggedDistanceTransform",
    "euclidean_distance_transform_ragged",
    "chamfer_distance_transform_ragged",
]

TM_RG_EDT, TM_RG_CDT_CHESSBOARD, TM_RG_CDT_TAXICAB = 0, 1, 2
_MAX_NDIM = 8

_OPS = {
    "edt": TM_RG_EDT,
    "euclidean": TM_RG_EDT,
    "cdt": TM_RG_CDT_CHESSBOARD,
    "cdt_chessboard": TM_RG_CDT_CHESSBOARD,
    "chessboard": TM_RG_CDT_CHESSBOARD,
    "cdt_taxicab": TM_RG_CDT_TAXICAB,
    "taxicab": TM_RG_CDT_TAXICAB,
    "cityblock": TM_RG_CDT_TAXICAB,
    "manhattan": TM_RG_CDT_TAXICAB,
}

_SCHEDULES = {"persistent": 0, "per_level": 1}


class TmRaggedJob(ctypes.Structure):
    """ctypes mirror of ``tm_ragged_job``."""

    _fields_ = [
        ("op", ctypes.c_int32),
        ("ndim", ctypes.c_int32),
        ("shape", ctypes.c_int64 * _MAX_NDIM),
        ("sampling", ctypes.c_double * _MAX_NDIM),
        ("in_offset", ctypes.c_int64),
        ("out_offset", ctypes.c_int64),
        ("idx_offset", ctypes.c_int64),
    ]


def _item_sampling(sampling, shapes, k: int) -> list[float]:
    nd = len(shapes[k])
    if sampling is None:
        vals = [1.0] * nd
    elif isinstance(sampling, (int, float)):
        vals = [float(sampling)] * nd
    else:
        s = sampling[k] if len(sampling) == len(shapes) and not isinstance(sampling[0], (int, float)) \
            else sampling
        if isinstance(s, (int, float)):
            vals = [float(s)] * nd
        else:
            vals = [float(v) for v in s]
            if len(vals) == 1:
                vals *= nd
            elif len(vals) != nd:
                raise ValueError(f"sampling of mask {k} must have length 1 or {nd}, got {len(vals)}")
    if any(not math.isfinite(v) or v <= 0 for v in vals):
        raise ValueError("sampling values must be finite and greater than zero")
    return vals


class RaggedDistanceTransform:
    """A fixed list of mask shapes and ops, with its outputs and workspace.

    Args:
        shapes: spatial shape of every mask (1..8 axes, each 1..65535).
        ops: op names, any of ``"edt"``, ``"cdt_chessboard"`` (``"cdt"``),
            ``"cdt_taxicab"``; every op runs on every mask.
        sampling: EDT spacing -- None, one number, one sequence for every mask,
            or one number/sequence per mask.
        return_distances / return_indices: outputs to produce.
        index_dtype: torch.int32 or torch.int64.
        schedule: ``"persistent"`` (one launch, device dependency counters) or
            ``"per_level"`` (one launch per pass level).
        device: CUDA device of the masks and outputs.
    """

    def __init__(self, shapes: Sequence[Sequence[int]], ops: Sequence[str] = ("edt",),
                 sampling=None, return_distances: bool = True, return_indices: bool = False,
                 index_dtype: torch.dtype = torch.int64, schedule: str = "persistent",
                 device: torch.device | str | None = None):
        if not shapes:
            raise ValueError("at least one mask shape is required")
        if not return_distances and not return_indices:
            raise ValueError("at least one of return_distances/return_indices must be True")
        if index_dtype not in (torch.int32, torch.int64):
            raise ValueError("index_dtype must be torch.int32 or torch.int64")
        if schedule not in _SCHEDULES:
            raise ValueError(f"schedule must be one of {sorted(_SCHEDULES)}")
        self.shapes = [tuple(int(s) for s in sh) for sh in shapes]
        for sh in self.shapes:
            if not 1 <= len(sh) <= _MAX_NDIM or any(s < 1 or s > 65535 for s in sh):
                raise ValueError(f"mask shape {sh}: 1..8 axes of size 1..65535 are supported")
        try:
            self.ops = [op.lower() for op in ops]
            codes = [_OPS[op] for op in self.ops]
        except KeyError as e:
            raise ValueError(f"unknown op {e.args[0]!r}; choose from {sorted(_OPS)}") from None
        dev = torch.device(device) if device is not None else torch.device("cuda")
        if dev.type != "cuda":
            raise ValueError("the ragged distance transforms run on CUDA devices only")
        self.device = dev if dev.index is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.return_distances = return_distances
        self.return_indices = return_indices
        self.index_dtype = index_dtype
        self.schedule = _SCHEDULES[schedule]

        numels = [math.prod(sh) for sh in self.shapes]
        self.in_offsets = [0]
        for n in numels:
            self.in_offsets.append(self.in_offsets[-1] + n)
        n_items = len(self.shapes)
        jobs = (TmRaggedJob * (len(codes) * n_items))()
        out_off = 0
        idx_off = 0
        self._views = []  # per op: list of (out_offset, idx_offset)
        for o, code in enumerate(codes):
            views = []
            for k, sh in enumerate(self.shapes):
                j = jobs[o * n_items + k]
                j.op = code
                j.ndim = len(sh)
                for a, s in enumerate(sh):
                    j.shape[a] = s
                samp = _item_sampling(sampling, self.shapes, k) if code == TM_RG_EDT \
                    else [1.0] * len(sh)
                for a, s in enumerate(samp):
                    j.sampling[a] = s
                j.in_offset = self.in_offsets[k]
                j.out_offset = out_off if return_distances else -1
                j.idx_offset = idx_off if return_indices else -1
                views.append((j.out_offset, j.idx_offset))
                if return_distances:
                    out_off += numels[k]
                if return_indices:
                    idx_off += numels[k] * len(sh)
            self._views.append(views)
        self._jobs = jobs
        self._njobs = len(jobs)
        L = _lib.lib()
        need = int(L.tm_ragged_dt_workspace_bytes(jobs, self._njobs))
        if need == 0:
            raise RuntimeError(f"tm_ragged_dt_workspace_bytes: {_lib.last_error()}")
        # 256-byte aligned workspace (the caching allocator aligns to 512)
        self.workspace = torch.empty(need, dtype=torch.uint8, device=self.device)
        self.distances = torch.empty(out_off, dtype=torch.float32, device=self.device) \
            if return_distances else None
        self.indices = torch.empty(idx_off, dtype=index_dtype, device=self.device) \
            if return_indices else None
        self._uploaded = False

    @property
    def numel(self) -> int:
        return self.in_offsets[-1]

    def pack(self, masks: Sequence[Tensor]) -> Tensor:
        """Flat uint8 (or float32) buffer of the masks in list order."""
        if len(masks) != len(self.shapes):
            raise ValueError(f"expected {len(self.shapes)} masks, got {len(masks)}")
        for k, m in enumerate(masks):
            if tuple(m.shape) != self.shapes[k]:
                raise ValueError(f"mask {k} has shape {tuple(m.shape)}, expected {self.shapes[k]}")
        if all(m.dtype == torch.float32 for m in masks):
            return torch.cat([m.reshape(-1) for m in masks]).to(self.device)
        parts = [(m != 0).reshape(-1).to(torch.uint8) for m in masks]
        return torch.cat(parts).to(self.device)

    def run(self, flat: Tensor, check: bool = True) -> None:
        """Runs every job on a packed input; results land in ``self.distances``
        / ``self.indices``.  ``check`` synchronises and reports an aborted
        dependency wait."""
        if not flat.is_cuda or flat.device != self.device:
            raise ValueError(f"input must be on {self.device}")
        if flat.dim() != 1 or flat.numel() != self.numel or not flat.is_contiguous():
            raise ValueError(f"packed input must be a contiguous 1-D tensor of {self.numel} elements")
        if flat.dtype == torch.bool:
            flat = flat.view(torch.uint8)
        if flat.dtype == torch.uint8:
            dt = _lib.TM_U8
        elif flat.dtype == torch.float32:
            dt = _lib.TM_F32
        else:
            raise ValueError("packed input must be bool, uint8 or float32")
        L = _lib.lib()
        with torch.cuda.device(self.device):
            st = _lib.stream_of(self.device)
            _lib.check(L.tm_ragged_dt(
                self._jobs, self._njobs, _lib.ptr(flat), dt, _lib.ptr(self.distances),
                _lib.ptr(self.indices),
                _lib.TM_I64 if self.index_dtype == torch.int64 else _lib.TM_I32,
                self.schedule, 0 if self._uploaded else 1, _lib.ptr(self.workspace),
                self.workspace.numel(), st), "tm_ragged_dt")
            self._uploaded = True
            if check:
                _lib.check(L.tm_ragged_dt_check(_lib.ptr(self.workspace), st), "tm_ragged_dt")

    def results(self):
        """One list per op, one entry per mask: distances, indices
        ``(ndim, *shape)``, or ``(distances, indices)`` -- views of the flat
        outputs."""
        out = []
        for views in self._views:
            per = []
            for k, (do, io) in enumerate(views):
                sh = self.shapes[k]
                n = math.prod(sh)
                d = self.distances[do:do + n].view(sh) if self.return_distances else None
                i = self.indices[io:io + n * len(sh)].view((len(sh),) + sh) \
                    if self.return_indices else None
                per.append(d if i is None else i if d is None else (d, i))
            out.append(per)
        return out

    def __call__(self, masks: Sequence[Tensor] | Tensor, check: bool = True):
        flat = masks if isinstance(masks, Tensor) else self.pack(masks)
        self.run(flat, check=check)
        return self.results()


def euclidean_distance_transform_ragged(masks: Sequence[Tensor], sampling=None,
                                        return_distances: bool = True,
                                        return_indices: bool = False):
    """EDT of every mask of a ragged list (spatial shapes only, no batch or
    channel axes) in one launch; one result per mask, as
    ``euclidean_distance_transform`` returns it for that mask."""
    plan = RaggedDistanceTransform([m.shape for m in masks], ("edt",), sampling,
                                   return_distances, return_indices, device=masks[0].device)
    return plan(masks)[0]


def chamfer_distance_transform_ragged(masks: Sequence[Tensor], metric: str = "chessboard",
                                      return_distances: bool = True,
                                      return_indices: bool = False):
    """Chamfer DT of every mask of a ragged list in one launch."""
    m = metric.lower()
    if m not in ("chessboard", "taxicab", "cityblock", "manhattan"):
        raise ValueError("metric must be 'chessboard', 'taxicab', 'cityblock' or 'manhattan'")
    plan = RaggedDistanceTransform([x.shape for x in masks], (m,), None, return_distances,
                                   return_indices, device=masks[0].device)
    return plan(masks)[0]
