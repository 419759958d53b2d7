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

Grey morphology has the same form (``tm_ragged_grey``): a list of float items
of one spatial rank and one structuring element, every tile of every item in
one launch, each result bit-identical to ``grey_erosion`` / ``grey_dilation``
on that item (reference grey.py:29-197 takes one dense tensor per call).

    plan = RaggedGreyMorphology([v.shape for v in volumes], "opening", size=3)
    opened = plan(volumes)                 # one tensor per volume

Binary morphology of a ragged list (``RaggedBinaryMorphology``) runs its steps
on the same kernels: one step of binary erosion / dilation (reference
binary_kernel.cu:56-127, binary.py:15-25) is grey erosion / dilation of the
0/1 mask with the structure as a flat footprint, constant mode and cval =
border_value (AND = min and OR = max on {0, 1}, the same element offsets).
"""
from __future__ import annotations

import ctypes
import math
from collections.abc import Sequence

import torch
from torch import Tensor

from . import _lib
from .morphology.grey import _resolve_element_nd

__all__ = [
    "RaggedDistanceTransform",
    "euclidean_distance_transform_ragged",
    "chamfer_distance_transform_ragged",
    "RaggedGreyMorphology",
    "grey_erosion_ragged",
    "grey_dilation_ragged",
    "grey_opening_ragged",
    "grey_closing_ragged",
    "RaggedBinaryMorphology",
    "binary_erosion_ragged",
    "binary_dilation_ragged",
    "binary_opening_ragged",
    "binary_closing_ragged",
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

_SCHEDULES = {"persistent": 0, "per_level": 1, "per_pass": 2}


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
        schedule: ``"per_pass"`` (default: one launch per pass level and kind
            of pass, one line per lane, row-staged passes along the contiguous
            axis), ``"persistent"`` (one launch, device dependency counters) or
            ``"per_level"`` (one launch per pass level).
        graph: per-pass schedule only -- after the first call (which uploads
            the plan's tables and lets the autotuner pick the row-pass chunk
            depth for this shape class), the plan's launches are captured once
            per input buffer as a HIP graph and replayed.  ``__call__`` then
            copies the masks into one persistent input buffer.
        device: CUDA device of the masks and outputs.
    """

    def __init__(self, shapes: Sequence[Sequence[int]], ops: Sequence[str] = ("edt",),
                 sampling=None, return_distances: bool = True, return_indices: bool = False,
                 index_dtype: torch.dtype = torch.int64, schedule: str = "per_pass",
                 graph: bool = False, device: torch.device | str | None = None):
        if not shapes:
            raise ValueError("at least one mask shape is required")
        if graph and schedule != "per_pass":
            raise ValueError("graph replay is offered for the per_pass schedule only")
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
        self.graph = bool(graph)
        self._graphs: dict = {}       # (input pointer, dtype) -> captured launches
        self._inbuf: Tensor | None = None

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
            if self.graph and self._uploaded:
                key = (flat.data_ptr(), flat.dtype)
                g = self._graphs.get(key)
                if g is None:
                    g = torch.cuda.CUDAGraph()
                    side = torch.cuda.Stream(device=self.device)
                    side.wait_stream(torch.cuda.current_stream(self.device))
                    with torch.cuda.graph(g, stream=side):
                        self._launch(L, flat, dt)
                    torch.cuda.current_stream(self.device).wait_stream(side)
                    self._graphs[key] = g
                g.replay()
            else:
                self._launch(L, flat, dt)
            if check:
                st = _lib.stream_of(self.device)
                _lib.check(L.tm_ragged_dt_check(_lib.ptr(self.workspace), st), "tm_ragged_dt")

    def _launch(self, L, flat: Tensor, dt: int) -> None:
        st = _lib.stream_of(self.device)
        _lib.check(L.tm_ragged_dt(
            self._jobs, self._njobs, _lib.ptr(flat), dt, _lib.ptr(self.distances),
            _lib.ptr(self.indices),
            _lib.TM_I64 if self.index_dtype == torch.int64 else _lib.TM_I32,
            self.schedule, 0 if self._uploaded else 1, _lib.ptr(self.workspace),
            self.workspace.numel(), st), "tm_ragged_dt")
        self._uploaded = True

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
        if self.graph and not isinstance(masks, Tensor):
            # one persistent input buffer, so the captured graph is reused
            if self._inbuf is None or self._inbuf.dtype != flat.dtype:
                self._inbuf = torch.empty_like(flat)
            self._inbuf.copy_(flat)
            flat = self._inbuf
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


# operation -> dilation flag of each pass (opening = dilation of the erosion,
# closing = erosion of the dilation, reference grey.py:162-197)
_GREY_PASSES = {"erosion": (0,), "dilation": (1,), "opening": (0, 1), "closing": (1, 0)}


class RaggedGreyMorphology:
    """Grey morphology of a fixed list of item shapes, with its outputs and
    workspace.

    Args:
        shapes: spatial shape of every item (1..3 axes, the same number for
            every item; no batch or channel axes).
        operation: ``"erosion"``, ``"dilation"``, ``"opening"`` or ``"closing"``.
        size, footprint, structure, mode, cval, origin: as ``grey_erosion``
            (structure > footprint > size; reference grey.py:42-89), for items
            of that many axes.
        device: CUDA device of the items and outputs.
    """

    def __init__(self, shapes: Sequence[Sequence[int]], operation: str = "erosion", size=None,
                 footprint=None, structure=None, mode: str = "reflect", cval: float = 0.0,
                 origin=0, device: torch.device | str | None = None):
        if not shapes:
            raise ValueError("at least one item shape is required")
        if operation not in _GREY_PASSES:
            raise ValueError(f"operation must be one of {sorted(_GREY_PASSES)}")
        self.shapes = [tuple(int(v) for v in sh) for sh in shapes]
        nd = len(self.shapes[0])
        if not 1 <= nd <= 3:
            raise ValueError(f"ragged grey morphology takes items of 1..3 axes, got {nd}")
        for sh in self.shapes:
            if len(sh) != nd:
                raise ValueError(f"every item must have {nd} axes (one structuring element "
                                 f"serves the list); got shape {sh}")
            if any(v < 1 for v in sh):
                raise ValueError(f"item shape {sh} has an empty axis")
        dev = torch.device(device) if device is not None else torch.device("cuda")
        if dev.type != "cuda":
            raise ValueError("ragged grey morphology runs on CUDA devices only")
        self.device = dev if dev.index is not None else torch.device(
            "cuda", torch.cuda.current_device())
        self.operation = operation
        struct, fp, origin_t, mode_int, se_shape = _resolve_element_nd(
            nd, size, footprint, structure, mode, origin)
        # host arrays handed to the library on every call
        self._struct = None if struct is None else struct.numpy()
        self._fp = None if fp is None else fp.to(torch.uint8).numpy()
        self._se_shape = _lib.i64_array(se_shape)
        self._origin = _lib.i64_array(origin_t)
        self._mode = mode_int
        self._cval = float(cval)

        self.offsets = [0]
        for sh in self.shapes:
            self.offsets.append(self.offsets[-1] + math.prod(sh))
        jobs = (TmRaggedJob * len(self.shapes))()
        for k, sh in enumerate(self.shapes):
            j = jobs[k]
            j.op = 0
            j.ndim = nd
            for a, v in enumerate(sh):
                j.shape[a] = v
            j.in_offset = self.offsets[k]
            j.out_offset = self.offsets[k]
            j.idx_offset = -1
        self._jobs = jobs
        L = _lib.lib()
        need = int(L.tm_ragged_grey_workspace_bytes(jobs, len(jobs)))
        if need == 0:
            raise RuntimeError(f"tm_ragged_grey_workspace_bytes: {_lib.last_error()}")
        self.workspace = torch.empty(need, dtype=torch.uint8, device=self.device)
        self.output = torch.empty(self.numel, dtype=torch.float32, device=self.device)
        self._mid = torch.empty(self.numel, dtype=torch.float32, device=self.device) \
            if len(_GREY_PASSES[operation]) == 2 else None
        self._uploaded = False

    @property
    def numel(self) -> int:
        return self.offsets[-1]

    def pack(self, items: Sequence[Tensor]) -> Tensor:
        """Flat float32 buffer of the items in list order."""
        if len(items) != len(self.shapes):
            raise ValueError(f"expected {len(self.shapes)} items, got {len(items)}")
        for k, x in enumerate(items):
            if tuple(x.shape) != self.shapes[k]:
                raise ValueError(f"item {k} has shape {tuple(x.shape)}, expected {self.shapes[k]}")
        return torch.cat([x.detach().reshape(-1).float() for x in items]).to(self.device)

    def _pass(self, src: Tensor, dst: Tensor, dilation: int, st) -> None:
        def arr(a):
            return None if a is None else a.ctypes.data_as(ctypes.c_void_p)

        L = _lib.lib()
        _lib.check(L.tm_ragged_grey(
            self._jobs, len(self._jobs), _lib.ptr(src), arr(self._struct), arr(self._fp),
            self._se_shape, self._origin, self._mode, self._cval, dilation, _lib.ptr(dst),
            0 if self._uploaded else 1, _lib.ptr(self.workspace), self.workspace.numel(), st),
            "tm_ragged_grey")
        self._uploaded = True

    def run(self, flat: Tensor) -> Tensor:
        """Runs the operation on a packed float32 input; returns ``self.output``."""
        if not flat.is_cuda or flat.device != self.device:
            raise ValueError(f"input must be on {self.device}")
        if flat.dtype != torch.float32 or flat.dim() != 1 or flat.numel() != self.numel \
                or not flat.is_contiguous():
            raise ValueError(f"packed input must be a contiguous 1-D float32 tensor of "
                             f"{self.numel} elements")
        passes = _GREY_PASSES[self.operation]
        with torch.cuda.device(self.device):
            st = _lib.stream_of(self.device)
            if len(passes) == 1:
                self._pass(flat, self.output, passes[0], st)
            else:
                self._pass(flat, self._mid, passes[0], st)
                self._pass(self._mid, self.output, passes[1], st)
        return self.output

    def results(self) -> list[Tensor]:
        """One float32 view of ``self.output`` per item."""
        return [self.output[self.offsets[k]:self.offsets[k + 1]].view(sh)
                for k, sh in enumerate(self.shapes)]

    def __call__(self, items: Sequence[Tensor] | Tensor) -> list[Tensor]:
        flat = items if isinstance(items, Tensor) else self.pack(items)
        self.run(flat)
        return self.results()


def _grey_ragged(items, operation, size, footprint, structure, mode, cval, origin):
    if not items:
        raise ValueError("at least one item is required")
    plan = RaggedGreyMorphology([x.shape for x in items], operation, size, footprint, structure,
                                mode, cval, origin, device=items[0].device)
    return plan(items)


def grey_erosion_ragged(items: Sequence[Tensor], size=None, footprint=None, structure=None,
                        mode: str = "reflect", cval: float = 0.0, origin=0) -> list[Tensor]:
    """Grey erosion of every item of a ragged list (spatial shapes only) in one
    launch; one float32 result per item, as ``grey_erosion`` returns it."""
    return _grey_ragged(items, "erosion", size, footprint, structure, mode, cval, origin)


def grey_dilation_ragged(items: Sequence[Tensor], size=None, footprint=None, structure=None,
                         mode: str = "reflect", cval: float = 0.0, origin=0) -> list[Tensor]:
    """Grey dilation of every item of a ragged list in one launch."""
    return _grey_ragged(items, "dilation", size, footprint, structure, mode, cval, origin)


def grey_opening_ragged(items: Sequence[Tensor], size=None, footprint=None, structure=None,
                        mode: str = "reflect", cval: float = 0.0, origin=0) -> list[Tensor]:
    """Grey opening (dilation of the erosion) of every item, two launches."""
    return _grey_ragged(items, "opening", size, footprint, structure, mode, cval, origin)


def grey_closing_ragged(items: Sequence[Tensor], size=None, footprint=None, structure=None,
                        mode: str = "reflect", cval: float = 0.0, origin=0) -> list[Tensor]:
    """Grey closing (erosion of the dilation) of every item, two launches."""
    return _grey_ragged(items, "closing", size, footprint, structure, mode, cval, origin)


# operation -> dilation flag of each block of `iterations` steps (reference
# binary.py:91-136, 231-290)
_BINARY_BLOCKS = {"erosion": (0,), "dilation": (1,), "opening": (0, 1), "closing": (1, 0)}


class RaggedBinaryMorphology:
    """Binary erosion / dilation / opening / closing of a fixed list of item
    shapes, every step of every item in one launch.

    Args:
        shapes: spatial shape of every item (1..3 axes, the same number for
            every item).
        operation: ``"erosion"``, ``"dilation"``, ``"opening"`` or ``"closing"``.
        structure: boolean structuring element (default: the cross of
            ``generate_binary_structure(ndim, 1)``).
        iterations: steps per erosion / dilation, >= 1 (repeating until stable
            and ``mask`` are not offered on ragged lists).
        border_value, origin: as ``binary_erosion``.
    """

    def __init__(self, shapes: Sequence[Sequence[int]], operation: str = "erosion",
                 structure: Tensor | None = None, iterations: int = 1,
                 border_value: bool = False, origin=0,
                 device: torch.device | str | None = None):
        from .morphology.binary import _normalize_structure
        from .morphology.structure import generate_binary_structure

        if operation not in _BINARY_BLOCKS:
            raise ValueError(f"operation must be one of {sorted(_BINARY_BLOCKS)}")
        if int(iterations) < 1:
            raise ValueError("ragged binary morphology needs iterations >= 1")
        if not shapes:
            raise ValueError("at least one item shape is required")
        nd = len(shapes[0])
        if structure is None:
            structure = generate_binary_structure(nd, 1)
        structure = _normalize_structure(structure, nd)
        self.operation = operation
        self.iterations = int(iterations)
        self.empty = not bool(structure.any().item())
        footprint = structure if not self.empty else torch.ones((1,) * nd, dtype=torch.bool)
        self._plan = RaggedGreyMorphology(shapes, "erosion", footprint=footprint,
                                          mode="constant", cval=1.0 if border_value else 0.0,
                                          origin=origin if not self.empty else 0, device=device)
        self.shapes = self._plan.shapes
        self.device = self._plan.device
        self._buf = torch.empty_like(self._plan.output)

    def pack(self, items: Sequence[Tensor]) -> Tensor:
        """Flat float32 0/1 buffer of the items in list order."""
        return self._plan.pack([(x != 0) for x in items])

    def run(self, flat: Tensor) -> Tensor:
        """Runs the operation on a packed 0/1 input; returns the flat float32
        0/1 result."""
        steps = [dil for dil in _BINARY_BLOCKS[self.operation] for _ in range(self.iterations)]
        if self.empty:
            # every step yields a constant (reference binary.py:54-58)
            self._plan.output.fill_(0.0 if steps[-1] else 1.0)
            return self._plan.output
        # step k writes bufs[(n - 1 - k) % 2]: the last step lands in the output
        bufs = (self._plan.output, self._buf)
        n = len(steps)
        src = flat
        with torch.cuda.device(self.device):
            st = _lib.stream_of(self.device)
            for k, dil in enumerate(steps):
                dst = bufs[(n - 1 - k) % 2]
                self._plan._pass(src, dst, dil, st)
                src = dst
        return src

    def __call__(self, items: Sequence[Tensor] | Tensor) -> list[Tensor]:
        flat = items if isinstance(items, Tensor) else self.pack(items)
        out = self.run(flat)
        return [out[self._plan.offsets[k]:self._plan.offsets[k + 1]].view(sh) != 0
                for k, sh in enumerate(self.shapes)]


def _binary_ragged(items, operation, structure, iterations, border_value, origin):
    if not items:
        raise ValueError("at least one item is required")
    plan = RaggedBinaryMorphology([x.shape for x in items], operation, structure, iterations,
                                  border_value, origin, device=items[0].device)
    return plan(items)


def binary_erosion_ragged(items: Sequence[Tensor], structure: Tensor | None = None,
                          iterations: int = 1, border_value: bool = False,
                          origin=0) -> list[Tensor]:
    """Binary erosion of every item of a ragged list; one bool result per item,
    as ``binary_erosion`` returns it for that item."""
    return _binary_ragged(items, "erosion", structure, iterations, border_value, origin)


def binary_dilation_ragged(items: Sequence[Tensor], structure: Tensor | None = None,
                           iterations: int = 1, border_value: bool = False,
                           origin=0) -> list[Tensor]:
    """Binary dilation of every item of a ragged list."""
    return _binary_ragged(items, "dilation", structure, iterations, border_value, origin)


def binary_opening_ragged(items: Sequence[Tensor], structure: Tensor | None = None,
                          iterations: int = 1, border_value: bool = False,
                          origin=0) -> list[Tensor]:
    """Binary opening (erosions, then dilations) of every item of a ragged list."""
    return _binary_ragged(items, "opening", structure, iterations, border_value, origin)


def binary_closing_ragged(items: Sequence[Tensor], structure: Tensor | None = None,
                          iterations: int = 1, border_value: bool = False,
                          origin=0) -> list[Tensor]:
    """Binary closing (dilations, then erosions) of every item of a ragged list."""
    return _binary_ragged(items, "closing", structure, iterations, border_value, origin)
