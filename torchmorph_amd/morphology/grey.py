"""Grey-scale morphology on (B, C, spatial...) GPU tensors.

Drop-in for reference torchmorph/morphology/grey.py:29-289 (same names,
arguments, structuring-element resolution, border modes and messages).  The
kernels read float32 input directly (other dtypes are converted, like the
reference's ``.float()``), write float32 and are bit-identical to SciPy's
ndimage on every border mode.

``differentiable=True`` (an opt-in the reference does not have) routes
erosion/dilation through hand-written backward kernels; the default keeps the
reference's non-differentiable behaviour (reference test_grey.py:218-222).
"""
from __future__ import annotations

import torch
from torch import Tensor

from .. import _lib
from .._validation import validate_bcs_input, validate_output
from .structure import check_origin, origin_tuple as _origin_tuple

_MODE_MAP = {"constant": 0, "reflect": 1, "nearest": 2, "mirror": 3, "wrap": 4}


def _element_kwargs(size, footprint, structure, mode, cval, origin) -> dict:
    return dict(size=size, footprint=footprint, structure=structure, mode=mode, cval=cval,
                origin=origin)


def _resolve_element(input: Tensor, size, footprint, structure, mode, origin):
    """Reference grey.py:42-89: structure > footprint > size, flat when only a
    size or footprint is given; returns CPU (structure f32, footprint u8 or
    None, origin tuple, mode int)."""
    return _resolve_element_nd(validate_bcs_input(input), size, footprint, structure, mode,
                               origin)


def _resolve_element_nd(spatial_ndim: int, size, footprint, structure, mode, origin):
    """``_resolve_element`` for items of ``spatial_ndim`` axes."""
    footprint_cpu = None
    if structure is not None:
        struct = structure.detach().to(device="cpu", dtype=torch.float32)
        if footprint is not None:
            mask = footprint.to(torch.bool)
            if tuple(mask.shape) != tuple(struct.shape):
                raise ValueError(f"Footprint shape {tuple(mask.shape)} must match structure "
                                 f"shape {tuple(struct.shape)}.")
            if not mask.any():
                raise ValueError("All-zero footprint is not supported.")
            footprint_cpu = mask.detach().to("cpu").contiguous()
    elif footprint is not None:
        mask = footprint.to(torch.bool)
        if not mask.any():
            raise ValueError("All-zero footprint is not supported.")
        struct = torch.zeros(tuple(mask.shape), dtype=torch.float32)
        footprint_cpu = mask.detach().to("cpu").contiguous()
    elif size is not None:
        if isinstance(size, int):
            size = (size,) * spatial_ndim
        if any(v <= 0 for v in size):
            raise ValueError("size values must be greater than zero")
        struct = torch.zeros(tuple(size), dtype=torch.float32)
    else:
        raise ValueError("At least one of size, footprint, or structure must be specified.")
    if struct.ndim != spatial_ndim:
        raise ValueError(f"Structure dimension {struct.ndim} must match spatial dimension "
                         f"{spatial_ndim}.")
    origin_tuple = _origin_tuple(origin, spatial_ndim)
    check_origin(origin_tuple, struct)
    if mode not in _MODE_MAP:
        raise ValueError(f"Unknown mode '{mode}'. Must be one of {list(_MODE_MAP.keys())}.")
    flat_box = structure is None and footprint_cpu is None
    return (None if flat_box else struct.contiguous(), footprint_cpu, origin_tuple,
            _MODE_MAP[mode], tuple(struct.shape))


def _launch(x: Tensor, struct, fp, se_shape, origin, mode_int, cval, dilation: bool,
            out: Tensor) -> None:
    L = _lib.lib()
    nd = x.ndim - 2
    batch = x.shape[0] * x.shape[1]
    shp = _lib.i64_array(x.shape[2:])
    ses = _lib.i64_array(se_shape)
    st_ptr = None
    fp_ptr = None
    st_np = fp_np = None
    if struct is not None:
        st_np = struct.numpy()
        st_ptr = st_np.ctypes.data_as(_lib.ctypes.c_void_p)
    if fp is not None:
        fp_np = fp.to(torch.uint8).numpy()
        fp_ptr = fp_np.ctypes.data_as(_lib.ctypes.c_void_p)
    nbytes = L.tm_grey_workspace_bytes(batch, nd, shp, ses)
    org = _lib.i64_array(origin)
    with torch.cuda.device(x.device):
        stream = _lib.stream_of(x.device)
        # the fused kernels need no workspace; the library reports
        # TM_ERR_WORKSPACE (-3) before launching anything when a fallback does
        status = L.tm_grey(_lib.ptr(x), batch, nd, shp, st_ptr, fp_ptr, ses, org, mode_int,
                           float(cval), int(dilation), _lib.ptr(out), None, 0, stream)
        if status == -3:
            ws = _lib.workspace(nbytes, x.device)
            status = L.tm_grey(_lib.ptr(x), batch, nd, shp, st_ptr, fp_ptr, ses, org, mode_int,
                               float(cval), int(dilation), _lib.ptr(out), _lib.ptr(ws), nbytes,
                               stream)
        _lib.check(status, "tm_grey")


def _host_element(struct, fp):
    st_np = None if struct is None else struct.numpy()
    fp_np = None if fp is None else fp.to(torch.uint8).numpy()
    return st_np, fp_np


def _cptr(a):
    return None if a is None else a.ctypes.data_as(_lib.ctypes.c_void_p)


def _fused_ok(nd, struct, fp, se_shape, origin) -> bool:
    st_np, fp_np = _host_element(struct, fp)
    L = _lib.lib()
    return all(L.tm_grey_fused_ok(nd, _cptr(st_np), _cptr(fp_np), _lib.i64_array(se_shape),
                                  _lib.i64_array(origin), dil) == 1 for dil in (0, 1))


def _launch_ex(x: Tensor, struct, fp, se_shape, origin, mode_int, cval, dilation: bool, epi: int,
               a: Tensor, b: Tensor | None, out: Tensor) -> None:
    """Erosion/dilation of x with a composite's final arithmetic fused into the
    store (tm_grey_ex): epi 1 -> acc - a, 2 -> a - acc, 3 -> (acc + a) - 2 b."""
    L = _lib.lib()
    st_np, fp_np = _host_element(struct, fp)
    with torch.cuda.device(x.device):
        _lib.check(L.tm_grey_ex(_lib.ptr(x), x.shape[0] * x.shape[1], x.ndim - 2,
                                _lib.i64_array(x.shape[2:]), _cptr(st_np), _cptr(fp_np),
                                _lib.i64_array(se_shape), _lib.i64_array(origin), mode_int,
                                float(cval), int(dilation), epi, _lib.ptr(a),
                                None if b is None else _lib.ptr(b), _lib.ptr(out),
                                _lib.stream_of(x.device)), "tm_grey_ex")


def _composite(input: Tensor, output, kw, first: bool, second: bool, second_on_first: bool,
               epi: int, a_is_first: bool, b_is_input: bool, fallback, one_pass: int = 0):
    """A composite in two launches: pass `first` (False erosion, True
    dilation) of the input into a temporary, then pass `second` -- of the
    temporary when `second_on_first`, else of the input -- with the epilogue
    fused into its store; epilogue operand a is the temporary or the input,
    b the input.  `one_pass` (4 gradient, 5 laplace): a flat box of odd sizes
    with zero origin runs in a single pass over the input instead.  Falls back
    to the plain composition (the same kernels and tensor arithmetic) where
    the fused kernels do not apply."""
    validate_bcs_input(input)
    validate_output(input, output)
    struct, fp, origin_t, mode_int, se_shape = _resolve_element(
        input, kw["size"], kw["footprint"], kw["structure"], kw["mode"], kw["origin"])
    nd = input.ndim - 2
    if not _fused_ok(nd, struct, fp, se_shape, origin_t):
        return _finish(fallback(), output)
    x = input.detach()
    x = x.contiguous() if x.dtype == torch.float32 else x.float().contiguous()
    if one_pass and struct is None and fp is None and _lib.lib().tm_grey_dual_ok(
            nd, _lib.i64_array(se_shape), _lib.i64_array(origin_t)) == 1:
        result = _result_buffer(x, output)
        _launch_ex(x, None, None, se_shape, origin_t, mode_int, kw["cval"], True, one_pass, x,
                   None, result)
        return _finish(result, output)
    tmp = torch.empty_like(x)
    _launch(x, struct, fp, se_shape, origin_t, mode_int, kw["cval"], first, tmp)
    result = _result_buffer(x, output)
    _launch_ex(tmp if second_on_first else x, struct, fp, se_shape, origin_t, mode_int, kw["cval"],
               second, epi, tmp if a_is_first else x, x if b_is_input else None, result)
    return _finish(result, output)


def _grey_morphology(input: Tensor, size=None, footprint=None, structure=None, output=None,
                     mode: str = "reflect", cval: float = 0.0, origin=0, *, operation: str,
                     differentiable: bool = False) -> Tensor:
    validate_bcs_input(input)
    validate_output(input, output)
    struct, fp, origin_t, mode_int, se_shape = _resolve_element(input, size, footprint,
                                                                structure, mode, origin)
    if differentiable and input.requires_grad:
        from .grad import grey_autograd

        result = grey_autograd(input, struct, fp, se_shape, origin_t, mode_int, cval,
                               operation == "dilation")
    else:
        x = input.detach()
        x = x.contiguous() if x.dtype == torch.float32 else x.float().contiguous()
        if _writable_in_place(output, x):
            _launch(x, struct, fp, se_shape, origin_t, mode_int, cval, operation == "dilation",
                    output)
            return output
        result = torch.empty(x.shape, dtype=torch.float32, device=x.device)
        _launch(x, struct, fp, se_shape, origin_t, mode_int, cval, operation == "dilation",
                result)
    if output is not None:
        output.copy_(result)
        return output
    return result


def _writable_in_place(output, x) -> bool:
    """A preallocated float32 contiguous output that shares no storage with
    the input is written by the kernels directly (no copy pass)."""
    return (output is not None and output.dtype == torch.float32 and output.is_contiguous()
            and output.untyped_storage().data_ptr() != x.untyped_storage().data_ptr())


def grey_erosion(input: Tensor, size=None, footprint=None, structure=None, output=None,
                 mode: str = "reflect", cval: float = 0.0, origin=0,
                 differentiable: bool = False) -> Tensor:
    """N-D grey erosion (reference grey.py:108-132); float32 result."""
    return _grey_morphology(input, size, footprint, structure, output, mode, cval, origin,
                            operation="erosion", differentiable=differentiable)


def grey_dilation(input: Tensor, size=None, footprint=None, structure=None, output=None,
                  mode: str = "reflect", cval: float = 0.0, origin=0,
                  differentiable: bool = False) -> Tensor:
    """N-D grey dilation (reference grey.py:135-159); float32 result."""
    return _grey_morphology(input, size, footprint, structure, output, mode, cval, origin,
                            operation="dilation", differentiable=differentiable)


def grey_opening(input: Tensor, size=None, footprint=None, structure=None, output=None,
                 mode: str = "reflect", cval: float = 0.0, origin=0) -> Tensor:
    """Dilation of the erosion (reference grey.py:162-178)."""
    kw = _element_kwargs(size, footprint, structure, mode, cval, origin)
    return grey_dilation(grey_erosion(input, **kw), output=output, **kw)


def grey_closing(input: Tensor, size=None, footprint=None, structure=None, output=None,
                 mode: str = "reflect", cval: float = 0.0, origin=0) -> Tensor:
    """Erosion of the dilation (reference grey.py:181-197)."""
    kw = _element_kwargs(size, footprint, structure, mode, cval, origin)
    return grey_erosion(grey_dilation(input, **kw), output=output, **kw)


def _finish(result, output):
    if output is not None and result is not output:
        output.copy_(result)
        return output
    return result


def _disjoint(a: Tensor, b: Tensor) -> bool:
    a0, b0 = a.data_ptr(), b.data_ptr()
    return (a0 + a.numel() * a.element_size() <= b0
            or b0 + b.numel() * b.element_size() <= a0)


def _result_buffer(x: Tensor, output) -> Tensor:
    """Where a composite's last pass writes: the caller's output itself when
    it is contiguous float32 memory apart from the input (the passes read the
    input to the end), else a new tensor that _finish copies into it."""
    if (output is not None and output.dtype == torch.float32 and output.is_contiguous()
            and output.shape == x.shape and _disjoint(output, x)):
        return output
    return torch.empty_like(x)


def morphological_gradient(input: Tensor, size=None, footprint=None, structure=None, output=None,
                           mode: str = "reflect", cval: float = 0.0, origin=0) -> Tensor:
    """dilation - erosion (reference grey.py:200-221): one pass for a flat box
    of odd sizes, else the subtraction fused into the dilation's store."""
    kw = _element_kwargs(size, footprint, structure, mode, cval, origin)
    return _composite(input, output, kw, first=False, second=True, second_on_first=False, epi=1,
                      a_is_first=True, b_is_input=False,
                      fallback=lambda: grey_dilation(input, **kw) - grey_erosion(input, **kw),
                      one_pass=4)


def white_tophat(input: Tensor, size=None, footprint=None, structure=None, output=None,
                 mode: str = "reflect", cval: float = 0.0, origin=0) -> Tensor:
    """input - opening (reference grey.py:224-243), the subtraction fused
    into the opening's dilation."""
    kw = _element_kwargs(size, footprint, structure, mode, cval, origin)
    return _composite(input, output, kw, first=False, second=True, second_on_first=True, epi=2,
                      a_is_first=False, b_is_input=False,
                      fallback=lambda: input.detach().float() - grey_opening(input, **kw))


def black_tophat(input: Tensor, size=None, footprint=None, structure=None, output=None,
                 mode: str = "reflect", cval: float = 0.0, origin=0) -> Tensor:
    """closing - input (reference grey.py:246-265), the subtraction fused
    into the closing's erosion."""
    kw = _element_kwargs(size, footprint, structure, mode, cval, origin)
    return _composite(input, output, kw, first=True, second=False, second_on_first=True, epi=1,
                      a_is_first=False, b_is_input=False,
                      fallback=lambda: grey_closing(input, **kw) - input.detach().float())


def morphological_laplace(input: Tensor, size=None, footprint=None, structure=None, output=None,
                          mode: str = "reflect", cval: float = 0.0, origin=0) -> Tensor:
    """dilation + erosion - 2 input (reference grey.py:268-289): one pass for
    a flat box of odd sizes, else the sums fused into the dilation's store."""
    kw = _element_kwargs(size, footprint, structure, mode, cval, origin)
    return _composite(input, output, kw, first=False, second=True, second_on_first=False, epi=3,
                      a_is_first=True, b_is_input=True,
                      fallback=lambda: (grey_dilation(input, **kw) + grey_erosion(input, **kw)
                                        - 2 * input.detach().float()),
                      one_pass=5)
