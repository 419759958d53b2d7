"""ctypes binding of libtorchmorph_hip.so (the C ABI in include/torchmorph_hip.h).

torch is imported before the library is opened so that the library's
``libamdhip64.so.7`` dependency resolves to the HIP runtime torch already
loaded: device pointers and ``hipStream_t`` handles from torch are then valid
in the library.  There is no fallback: a missing library raises.
"""
from __future__ import annotations

import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtorchmorph_hip.so")

TM_U8, TM_F32, TM_I32, TM_I64, TM_F64 = 0, 1, 2, 3, 4
TM_OK, TM_ERR_ARG, TM_ERR_HIP, TM_ERR_WORKSPACE = 0, -1, -2, -3  # enum tm_status

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_SZ = ctypes.c_size_t
_D = ctypes.c_double
_F = ctypes.c_float
_I64P = ctypes.POINTER(ctypes.c_int64)
_DP = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); every C-ABI entry point is declared here
SIGNATURES: dict[str, tuple] = {
    "tm_last_error": (_I, [ctypes.c_char_p, _SZ]),
    "tm_version": (ctypes.c_char_p, []),
    "tm_autotune_enable": (None, [_I]),
    "tm_autotune_clear": (None, []),
    "tm_autotune_export": (_I, [ctypes.c_char_p, _SZ]),
    "tm_autotune_import": (_I, [ctypes.c_char_p]),
    "tm_edt_workspace_bytes": (_SZ, [_I64, _I, _I64P, _DP]),
    "tm_edt": (_I, [_P, _I, _I64, _I, _I64P, _DP, _P, _P, _I, _P, _SZ, _P]),
    "tm_edt_noalloc": (_I, [_P, _I, _I64, _I, _I64P, _DP, _P, _P, _I, _P]),
    "tm_cdt_workspace_bytes": (_SZ, [_I64, _I, _I64P]),
    "tm_cdt": (_I, [_P, _I, _I64, _I, _I64P, _I, _P, _P, _I, _P, _SZ, _P]),
    "tm_bfdt": (_I, [_P, _I, _I64, _I, _I64P, _DP, _I, _P, _P, _I, _P]),
    "tm_grey_workspace_bytes": (_SZ, [_I64, _I, _I64P, _I64P]),
    "tm_grey": (_I, [_P, _I64, _I, _I64P, _P, _P, _I64P, _I64P, _I, _F, _I, _P, _P, _SZ, _P]),
    "tm_sinkhorn_hist": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _I64, _I, _I, _D,
                              ctypes.POINTER(ctypes.c_int), _P]),
    "tm_sinkhorn_hist_workspace_bytes": (_SZ, [_I, _I64, _I, _D, _I]),
    "tm_sinkhorn_hist_solve": (_I, [_I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I64, _I, _I, _D, _D,
                                    _P, _P, _P, ctypes.POINTER(ctypes.c_int),
                                    ctypes.POINTER(ctypes.c_int), _P, _SZ, _P]),
    "tm_sinkhorn_hist_plan": (_I, [_I, _I, _P, _P, _P, _I64, _I, _D, _P, _P]),
    "tm_sinkhorn_points_workspace_bytes": (_SZ, [_I, ctypes.POINTER(ctypes.c_int32),
                                                 ctypes.POINTER(ctypes.c_int32)]),
    "tm_sinkhorn_points": (_I, [_P, _P, _P, _P, _I, _I, ctypes.POINTER(ctypes.c_int32),
                                ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float),
                                _F, _I, _I, _P, _P, _P, _P, _P, _SZ, _P]),
    "tm_grey_fused_ok": (_I, [_I, _P, _P, _I64P, _I64P, _I]),
    "tm_grey_dual_ok": (_I, [_I, _I64P, _I64P]),
    "tm_grey_ex": (_I, [_P, _I64, _I, _I64P, _P, _P, _I64P, _I64P, _I, _F, _I, _I, _P, _P, _P,
                        _P]),
    "tm_grey_backward_workspace_bytes": (_SZ, [_I64, _I, _I64P]),
    "tm_grey_backward": (_I, [_P, _I64, _I, _I64P, _P, _P, _I64P, _I64P, _I, _F, _I, _P, _P, _P,
                              _SZ, _P]),
    "tm_sinkhorn_points_grad_workspace_bytes": (_SZ, [_I, ctypes.POINTER(ctypes.c_int32),
                                                      ctypes.POINTER(ctypes.c_int32)]),
    "tm_sinkhorn_points_grad": (_I, [_P, _P, _P, _P, _I, _I, ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_int32),
                                     ctypes.POINTER(ctypes.c_float), _P, _P, _P, _P, _P, _SZ,
                                     _P]),
    "tm_sinkhorn_points_planmat": (_I, [_P, _P, _P, _P, _I, _I, ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_float), _P, _P, _I, _I, _P, _P,
                                        _P, _SZ, _P]),
    "tm_mlp_gemm_bf16": (_I, [_P, _I64, _I64, _P, _I64, _I64, _P, _I, _I, _I, _P, _SZ, _P]),
    "tm_cast_bf16": (_I, [_P, _P, _I64, _P]),
    "tm_warmstart_param_count": (_I64, []),
    "tm_warmstart_workspace_bytes": (_SZ, [_I, ctypes.POINTER(ctypes.c_int32),
                                           ctypes.POINTER(ctypes.c_int32)]),
    "tm_warmstart_predict": (_I, [_P, _P, _P, _I, ctypes.POINTER(ctypes.c_int32),
                                  ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float),
                                  _P, _P, _SZ, _P]),
    "tm_warmstart_grad": (_I, [_P, _P, _P, _I, ctypes.POINTER(ctypes.c_int32),
                               ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_float), _P,
                               _F, _P, _P, _P, _SZ, _P]),
    "tm_warmstart_adam": (_I, [_P, _P, _P, _P, _I64, _I, _F, _F, _F, _F, _F, _P]),
    "tm_ragged_dt_workspace_bytes": (_SZ, [_P, _I64]),
    "tm_ragged_dt": (_I, [_P, _I64, _P, _I, _P, _P, _I, _I, _I, _P, _SZ, _P]),
    "tm_ragged_dt_check": (_I, [_P, _P]),
    "tm_ragged_grey_workspace_bytes": (_SZ, [_P, _I64]),
    "tm_ragged_grey": (_I, [_P, _I64, _P, _P, _P, _I64P, _I64P, _I, _F, _I, _P, _I, _P, _SZ, _P]),
    "tm_binary_workspace_bytes": (_SZ, [_I64, _I, _I64P]),
    "tm_binary_fill_holes": (_I, [_P, _I, _I64, _I, _I64P, _P, _I64P, _I64P, _P, _P, _SZ, _P]),
    "tm_binary_hit_or_miss": (_I, [_P, _I, _I64, _I, _I64P, _P, _I64P, _I64P, _P, _I64P, _I64P,
                                   _P, _P, _SZ, _P]),
    "tm_binary": (_I, [_P, _I, _I64, _I, _I64P, _P, _I64P, _I64P, _I, _I,
                       ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), _P, _P, _P,
                       _SZ, _P]),
}

_lib = None


def lib():
    """Load the HIP library (raises RuntimeError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"torchmorph_amd: {LIB_PATH} is missing; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'` or "
            "`python -m torchmorph_amd._build`")
    handle = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_LOCAL)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(handle, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args
    _lib = handle
    return _lib


def last_error() -> str:
    buf = ctypes.create_string_buffer(1024)
    lib().tm_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int, what: str) -> None:
    if status != 0:
        raise RuntimeError(f"{what}: {last_error()} (status {status})")


def i64_array(values) -> ctypes.Array:
    vals = [int(v) for v in values]
    return (ctypes.c_int64 * max(1, len(vals)))(*vals)


def f64_array(values) -> ctypes.Array:
    vals = [float(v) for v in values]
    return (ctypes.c_double * max(1, len(vals)))(*vals)


def ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_of(device: torch.device) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def workspace(nbytes: int, device: torch.device) -> torch.Tensor | None:
    if nbytes <= 0:
        return None
    return torch.empty(int(nbytes), dtype=torch.uint8, device=device)


# Per-(device, stream) workspaces of the hot entry points.  Calls on one
# stream run in order, so they can share one buffer; a call on another stream
# gets its own.  A buffer is never freed once handed out (a captured graph may
# still reference it): growth keeps the old one and allocates double.
_STREAM_WS: dict = {}
_STREAM_WS_KEEP: list = []


def stream_workspace(nbytes: int, device: torch.device, stream) -> torch.Tensor | None:
    """A workspace of at least `nbytes` reused across calls on `stream`
    (a torch.cuda.Stream)."""
    if nbytes <= 0:
        return None
    key = (device.index, stream.cuda_stream)
    ws = _STREAM_WS.get(key)
    if ws is None or ws.numel() < nbytes:
        if ws is not None:
            _STREAM_WS_KEEP.append(ws)
        size = max(int(nbytes), 2 * (ws.numel() if ws is not None else 0))
        ws = torch.empty(size, dtype=torch.uint8, device=device)
        _STREAM_WS[key] = ws
    return ws


def mask_operand(x: torch.Tensor) -> tuple[torch.Tensor, int]:
    """Contiguous device view of a mask for the kernels: bool/uint8 are read
    as bytes, float32 directly; any other dtype is reduced to ``x != 0``."""
    if x.dtype == torch.bool:
        x = x.contiguous().view(torch.uint8)
        return x, TM_U8
    if x.dtype == torch.uint8:
        return x.contiguous(), TM_U8
    if x.dtype == torch.float32:
        return x.contiguous(), TM_F32
    return (x != 0).contiguous().view(torch.uint8), TM_U8
