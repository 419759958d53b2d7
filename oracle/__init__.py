"""CPU parity oracle for the torchmorph operator path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this package, and only as the checker or as the timed
CPU baseline.  The shipped package (``torchmorph_amd``) never imports it.

Contents:
  * ctypes bindings to ``libtm_oracle.so`` (built from ``tm_oracle.c`` by
    ``make -C oracle``): EDT, CDT, brute-force DT, grey and binary
    morphology, one item at a time, looped over the leading (B, C) axes here;
  * numpy FP64 restatements of the Sinkhorn iterations
    (reference ``torchmorph/optimal_transport.py:71-125``) and of the
    point-cloud log-domain solver the build adds.

Parity pins (what each function is checked against) are listed in
``tests/test_oracle_*.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtm_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile ``libtm_oracle.so`` with gcc (make -C oracle)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64p = ctypes.POINTER(ctypes.c_int64)
        lib.tmo_edt.argtypes = [P, ctypes.c_int, i64p, P, P, P]
        lib.tmo_cdt.argtypes = [P, ctypes.c_int, i64p, ctypes.c_int, P, P]
        lib.tmo_grey.argtypes = [P, ctypes.c_int, i64p, P, P, i64p, i64p, ctypes.c_int,
                                 ctypes.c_float, ctypes.c_int, P]
        lib.tmo_binary.argtypes = [P, ctypes.c_int, i64p, P, i64p, i64p, ctypes.c_int,
                                   ctypes.c_int, P]
        lib.tmo_bfdt.argtypes = [P, ctypes.c_int, i64p, P, ctypes.c_int, P, P]
        for f in (lib.tmo_edt, lib.tmo_cdt, lib.tmo_grey, lib.tmo_binary, lib.tmo_bfdt):
            f.restype = ctypes.c_int
        _lib = lib
    return _lib


def _i64(seq):
    arr = (ctypes.c_int64 * len(seq))(*[int(v) for v in seq])
    return arr


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def _items(x: np.ndarray, nlead: int):
    lead = x.shape[:nlead]
    return lead, x.reshape((-1,) + x.shape[nlead:])


# --------------------------------------------------------------------- EDT --
def edt(mask: np.ndarray, sampling=None, lead: int = 2):
    """Exact EDT of every (B, C) item.  Returns (float64 distances,
    int32 indices with shape (ndim, *mask.shape))."""
    lib = _load()
    mask = np.ascontiguousarray(mask != 0, dtype=np.uint8)
    lead_shape, items = _items(mask, lead)
    spatial = items.shape[1:]
    nd = len(spatial)
    samp = None
    if sampling is not None:
        samp = np.ascontiguousarray(np.broadcast_to(np.asarray(sampling, np.float64), (nd,)))
    dist = np.empty(items.shape, np.float64)
    feat = np.empty((items.shape[0], nd) + spatial, np.int32)
    shp = _i64(spatial)
    for i in range(items.shape[0]):
        it = np.ascontiguousarray(items[i])
        fi = np.empty((nd,) + spatial, np.int32)
        di = np.empty(spatial, np.float64)
        rc = lib.tmo_edt(_ptr(it), nd, shp, _ptr(samp) if samp is not None else None,
                         _ptr(fi), _ptr(di))
        assert rc == 0
        dist[i] = di
        feat[i] = fi
    dist = dist.reshape(lead_shape + spatial)
    feat = np.moveaxis(feat, 1, 0).reshape((nd,) + lead_shape + spatial)
    return dist, feat


def bfdt(mask: np.ndarray, metric: str = "euclidean", sampling=None, lead: int = 2):
    """Brute-force DT of every item (reference bfdt semantics).  Returns
    (float64 distances, int32 indices with shape (ndim, *mask.shape))."""
    lib = _load()
    code = {"euclidean": 0, "taxicab": 1, "chessboard": 2}[metric]
    mask = np.ascontiguousarray(mask != 0, dtype=np.uint8)
    lead_shape, items = _items(mask, lead)
    spatial = items.shape[1:]
    nd = len(spatial)
    samp = None
    if sampling is not None:
        samp = np.ascontiguousarray(np.broadcast_to(np.asarray(sampling, np.float64), (nd,)))
    dist = np.empty(items.shape, np.float64)
    feat = np.empty((items.shape[0], nd) + spatial, np.int32)
    shp = _i64(spatial)
    for i in range(items.shape[0]):
        it = np.ascontiguousarray(items[i])
        fi = np.empty((nd,) + spatial, np.int32)
        di = np.empty(spatial, np.float64)
        rc = lib.tmo_bfdt(_ptr(it), nd, shp, _ptr(samp) if samp is not None else None, code,
                          _ptr(di), _ptr(fi))
        assert rc == 0
        dist[i] = di
        feat[i] = fi
    dist = dist.reshape(lead_shape + spatial)
    feat = np.moveaxis(feat, 1, 0).reshape((nd,) + lead_shape + spatial)
    return dist, feat


# --------------------------------------------------------------------- CDT --
def cdt(mask: np.ndarray, metric: str = "chessboard", lead: int = 2):
    """Chamfer DT of every item.  Returns (int32 distances, int32 indices)."""
    lib = _load()
    m = {"chessboard": 0, "taxicab": 1, "cityblock": 1, "manhattan": 1}[metric]
    mask = np.ascontiguousarray(mask != 0, dtype=np.uint8)
    lead_shape, items = _items(mask, lead)
    spatial = items.shape[1:]
    nd = len(spatial)
    dist = np.empty(items.shape, np.int32)
    feat = np.empty((items.shape[0], nd) + spatial, np.int32)
    shp = _i64(spatial)
    for i in range(items.shape[0]):
        it = np.ascontiguousarray(items[i])
        di = np.empty(spatial, np.int32)
        fi = np.empty((nd,) + spatial, np.int32)
        assert lib.tmo_cdt(_ptr(it), nd, shp, m, _ptr(di), _ptr(fi)) == 0
        dist[i] = di
        feat[i] = fi
    return (dist.reshape(lead_shape + spatial),
            np.moveaxis(feat, 1, 0).reshape((nd,) + lead_shape + spatial))


# ---------------------------------------------------------------- grey / binary --
_MODES = {"constant": 0, "reflect": 1, "nearest": 2, "mirror": 3, "wrap": 4}


def grey(x: np.ndarray, structure: np.ndarray | None, footprint: np.ndarray | None,
         se_shape, origin, mode: str = "reflect", cval: float = 0.0,
         dilation: bool = False, lead: int = 2) -> np.ndarray:
    lib = _load()
    x = np.ascontiguousarray(x, dtype=np.float32)
    lead_shape, items = _items(x, lead)
    spatial = items.shape[1:]
    nd = len(spatial)
    st = None if structure is None else np.ascontiguousarray(structure, np.float32)
    fp = None if footprint is None else np.ascontiguousarray(footprint != 0, np.uint8)
    out = np.empty_like(items)
    for i in range(items.shape[0]):
        it = np.ascontiguousarray(items[i])
        oi = np.empty_like(it)
        rc = lib.tmo_grey(_ptr(it), nd, _i64(spatial), _ptr(st) if st is not None else None,
                          _ptr(fp) if fp is not None else None, _i64(se_shape), _i64(origin),
                          _MODES[mode], float(cval), int(dilation), _ptr(oi))
        assert rc == 0
        out[i] = oi
    return out.reshape(lead_shape + spatial)


def binary(x: np.ndarray, structure: np.ndarray, origin, border_value: bool = False,
           dilation: bool = False, iterations: int = 1, mask: np.ndarray | None = None,
           lead: int = 2) -> np.ndarray:
    """Binary erosion/dilation, iterated like reference binary.py:79-82
    (mask re-applied after every step).  iterations < 1 iterates until stable
    (binary.py:67-77)."""
    lib = _load()
    x0 = np.ascontiguousarray(x != 0, dtype=np.uint8)
    lead_shape, items = _items(x0, lead)
    spatial = items.shape[1:]
    nd = len(spatial)
    st = np.ascontiguousarray(structure != 0, np.uint8)
    mk = None
    if mask is not None:
        mk = np.broadcast_to(mask != 0, x0.shape).reshape(items.shape)
    out = np.empty_like(items)
    for i in range(items.shape[0]):
        cur = np.ascontiguousarray(items[i])
        orig = cur.copy()
        prev = None
        k = 0
        while True:
            nxt = np.empty_like(cur)
            if st.any():
                assert lib.tmo_binary(_ptr(cur), nd, _i64(spatial), _ptr(st), _i64(st.shape),
                                      _i64(origin), int(border_value), int(dilation),
                                      _ptr(nxt)) == 0
            else:
                nxt[...] = 0 if dilation else 1
            if mk is not None:
                nxt = np.where(mk[i], nxt, orig).astype(np.uint8)
            cur = np.ascontiguousarray(nxt)
            k += 1
            if iterations >= 1:
                if k >= iterations:
                    break
            else:
                if prev is not None and np.array_equal(cur, prev):
                    break
                prev = cur.copy()
        out[i] = cur
    return out.reshape(lead_shape + spatial).astype(bool)


# ---------------------------------------------------------------- Sinkhorn --
def _lse(z: np.ndarray, axis: int) -> np.ndarray:
    m = np.max(z, axis=axis, keepdims=True)
    m = np.where(np.isfinite(m), m, 0.0)
    return (np.log(np.sum(np.exp(z - m), axis=axis, keepdims=True)) + m).squeeze(axis)


def sinkhorn_log_hist(a, b, cost, eps, n_iter):
    """Log-domain Sinkhorn on (n, d) histograms sharing one (d, d) cost,
    restating reference optimal_transport.py:96-125 (threshold 0) in FP64.
    Returns (log_u, log_v)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    log_k = -np.asarray(cost, np.float64) / eps
    la, lb = np.log(a), np.log(b)
    lu = np.zeros_like(a)
    lv = np.zeros_like(b)
    for _ in range(n_iter):
        lu = la - _lse(log_k[None, :, :] + lv[:, None, :], axis=-1)
        lv = lb - _lse(log_k.T[None, :, :] + lu[:, None, :], axis=-1)
    return lu, lv


def sinkhorn_scaling_hist(a, b, cost, eps, n_iter):
    """Scaling-form Sinkhorn (reference optimal_transport.py:71-93) in FP64."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    k = np.exp(-np.asarray(cost, np.float64) / eps)
    u = np.ones_like(a)
    v = np.ones_like(b)
    for _ in range(n_iter):
        u = a / (v @ k.T + 1e-12)
        v = b / (u @ k + 1e-12)
    return np.log(u), np.log(v)


def sqeuclid(x, y):
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    return ((x[:, None, :] - y[None, :, :]) ** 2).sum(-1)


def sinkhorn_points(x, y, a, b, eps, tol=1e-5, max_iter=100000, f0=None, g0=None,
                    check_every=1):
    """FP64 log-domain Sinkhorn between weighted point clouds with squared
    Euclidean cost, in dual-potential form: f = -eps*LSE_j((g_j - C_ij)/eps
    + log b_j), g likewise; stops once the row-marginal L1 error, measured
    after the g-update, is <= tol.  Returns (f, g, iterations, err)."""
    c = sqeuclid(x, y)
    la = np.log(np.asarray(a, np.float64))
    lb = np.log(np.asarray(b, np.float64))
    f = np.zeros(c.shape[0]) if f0 is None else np.asarray(f0, np.float64).copy()
    g = np.zeros(c.shape[1]) if g0 is None else np.asarray(g0, np.float64).copy()
    if f0 is not None and g0 is None:
        g = -eps * _lse((f[:, None] - c) / eps + la[:, None], axis=0)
    err = np.inf
    it = 0
    while it < max_iter:
        f = -eps * _lse((g[None, :] - c) / eps + lb[None, :], axis=1)
        g = -eps * _lse((f[:, None] - c) / eps + la[:, None], axis=0)
        it += 1
        if it % check_every == 0 or it == max_iter:
            logp = (f[:, None] + g[None, :] - c) / eps + la[:, None] + lb[None, :]
            err = np.abs(np.exp(logp).sum(1) - np.exp(la)).sum()
            if err <= tol:
                break
    return f, g, it, err


def plan_points(x, y, a, b, f, g, eps):
    c = sqeuclid(x, y)
    return np.exp((f[:, None] + g[None, :] - c) / eps) * np.asarray(a)[:, None] * np.asarray(b)[None, :]


# ------------------------------------------------ grey morphology (numpy) --
def _map_coord_np(c, size, mode):
    """Border modes of reference grey_kernel.cu:63-97, vectorised; returns
    (mapped, use_cval)."""
    c = np.asarray(c, np.int64)
    inb = (c >= 0) & (c < size)
    if mode == "constant":
        return np.where(inb, c, 0), ~inb
    if mode == "nearest":
        return np.clip(c, 0, size - 1), np.zeros_like(inb)
    if mode == "reflect":
        period = 2 * size
        m = np.mod(c, period)
        return np.where(m >= size, period - 1 - m, m), np.zeros_like(inb)
    if mode == "mirror":
        period = 2 * (size - 1)
        if period <= 0:
            return np.zeros_like(c), np.zeros_like(inb)
        m = np.mod(c, period)
        return np.where(m >= size, period - m, m), np.zeros_like(inb)
    if mode == "wrap":
        return np.mod(c, size), np.zeros_like(inb)
    raise ValueError(mode)


def _grey_table(structure, footprint, se_shape, origin, dilation):
    offs, vals = [], []
    for e in np.ndindex(*se_shape):
        if footprint is not None and not footprint[e]:
            continue
        d = [e[j] - (se_shape[j] // 2 + origin[j]) for j in range(len(se_shape))]
        s = 0.0 if structure is None else float(structure[e])
        offs.append([-v for v in d] if dilation else d)
        vals.append(s if dilation else -s)
    return np.array(offs, np.int64), np.array(vals, np.float64)


def _grey_candidates(x, offs, vals, mode, cval):
    """Per element t and output p: (value, flat source index or -1)."""
    shape = x.shape
    coords = np.indices(shape).reshape(len(shape), -1)
    n = coords.shape[1]
    V = np.empty((len(offs), n), np.float64)
    SRC = np.empty((len(offs), n), np.int64)
    xf = x.reshape(-1).astype(np.float64)
    for t, (o, v) in enumerate(zip(offs, vals)):
        flat = np.zeros(n, np.int64)
        usec = np.zeros(n, bool)
        for j in range(len(shape)):
            m, uc = _map_coord_np(coords[j] + o[j], shape[j], mode)
            flat = flat * shape[j] + m
            usec |= uc
        V[t] = np.where(usec, cval, xf[flat]) + v
        SRC[t] = np.where(usec, -1, flat)
    return V, SRC


def grey_fp64(x, structure, footprint, se_shape, origin, mode, cval, dilation):
    """Grey erosion/dilation of one item in float64 (same element semantics
    as tmo_grey), for finite-difference checks."""
    offs, vals = _grey_table(structure, footprint, se_shape, origin, dilation)
    V, _ = _grey_candidates(x, offs, vals, mode, cval)
    return (V.max(0) if dilation else V.min(0)).reshape(x.shape)


def grey_backward(x, structure, footprint, se_shape, origin, mode, cval, dilation, dy):
    """Gradient of one item: dy[p] goes to the source of the first element
    attaining the extremum (np.argmax/argmin return the first occurrence),
    nothing for cval reads.  Same rule as the HIP backward kernels."""
    offs, vals = _grey_table(structure, footprint, se_shape, origin, dilation)
    V, SRC = _grey_candidates(x.astype(np.float32), offs, vals.astype(np.float32), mode,
                              np.float32(cval))
    V = V.astype(np.float32)
    arg = V.argmax(0) if dilation else V.argmin(0)
    src = SRC[arg, np.arange(V.shape[1])]
    w = np.asarray(dy, np.float64).reshape(-1)
    keep = src >= 0
    dx = np.bincount(src[keep], weights=w[keep], minlength=x.size)
    return dx.reshape(x.shape)
