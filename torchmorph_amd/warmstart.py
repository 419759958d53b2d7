"""Learned warm start for point-cloud Sinkhorn.

A DeepSets network (~182k parameters, BF16 MFMA forward/backward with fp32
accumulation, all in libtorchmorph_hip: tm_warmstart.hip + tm_mlp.hip)
predicts the dual potential f of every source point from both clouds and eps.
:func:`solve` feeds the prediction to :func:`sinkhorn_points` as its starting
point; plain Sinkhorn still iterates to the same marginal tolerance, so the
prediction only changes how many iterations are needed, never the accuracy
of the returned solution.  A prediction that is not finite is discarded and
the pair starts cold (the fallback).

No reference counterpart (the reference has no learned component).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .optimal_transport import SinkhornResult, _as_ragged, sinkhorn_points

K_IN, H, G = 4, 128, 256
LAYOUT = [("W0", (K_IN, H)), ("b0", (H,)), ("W1", (H, H)), ("b1", (H,)),
          ("W2a", (H, G)), ("W2x", (H, G)), ("W2y", (H, G)), ("w2l", (G,)), ("b2", (G,)),
          ("W3", (G, G)), ("b3", (G,)), ("w4", (G,)), ("b4", (1,))]


def n_params() -> int:
    return int(_lib.lib().tm_warmstart_param_count())


def init_params(seed: int = 0, device="cuda") -> torch.Tensor:
    """He-normal weights, zero biases, as one flat float32 vector."""
    g = torch.Generator().manual_seed(seed)
    parts = []
    for name, shape in LAYOUT:
        if len(shape) == 2:
            parts.append(torch.randn(shape, generator=g, dtype=torch.float64) *
                         math.sqrt(2.0 / shape[0]))
        elif name in ("w2l", "w4"):
            parts.append(torch.randn(shape, generator=g, dtype=torch.float64) *
                         math.sqrt(1.0 / shape[0]))
        else:
            parts.append(torch.zeros(shape, dtype=torch.float64))
    flat = torch.cat([p.reshape(-1) for p in parts]).float()
    assert flat.numel() == n_params()
    return flat.to(device)


def _sizes(xs, ys):
    x, n = _as_ragged(xs)
    y, m = _as_ragged(ys)
    if x.shape[1] != 3 or y.shape[1] != 3:
        raise ValueError("the warm-start network takes 3-D point clouds")
    return x.detach().float().contiguous(), y.detach().float().contiguous(), n, m


def _eps_array(eps, P):
    if isinstance(eps, torch.Tensor):
        vals = eps.detach().float().reshape(-1).cpu().tolist()
    elif isinstance(eps, (list, tuple, np.ndarray)):
        vals = [float(v) for v in eps]
    else:
        vals = [float(eps)] * P
    if len(vals) != P:
        raise ValueError("eps must be a scalar or one value per pair")
    return (_lib.ctypes.c_float * P)(*vals), torch.tensor(vals)


@dataclass
class WarmStartNet:
    params: torch.Tensor
    m1: torch.Tensor = field(default=None)
    m2: torch.Tensor = field(default=None)
    step: int = 0

    @classmethod
    def create(cls, seed: int = 0, device="cuda") -> "WarmStartNet":
        p = init_params(seed, device)
        return cls(p, torch.zeros_like(p), torch.zeros_like(p), 0)

    # ----------------------------------------------------------- inference
    @torch.no_grad()
    def predict(self, xs, ys, eps) -> torch.Tensor:
        """Predicted f (concatenated over pairs, cost units)."""
        x, y, n, m = _sizes(xs, ys)
        P = len(n)
        L = _lib.lib()
        na = (_lib.ctypes.c_int32 * P)(*n)
        ma = (_lib.ctypes.c_int32 * P)(*m)
        ea, _ = _eps_array(eps, P)
        f = torch.empty(sum(n), device=x.device)
        with torch.cuda.device(x.device):
            nb = L.tm_warmstart_workspace_bytes(P, na, ma)
            ws = _lib.workspace(nb, x.device)
            _lib.check(L.tm_warmstart_predict(_lib.ptr(self.params), _lib.ptr(x), _lib.ptr(y), P,
                                              na, ma, ea, _lib.ptr(f), _lib.ptr(ws), nb,
                                              _lib.stream_of(x.device)), "tm_warmstart_predict")
        return f

    # ------------------------------------------------------------ training
    def loss_and_grad(self, xs, ys, eps, f_target, loss_alpha: float = 1.0):
        x, y, n, m = _sizes(xs, ys)
        P = len(n)
        L = _lib.lib()
        na = (_lib.ctypes.c_int32 * P)(*n)
        ma = (_lib.ctypes.c_int32 * P)(*m)
        ea, _ = _eps_array(eps, P)
        t = f_target.detach().float().reshape(-1).contiguous()
        grad = torch.empty_like(self.params)
        loss = torch.empty(1, device=x.device)
        with torch.cuda.device(x.device):
            nb = L.tm_warmstart_workspace_bytes(P, na, ma)
            ws = _lib.workspace(nb, x.device)
            _lib.check(L.tm_warmstart_grad(_lib.ptr(self.params), _lib.ptr(x), _lib.ptr(y), P,
                                           na, ma, ea, _lib.ptr(t), float(loss_alpha),
                                           _lib.ptr(grad),
                                           _lib.ptr(loss), _lib.ptr(ws), nb,
                                           _lib.stream_of(x.device)), "tm_warmstart_grad")
        return loss, grad

    def train_step(self, xs, ys, eps, f_target, lr=1e-3, betas=(0.9, 0.999), adam_eps=1e-8,
                   weight_decay=0.0, loss_alpha: float = 1.0) -> torch.Tensor:
        loss, grad = self.loss_and_grad(xs, ys, eps, f_target, loss_alpha)
        self.step += 1
        _lib.check(_lib.lib().tm_warmstart_adam(
            _lib.ptr(self.params), _lib.ptr(grad), _lib.ptr(self.m1), _lib.ptr(self.m2),
            self.params.numel(), self.step, float(lr), float(betas[0]), float(betas[1]),
            float(adam_eps), float(weight_decay), _lib.stream_of(self.params.device)),
            "tm_warmstart_adam")
        return loss

    # ------------------------------------------------------ checkpointing
    def state_dict(self):
        return {"params": self.params.detach().cpu(), "m1": self.m1.detach().cpu(),
                "m2": self.m2.detach().cpu(), "step": self.step}

    @classmethod
    def from_state_dict(cls, sd, device="cuda") -> "WarmStartNet":
        return cls(sd["params"].to(device), sd["m1"].to(device), sd["m2"].to(device),
                   int(sd["step"]))

    def save(self, path):
        torch.save(self.state_dict(), path)

    @classmethod
    def load(cls, path, device="cuda") -> "WarmStartNet":
        return cls.from_state_dict(torch.load(path, map_location="cpu", weights_only=True),
                                   device)


def solve(net: WarmStartNet | None, xs, ys, eps, tol=1e-5, max_iter=20000,
          chunk=10) -> SinkhornResult:
    """Warm-started Sinkhorn: the network's f is the starting point, plain
    Sinkhorn runs to `tol`.  Non-finite predictions fall back to a cold start
    for their pair; `net=None` is the cold start."""
    f0 = None
    if net is not None:
        f0 = net.predict(xs, ys, eps)
        if not torch.isfinite(f0).all():
            _, n = _as_ragged(xs)
            parts = list(torch.split(f0, n))
            f0 = torch.cat([p if torch.isfinite(p).all() else torch.zeros_like(p)
                            for p in parts])
    return sinkhorn_points(xs, ys, eps, tol=tol, max_iter=max_iter, f_init=f0, chunk=chunk)


def gmm_clouds(rng: np.random.Generator, sizes, d=3):
    """Point clouds from 4-component Gaussian mixtures (means N(0, 4I),
    isotropic sigma U(0.3, 1), Dirichlet(1) weights), one per size."""
    out = []
    for k in sizes:
        means = rng.normal(0.0, 2.0, size=(4, d))
        sig = rng.uniform(0.3, 1.0, size=4)
        w = rng.dirichlet(np.ones(4))
        comp = rng.choice(4, size=k, p=w)
        out.append((means[comp] + rng.normal(size=(k, d)) * sig[comp, None]).astype(np.float32))
    return out
