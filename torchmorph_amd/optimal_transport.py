"""Entropic optimal transport (log-domain Sinkhorn) on MI355X.

Two entry points:

* :class:`SinkhornSolver` / :func:`build_cost_matrix` -- the reference's
  histogram API (reference torchmorph/optimal_transport.py:7-18,128-249):
  (n, d) histograms sharing one (d, d) cost, same arguments, preprocessing,
  envelope-theorem backward and messages.  GPU tensors run
  ``tm_sinkhorn_hist_solve``: one update kernel per half-iteration, the value
  <P, C> and the centred potentials fused into the last update and a finish
  kernel (no n x d x d tensor), chunks of iterations replayed as a hipGraph
  for max_iter >= 100.  As in the reference, float32 runs exactly max_iter
  iterations (its fused kernels ignore ``threshold``) and every other dtype
  honours ``threshold``; float64 is solved in float64, float16 / bfloat16 in
  float32 with results cast back.  ``plan()`` is formed only when asked for
  (``tm_sinkhorn_hist_plan``).  CPU tensors run the same iterations with
  torch operations, batch rows in chunks.
* :func:`sinkhorn_points` / :class:`SinkhornPointCloud` -- no reference
  counterpart: ragged batches of weighted point-cloud pairs with the squared
  Euclidean cost computed on the fly, solved to a row-marginal L1 error
  tolerance with device-side convergence (``tm_sinkhorn_points``), optional
  warm-start potentials, and a differentiable OT value (gradients from the
  converged plan, see :class:`SinkhornPointCloud`).
* :func:`sinkhorn_potentials` / :class:`SinkhornPotentials` -- no reference
  counterpart: the gauge-fixed potentials themselves, differentiable through
  the implicit adjoint of the fixed point (conjugate gradients) or reverse
  mode from the converged checkpoint, with every plan product formed on the
  fly (``tm_sinkhorn_points_planmat``).
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence

import torch
from torch import Tensor, nn

from . import _lib

# ------------------------------------------------------------------ histograms


def build_cost_matrix(shape, p=2, device=None) -> Tensor:
    """(d, d) L^p distances between the d points of a grid of ``shape``, points
    in row-major order (reference optimal_transport.py:7-18)."""
    axes = [torch.arange(int(s), device=device).float() for s in shape]
    points = torch.cartesian_prod(*axes) if len(axes) > 1 else axes[0][:, None]
    return torch.cdist(points, points, p=p)


# batch rows per CPU step: bounds the (rows, d, d) temporaries of the CPU path
_CPU_TERMS = 1 << 24


@dataclass
class _HistSolution:
    """State of a histogram solve: the final (log-)scalings in the compute
    dtype, and whichever of the centred potentials and the value were asked
    for (in the input dtype)."""

    log_space: bool
    eps: float
    u: Tensor
    v: Tensor
    cost: Tensor
    iterations: int
    f: Optional[Tensor] = None
    g: Optional[Tensor] = None
    value: Optional[Tensor] = None

    def log_scalings(self):
        return (self.u, self.v) if self.log_space else (self.u.log(), self.v.log())


def _row_chunks(n: int, d: int):
    step = max(1, _CPU_TERMS // max(1, d * d))
    return [(r, min(n, r + step)) for r in range(0, n, step)]


def _cpu_solve(a, b, cost, eps, max_iter, threshold, log_space, want) -> _HistSolution:
    """CPU tensors: the reference's torch iterations (optimal_transport.py:71-125).

    Log form ``log u = log a - LSE_j(log v_j - C_ij / eps)`` (then v from u),
    scaling form ``u = a / (K v + 1e-12)`` with K = exp(-C / eps); with
    ``threshold`` > 0 the run stops after iteration i (i % 10 == 0) once the
    largest row L1 change of u is <= threshold.  Log-sum-exps are taken over
    batch chunks so no (n, d, d) temporary exists at once."""
    n, d = a.shape
    chunks = _row_chunks(n, d)
    if log_space:
        log_k = -cost / eps
        log_kt = log_k.mT
        la, lb = a.log(), b.log()

        def half(num, lk, scale):
            return torch.cat([num[r0:r1] - torch.logsumexp(lk + scale[r0:r1, None, :], dim=-1)
                              for r0, r1 in chunks]) if n else num.clone()

        u_step = lambda v: half(la, log_k, v)  # noqa: E731
        v_step = lambda u: half(lb, log_kt, u)  # noqa: E731
        u, v = torch.zeros_like(a), torch.zeros_like(b)
    else:
        k = torch.exp(-cost / eps)
        u_step = lambda v: a / (v @ k.mT + 1e-12)  # noqa: E731
        v_step = lambda u: b / (u @ k + 1e-12)  # noqa: E731
        u, v = torch.ones_like(a), torch.ones_like(b)
    done = max_iter
    for it in range(max_iter):
        before = u
        u = u_step(v)
        v = v_step(u)
        if threshold > 0 and it % 10 == 0 and n:
            if bool((u - before).abs().sum(dim=-1).max() <= threshold):
                done = it + 1
                break
    sol = _HistSolution(log_space, eps, u, v, cost, done)
    lu, lv = sol.log_scalings()
    if "potentials" in want:
        sol.f = eps * lu - (eps * lu).mean(dim=-1, keepdim=True)
        sol.g = eps * lv - (eps * lv).mean(dim=-1, keepdim=True)
    if "value" in want:
        parts = []
        for r0, r1 in chunks:
            plan = torch.exp(lu[r0:r1, :, None] - cost / eps + lv[r0:r1, None, :])
            parts.append((plan * cost).sum(dim=(-2, -1)))
        sol.value = torch.cat(parts) if parts else a.new_zeros((0,))
    return sol


def _gpu_solve(a, b, cost, eps, max_iter, threshold, log_space, want) -> _HistSolution:
    """GPU tensors: ``tm_sinkhorn_hist_solve`` (HIP), computing in float32 or
    float64; float32 ignores ``threshold`` as the reference's fused kernels do."""
    out_dtype = a.dtype
    cdt = a.dtype if a.dtype in (torch.float32, torch.float64) else torch.float32
    if a.dtype == torch.float32:
        threshold = 0.0
    a = a.detach().to(cdt).contiguous()
    b = b.detach().to(cdt).contiguous()
    cost = cost.detach().to(cdt).contiguous()
    n, d = a.shape
    dev = a.device
    if n == 0:
        empty = a.new_zeros((0, d))
        return _HistSolution(log_space, eps, empty, empty, cost, 0, empty.to(out_dtype),
                             empty.to(out_dtype), a.new_zeros((0,), dtype=out_dtype))
    cost_t = cost.mT.contiguous()
    if log_space:
        num_a, num_b, M, M_t = a.log(), b.log(), cost, cost_t
        u, v = torch.zeros_like(a), torch.zeros_like(b)
    else:
        num_a, num_b = a, b
        M = torch.exp(-cost / eps).contiguous()
        M_t = M.mT.contiguous()
        u, v = torch.ones_like(a), torch.ones_like(b)
    f = torch.empty_like(a) if "potentials" in want else None
    g = torch.empty_like(b) if "potentials" in want else None
    value = a.new_empty((n,)) if "value" in want else None
    L = _lib.lib()
    dtype = _lib.TM_F32 if cdt == torch.float32 else _lib.TM_F64
    iters = _lib.ctypes.c_int(0)
    fb = _lib.ctypes.c_int(0)
    with torch.cuda.device(dev):
        nbytes = L.tm_sinkhorn_hist_workspace_bytes(dtype, n, d, float(threshold),
                                                    int(value is not None))
        ws = _lib.workspace(nbytes, dev)
        _lib.check(L.tm_sinkhorn_hist_solve(
            dtype, int(log_space), _lib.ptr(num_a), _lib.ptr(num_b), _lib.ptr(M), _lib.ptr(M_t),
            _lib.ptr(cost), _lib.ptr(cost_t), _lib.ptr(u), _lib.ptr(v), n, d, int(max_iter),
            float(eps), float(threshold), _lib.ptr(f), _lib.ptr(g), _lib.ptr(value),
            _lib.ctypes.byref(iters), _lib.ctypes.byref(fb), _lib.ptr(ws), nbytes,
            _lib.stream_of(dev)), "tm_sinkhorn_hist_solve")
    cast = (lambda t: None if t is None else t.to(out_dtype))  # noqa: E731
    return _HistSolution(log_space, eps, u, v, cost, iters.value, cast(f), cast(g), cast(value))


def _hist_solve(a, b, cost, eps, max_iter, threshold, log_space, want=()) -> _HistSolution:
    solver = _gpu_solve if a.is_cuda else _cpu_solve
    return solver(a, b, cost, float(eps), int(max_iter), float(threshold), bool(log_space),
                  tuple(want))


def _hist_plan(sol: _HistSolution, out_dtype) -> Tensor:
    """(n, d, d) plan exp(log u_i - C_ij / eps + log v_j) of a solution
    (reference optimal_transport.py:228-233)."""
    n, d = sol.u.shape
    if not sol.u.is_cuda:
        lu, lv = sol.log_scalings()
        return torch.exp(lu[:, :, None] - sol.cost / sol.eps + lv[:, None, :]).to(out_dtype)
    plan = sol.u.new_empty((n, d, d))
    if n:
        dtype = _lib.TM_F32 if sol.u.dtype == torch.float32 else _lib.TM_F64
        with torch.cuda.device(sol.u.device):
            _lib.check(_lib.lib().tm_sinkhorn_hist_plan(
                dtype, int(sol.log_space), _lib.ptr(sol.u), _lib.ptr(sol.v), _lib.ptr(sol.cost),
                n, d, float(sol.eps), _lib.ptr(plan), _lib.stream_of(sol.u.device)),
                "tm_sinkhorn_hist_plan")
    return plan.to(out_dtype)


class _SinkhornDistance(torch.autograd.Function):
    """<P, C> per batch row; its gradients w.r.t. the two marginals are the
    centred dual potentials (envelope theorem, reference
    optimal_transport.py:128-150), the cost gets none."""

    @staticmethod
    def forward(ctx, a, b, cost_matrix, epsilon, max_iter, threshold, log_space):
        sol = _hist_solve(a, b, cost_matrix, epsilon, max_iter, threshold, log_space,
                          want=("value", "potentials"))
        ctx.save_for_backward(sol.f, sol.g)
        return sol.value

    @staticmethod
    def backward(ctx, grad_output):
        f, g = ctx.saved_tensors
        scale = grad_output[:, None]
        return scale * f, scale * g, None, None, None, None, None


def _unit_mass(h: Tensor) -> Tensor:
    h = h.clamp(min=0)
    return h / h.sum(dim=-1, keepdim=True).clamp(min=1e-12)


class SinkhornSolver(nn.Module):
    """Entropy-regularised balanced OT between (n, d) histograms sharing one
    (d, d) cost (reference optimal_transport.py:153-249): forward returns the
    (n,) values <P, C>, differentiable w.r.t. both histograms."""

    def __init__(self, epsilon=1.0, max_iter=100, threshold=0.0, p=2, log_space=False):
        super().__init__()
        if epsilon <= 0:
            raise ValueError("epsilon must be positive.")
        if max_iter <= 0:
            raise ValueError("max_iter must be positive.")
        if threshold < 0:
            raise ValueError("threshold must be non-negative.")
        self.epsilon = epsilon
        self.max_iter = max_iter
        self.threshold = threshold
        self.p = p
        self.log_space = log_space

    def extra_repr(self):
        return (f"epsilon={self.epsilon}, max_iter={self.max_iter}, "
                f"threshold={self.threshold}, log_space={self.log_space}")

    def data_preprocess(self, source, target, cost_matrix: Optional[Tensor] = None):
        """Checks the (n, d) pair, resolves the cost (default: the d bins as
        points on a line) and scales each row to unit mass after clamping
        negative entries (differentiable)."""
        if source.ndim != 2 or source.shape != target.shape:
            raise ValueError("source and target must be (n, d) tensors of the same shape.")
        if not source.is_floating_point():
            raise ValueError("source and target must be floating-point tensors.")
        d = source.shape[1]
        if cost_matrix is not None and cost_matrix.shape != (d, d):
            raise ValueError(
                f"cost_matrix must have shape ({d}, {d}), got {tuple(cost_matrix.shape)}.")
        cost = build_cost_matrix((d,), self.p, source.device) if cost_matrix is None \
            else cost_matrix
        cost = cost.to(device=source.device, dtype=source.dtype).contiguous()
        return _unit_mass(source), _unit_mass(target), cost

    def _solve(self, source, target, cost_matrix, want):
        a, b, cost = self.data_preprocess(source, target, cost_matrix)
        return _hist_solve(a, b, cost, self.epsilon, self.max_iter, self.threshold,
                           self.log_space, want)

    def forward(self, source: Tensor, target: Tensor, cost_matrix: Optional[Tensor] = None):
        """(n,) transport values between the rows of source and target."""
        a, b, cost = self.data_preprocess(source, target, cost_matrix)
        return _SinkhornDistance.apply(a, b, cost, self.epsilon, self.max_iter, self.threshold,
                                       self.log_space)

    @torch.no_grad()
    def plan(self, source: Tensor, target: Tensor, cost_matrix: Optional[Tensor] = None):
        """(n, d, d) transport plans, formed from the final log-scalings."""
        return _hist_plan(self._solve(source, target, cost_matrix, ()), source.dtype)

    @torch.no_grad()
    def potentials(self, source: Tensor, target: Tensor, cost_matrix: Optional[Tensor] = None):
        """Centred dual potentials (f, g), each (n, d): the gradients of the
        value w.r.t. the marginals."""
        sol = self._solve(source, target, cost_matrix, ("potentials",))
        return sol.f, sol.g


# ---------------------------------------------------------------- point clouds


@dataclass
class SinkhornResult:
    """Solution of a batch of point-cloud pairs (concatenated layout).

    f, g: potentials of every point (sum n_p,), (sum m_p,); iterations and
    err: per pair.  split_f()/split_g() return per-pair views."""

    f: Tensor
    g: Tensor
    iterations: Tensor
    err: Tensor
    n: list
    m: list
    eps: Tensor

    def split_f(self):
        return list(torch.split(self.f, self.n))

    def split_g(self):
        return list(torch.split(self.g, self.m))


def _as_ragged(clouds) -> tuple[Tensor, list]:
    if isinstance(clouds, Tensor):
        if clouds.ndim != 3:
            raise ValueError("point clouds must be (B, n, d) or a list of (n_p, d) tensors")
        return clouds.reshape(-1, clouds.shape[-1]), [clouds.shape[1]] * clouds.shape[0]
    sizes = [int(c.shape[0]) for c in clouds]
    return torch.cat(list(clouds), dim=0), sizes


def _weights(w, sizes, device) -> Tensor:
    if w is None:
        # uniform 1 / n_p per pair in one fill (equal sizes) or one
        # repeat_interleave, not a fill per pair; 1 / n_p rounds to float32
        # as a fill of it does
        total = sum(sizes)
        if len(set(sizes)) == 1:
            return torch.full((total,), 1.0 / sizes[0], device=device)
        vals = torch.tensor([1.0 / s for s in sizes], dtype=torch.float32).to(device)
        reps = torch.tensor(sizes, dtype=torch.int64).to(device)
        return torch.repeat_interleave(vals, reps, output_size=total)
    w = torch.cat(list(w)) if not isinstance(w, Tensor) else w.reshape(-1)
    return w.to(device=device, dtype=torch.float32)


def sinkhorn_points(xs, ys, eps, a=None, b=None, tol: float = 1e-5, max_iter: int = 20000,
                    f_init: Optional[Tensor] = None, chunk: int = 10) -> SinkhornResult:
    """Solves entropic OT between every pair (xs[p], ys[p]) with cost
    |x - y|^2, to an L1 row-marginal error <= tol measured after a g-step.

    xs, ys: (B, n, d) / (B, m, d) tensors or lists of (n_p, d) / (m_p, d);
    eps: float or (B,) tensor; a, b: weights (default uniform), same layouts;
    f_init: concatenated initial f (warm start), default zeros.
    Runs entirely on the GPU (tm_sinkhorn_points): one launch per half-step
    for the whole ragged batch, convergence checked on the device, the
    iteration loop replayed as a hipGraph of `chunk` iterations.
    """
    x, n = _as_ragged(xs)
    y, m = _as_ragged(ys)
    if len(n) != len(m):
        raise ValueError("xs and ys must hold the same number of clouds")
    if not x.is_cuda:
        raise ValueError("sinkhorn_points runs on GPU tensors")
    dev = x.device
    x = x.detach().to(torch.float32)
    y = y.detach().to(torch.float32)
    if x.shape[1] != y.shape[1] or not 1 <= x.shape[1] <= 4:
        raise ValueError("point dimension must match and be in 1..4")
    x, y = x.contiguous(), y.contiguous()
    P = len(n)
    aw = _weights(a, n, dev).contiguous()
    bw = _weights(b, m, dev).contiguous()
    if isinstance(eps, Tensor):
        eps_t = eps.detach().to(torch.float32).reshape(-1).cpu()
    elif isinstance(eps, (list, tuple)) or hasattr(eps, "__len__"):
        eps_t = torch.tensor([float(v) for v in eps], dtype=torch.float32)
    else:
        eps_t = torch.full((P,), float(eps))
    if eps_t.numel() != P:
        raise ValueError("eps must be a scalar or hold one value per pair")
    f = (torch.zeros(sum(n), device=dev) if f_init is None
         else f_init.detach().to(device=dev, dtype=torch.float32).reshape(-1).clone())
    g = torch.zeros(sum(m), device=dev)
    iters = torch.zeros(P, dtype=torch.int32, device=dev)
    err = torch.zeros(P, dtype=torch.float32, device=dev)
    L = _lib.lib()
    n_arr = (_lib.ctypes.c_int32 * P)(*n)
    m_arr = (_lib.ctypes.c_int32 * P)(*m)
    e_arr = (_lib.ctypes.c_float * P)(*[float(v) for v in eps_t.tolist()])
    with torch.cuda.device(dev):
        nbytes = L.tm_sinkhorn_points_workspace_bytes(P, n_arr, m_arr)
        ws = _lib.workspace(nbytes, dev)
        _lib.check(L.tm_sinkhorn_points(_lib.ptr(x), _lib.ptr(y), _lib.ptr(aw), _lib.ptr(bw),
                                        x.shape[1], P, n_arr, m_arr, e_arr, float(tol),
                                        int(max_iter), int(chunk), _lib.ptr(f), _lib.ptr(g),
                                        _lib.ptr(iters), _lib.ptr(err), _lib.ptr(ws), nbytes,
                                        _lib.stream_of(dev)), "tm_sinkhorn_points")
    return SinkhornResult(f, g, iters, err, n, m, eps_t)


def median_sqeuclid(x: Tensor, y: Tensor) -> Tensor:
    """Median of |x_i - y_j|^2 over all pairs of one (n, d) / (m, d) pair."""
    return torch.cdist(x.float(), y.float()).square().flatten().median()


def sinkhorn_points_grad(xs, ys, res: SinkhornResult, a=None, b=None):
    """Gradients of every pair's OT value <f,a> + <g,b> w.r.t. its points,
    from a converged :class:`SinkhornResult` (HIP kernel, plan formed on the
    fly).  Returns concatenated (gx, gy)."""
    x, n = _as_ragged(xs)
    y, m = _as_ragged(ys)
    dev = x.device
    x = x.detach().to(torch.float32).contiguous()
    y = y.detach().to(torch.float32).contiguous()
    aw = _weights(a, n, dev).contiguous()
    bw = _weights(b, m, dev).contiguous()
    P = len(n)
    gx = torch.empty_like(x)
    gy = torch.empty_like(y)
    L = _lib.lib()
    n_arr = (_lib.ctypes.c_int32 * P)(*n)
    m_arr = (_lib.ctypes.c_int32 * P)(*m)
    e_arr = (_lib.ctypes.c_float * P)(*[float(v) for v in res.eps.tolist()])
    with torch.cuda.device(dev):
        nbytes = L.tm_sinkhorn_points_grad_workspace_bytes(P, n_arr, m_arr)
        ws = _lib.workspace(nbytes, dev)
        _lib.check(L.tm_sinkhorn_points_grad(
            _lib.ptr(x), _lib.ptr(y), _lib.ptr(aw), _lib.ptr(bw), x.shape[1], P, n_arr, m_arr,
            e_arr, _lib.ptr(res.f.contiguous()), _lib.ptr(res.g.contiguous()), _lib.ptr(gx),
            _lib.ptr(gy), _lib.ptr(ws), nbytes, _lib.stream_of(dev)), "tm_sinkhorn_points_grad")
    return gx, gy


class SinkhornPointCloud(torch.autograd.Function):
    """Entropic OT value <f, a> + <g, b> between two point clouds (one pair),
    differentiable w.r.t. the points and the weights.

    Backward (HIP, tm_sinkhorn_points_grad): at the converged fixed point the
    value's derivatives are d/da = f, d/db = g (centred: the mass-preserving
    direction), d/dx_i = 2 sum_j P_ij (x_i - y_j), d/dy_j = 2 sum_i P_ij
    (y_j - x_i) -- the envelope theorem applied to the dual problem; tests
    check them against FP64 central differences of a tightly converged
    solve."""

    @staticmethod
    def forward(ctx, x, y, a, b, eps, tol, max_iter):
        res = sinkhorn_points(x[None], y[None], eps, a[None], b[None], tol=tol, max_iter=max_iter)
        value = (res.f * a.float()).sum() + (res.g * b.float()).sum()
        ctx.save_for_backward(x, y, a, b)
        ctx.res = res
        return value.to(x.dtype)

    @staticmethod
    def backward(ctx, grad):
        x, y, a, b = ctx.saved_tensors
        res = ctx.res
        gx, gy = sinkhorn_points_grad(x[None], y[None], res, a[None], b[None])
        f, g = res.f, res.g
        ga = f - f.mean()
        gb = g - g.mean()
        return (grad * gx).to(x.dtype), (grad * gy).to(y.dtype), (grad * ga).to(a.dtype), \
            (grad * gb).to(b.dtype), None, None, None


# ------------------------------------------------ implicit adjoint (one pair)


class PlanOps:
    """Products with the entropic plan P of one solved pair, P_ij = a_i b_j
    exp((f_i + g_j - |x_i - y_j|^2) / eps): ``rows(V) = P V`` ((m, k) ->
    (n, k)) and ``cols(U) = P^T U`` ((n, k) -> (m, k)).  The adjoint below
    needs nothing else, so the plan is never materialised on the GPU."""

    def rows(self, V: Tensor) -> Tensor:
        raise NotImplementedError

    def cols(self, U: Tensor) -> Tensor:
        raise NotImplementedError


class HipPlanOps(PlanOps):
    """PlanOps of GPU float32 tensors (``tm_sinkhorn_points_planmat``: the
    plan formed on the fly per tile, 8 vectors per launch)."""

    def __init__(self, x, y, a, b, f, g, eps):
        self.x, self.y = x.detach().float().contiguous(), y.detach().float().contiguous()
        self.a, self.b = a.detach().float().contiguous(), b.detach().float().contiguous()
        self.f, self.g = f.detach().float().contiguous(), g.detach().float().contiguous()
        self.n_arr = (_lib.ctypes.c_int32 * 1)(self.x.shape[0])
        self.m_arr = (_lib.ctypes.c_int32 * 1)(self.y.shape[0])
        self.e_arr = (_lib.ctypes.c_float * 1)(float(eps))
        L = _lib.lib()
        self.nbytes = L.tm_sinkhorn_points_grad_workspace_bytes(1, self.n_arr, self.m_arr)
        self.ws = _lib.workspace(self.nbytes, self.x.device)

    def _prod(self, V: Tensor, transpose: int) -> Tensor:
        V = V.detach().float()
        rows = self.y.shape[0] if transpose else self.x.shape[0]
        outs = []
        L = _lib.lib()
        with torch.cuda.device(self.x.device):
            for k0 in range(0, V.shape[1], 8):
                Vk = V[:, k0:k0 + 8].contiguous()
                out = torch.empty((rows, Vk.shape[1]), device=V.device, dtype=torch.float32)
                _lib.check(L.tm_sinkhorn_points_planmat(
                    _lib.ptr(self.x), _lib.ptr(self.y), _lib.ptr(self.a), _lib.ptr(self.b),
                    self.x.shape[1], 1, self.n_arr, self.m_arr, self.e_arr, _lib.ptr(self.f),
                    _lib.ptr(self.g), transpose, Vk.shape[1], _lib.ptr(Vk), _lib.ptr(out),
                    _lib.ptr(self.ws), self.nbytes, _lib.stream_of(self.x.device)),
                    "tm_sinkhorn_points_planmat")
                outs.append(out)
        return torch.cat(outs, dim=1)

    def rows(self, V):
        return self._prod(V, 0)

    def cols(self, U):
        return self._prod(U, 1)


def solve_adjoint(ops: PlanOps, a: Tensor, b: Tensor, vf: Tensor, vg: Tensor,
                  mode: str = "implicit", steps: int = 200, tol: float = 1e-7):
    """The adjoint (alpha, beta) of the Sinkhorn fixed point for upstream
    potential gradients (vf, vg), gauge invariant (sum vf = sum vg).

    The fixed point f = T_f(g), g = T_g(f) has Jacobian blocks
    df/dg = -P / a (rows) and dg/df = -(P / b)^T; with alpha = w_f / a and
    beta = w_g / b the adjoint system (I - J^T) w = v reads
        a * alpha + P beta = vf,    P^T alpha + b * beta = vg,
    a symmetric positive semi-definite system (null space (1, -1)).
    mode "implicit": conjugate gradients with the Jacobi preconditioner
    (a, b) to a relative residual `tol` (or `steps` iterations).
    mode "unrolled": reverse mode through `steps` Sinkhorn iterations
    recomputed from the converged checkpoint (each recomputed iterate equals
    the checkpoint, so every step uses the checkpoint's plan): the truncated
    Neumann series of the implicit solution."""
    vf = vf.reshape(-1, 1).to(a.dtype)
    vg = vg.reshape(-1, 1).to(a.dtype)
    a = a.reshape(-1, 1)
    b = b.reshape(-1, 1)
    if mode == "unrolled":
        alpha = torch.zeros_like(vf)
        beta = torch.zeros_like(vg)
        gbar = vg
        fbar = vf - ops.rows(gbar / b).to(a.dtype)
        for t in range(steps):
            beta = beta + gbar / b
            alpha = alpha + fbar / a
            gbar = -ops.cols(fbar / a).to(a.dtype)
            fbar = -ops.rows(gbar / b).to(a.dtype)
        return alpha.reshape(-1), beta.reshape(-1)
    if mode != "implicit":
        raise ValueError("mode must be 'implicit' or 'unrolled'")

    def apply_m(al, be):
        return a * al + ops.rows(be).to(a.dtype), ops.cols(al).to(a.dtype) + b * be

    al = torch.zeros_like(vf)
    be = torch.zeros_like(vg)
    rf, rg = vf.clone(), vg.clone()
    zf, zg = rf / a, rg / b
    pf, pg = zf.clone(), zg.clone()
    rz = (rf * zf).sum() + (rg * zg).sum()
    r0 = float(torch.sqrt((rf * rf).sum() + (rg * rg).sum()))
    if r0 == 0.0:
        return al.reshape(-1), be.reshape(-1)
    for _ in range(steps):
        qf, qg = apply_m(pf, pg)
        pq = (pf * qf).sum() + (pg * qg).sum()
        if float(pq) <= 0.0:
            break
        step = rz / pq
        al = al + step * pf
        be = be + step * pg
        rf = rf - step * qf
        rg = rg - step * qg
        if float(torch.sqrt((rf * rf).sum() + (rg * rg).sum())) <= tol * r0:
            break
        zf, zg = rf / a, rg / b
        rz_new = (rf * zf).sum() + (rg * zg).sum()
        pf = zf + (rz_new / rz) * pf
        pg = zg + (rz_new / rz) * pg
        rz = rz_new
    # the gauge component (1, -1) changes nothing downstream except a constant
    # in the weight gradients, which are centred; drop it anyway
    shift = (al.sum() - be.sum()) / (al.numel() + be.numel())
    return (al - shift).reshape(-1), (be + shift).reshape(-1)


def adjoint_param_grads(ops: PlanOps, x: Tensor, y: Tensor, a: Tensor, b: Tensor, eps: float,
                        alpha: Tensor, beta: Tensor):
    """Gradients through the fixed point for the adjoint (alpha, beta):
    dL/dC_ij = P_ij (alpha_i + beta_j), so dL/dx_i = 2 sum_j P_ij (alpha_i +
    beta_j)(x_i - y_j) (likewise y), dL/da_i = -eps (P beta)_i / a_i and
    dL/db_j = -eps (P^T alpha)_j / b_j.  Coordinates are taken relative to
    the y mean (the cost is translation invariant)."""
    dt = a.dtype
    ctr = y.mean(dim=0, keepdim=True)
    xc, yc = (x - ctr).to(dt), (y - ctr).to(dt)
    al, be = alpha.reshape(-1, 1).to(dt), beta.reshape(-1, 1).to(dt)
    one_m = torch.ones_like(be)
    one_n = torch.ones_like(al)
    R = ops.rows(torch.cat([one_m, yc, be, be * yc], dim=1)).to(dt)  # (n, 2D + 2)
    Cc = ops.cols(torch.cat([one_n, xc, al, al * xc], dim=1)).to(dt)  # (m, 2D + 2)
    D = x.shape[1]
    r0, ry, rb, rby = R[:, :1], R[:, 1:1 + D], R[:, 1 + D:2 + D], R[:, 2 + D:]
    c0, cx, ca, cax = Cc[:, :1], Cc[:, 1:1 + D], Cc[:, 1 + D:2 + D], Cc[:, 2 + D:]
    gx = 2 * (xc * (al * r0 + rb) - (al * ry + rby))
    gy = 2 * (yc * (be * c0 + ca) - (be * cx + cax))
    ga = -eps * rb.reshape(-1) / a.reshape(-1)
    gb = -eps * ca.reshape(-1) / b.reshape(-1)
    return gx, gy, ga, gb


class SinkhornPotentials(torch.autograd.Function):
    """Converged dual potentials (f, g) of one point-cloud pair (cost
    |x - y|^2, weights a, b summing to 1), differentiable w.r.t. x, y, a, b
    through the fixed point (no reference counterpart).

    Gauge: f, g are shifted by c = (<g, b> - <f, a>) / 2 so that <f, a> =
    <g, b>, which makes them unique.  Backward: the adjoint of the fixed
    point (:func:`solve_adjoint`, ``mode`` "implicit" or "unrolled" with
    ``steps``), then :func:`adjoint_param_grads`; weight gradients are
    centred (mass-preserving directions).  Plan-based losses differentiate
    through P = a b exp((f + g - C) / eps) built from these potentials with
    torch operations, whose direct terms autograd adds."""

    @staticmethod
    def forward(ctx, x, y, a, b, eps, tol=1e-6, max_iter=20000, mode="implicit", steps=300):
        res = sinkhorn_points(x[None], y[None], eps, a[None], b[None], tol=tol,
                              max_iter=max_iter)
        f = res.f.to(x.dtype)
        g = res.g.to(x.dtype)
        c = ((g * b).sum() - (f * a).sum()) / 2
        f, g = f + c, g - c
        ctx.save_for_backward(x, y, a, b, f, g)
        ctx.eps, ctx.mode, ctx.steps = float(eps), mode, int(steps)
        return f, g

    @staticmethod
    def backward(ctx, gf, gg):
        x, y, a, b, f, g = ctx.saved_tensors
        gf = torch.zeros_like(f) if gf is None else gf
        gg = torch.zeros_like(g) if gg is None else gg
        ops = HipPlanOps(x, y, a, b, f, g, ctx.eps)
        gx, gy, ga, gb = potentials_backward(ops, x.float(), y.float(), a.float(), b.float(),
                                             f.float(), g.float(), ctx.eps, gf.float(),
                                             gg.float(), ctx.mode, ctx.steps)
        return (gx.to(x.dtype), gy.to(y.dtype), ga.to(a.dtype), gb.to(b.dtype), None, None, None,
                None, None)


def potentials_backward(ops: PlanOps, x, y, a, b, f, g, eps, gf, gg, mode="implicit",
                        steps=300):
    """Gradients w.r.t. (x, y, a, b) of a loss of the gauge-fixed converged
    potentials (f, g) (<f, a> = <g, b>), given its gradients (gf, gg).  The
    gauge shift c = (<g, b> - <f, a>) / 2 passes s = sum gf - sum gg on:
    v = (gf - s a / 2, gg + s b / 2) (gauge invariant), plus direct weight
    terms -s f / 2 and s g / 2; weight gradients are centred."""
    s = gf.sum() - gg.sum()
    vf = gf - s * a / 2
    vg = gg + s * b / 2
    alpha, beta = solve_adjoint(ops, a, b, vf, vg, mode, steps)
    gx, gy, ga, gb = adjoint_param_grads(ops, x, y, a, b, eps, alpha, beta)
    ga = ga - s * f / 2
    gb = gb + s * g / 2
    return gx, gy, ga - ga.mean(), gb - gb.mean()


def sinkhorn_potentials(x: Tensor, y: Tensor, eps: float, a: Optional[Tensor] = None,
                        b: Optional[Tensor] = None, tol: float = 1e-6, max_iter: int = 20000,
                        mode: str = "implicit", steps: int = 300):
    """Differentiable converged potentials (f, g) of one pair of GPU point
    clouds; see :class:`SinkhornPotentials`."""
    n, m = x.shape[0], y.shape[0]
    a = torch.full((n,), 1.0 / n, device=x.device, dtype=x.dtype) if a is None else a
    b = torch.full((m,), 1.0 / m, device=x.device, dtype=x.dtype) if b is None else b
    return SinkhornPotentials.apply(x, y, a, b, eps, tol, max_iter, mode, steps)
