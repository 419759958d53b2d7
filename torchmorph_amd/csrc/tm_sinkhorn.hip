// Log-domain Sinkhorn for gfx950: ragged batches of point-cloud pairs with the
// squared-Euclidean cost computed on the fly (the histogram entry point of the
// reference, one shared (d, d) cost, is in tm_hist.hip).
//
// Point clouds (no reference counterpart: the reference only solves
// histograms sharing one cost, optimal_transport.py:189-213).  Dual potentials
// f (rows, points x, weights a) and g (columns, points y, weights b):
//   f_i <- -eps * log sum_j b_j exp((g_j - C_ij) / eps)
//   g_j <- -eps * log sum_i a_i exp((f_i - C_ij) / eps),  C_ij = |x_i - y_j|^2.
// Kernel arithmetic is in base 2: with L = log2(e),
//   (g_j - C_ij) L/eps + log2 b_j = h_j + xt_i . y_j - c_i,
//   h_j = L/eps (g_j - |y_j|^2) + log2 b_j,  xt_i = 2L/eps x_i,  c_i = L/eps |x_i|^2,
// so each term is three FMAs and one v_exp_f32; clouds are centred by the
// caller, which keeps the expansion accurate.  The sum for row i is taken
// relative to s_i = c_i - L/eps f_i(old); the resulting sum S_i is exactly
// the row-marginal ratio r_i / a_i of the current plan, so the L1 marginal
// error sum_i a_i |S_i - 1| of the state after the previous g-step comes for
// free with every f-update.  If S_i leaves [2^-60, 2^60] the workgroup
// recomputes with an exact per-row maximum.
//
// One iteration = f-update (into the other half of a ping-pong buffer, with
// the marginal error) -> g-update (in place) with the convergence check of
// the state before the f-update fused in front of it.  A pair whose error is
// <= tol (or that reached max_iter) is frozen at the state the error was
// measured on: its f, g and iteration count are copied out and later
// launches skip it.  The host replays a captured hipGraph of K iterations and
// reads the number of active pairs between replays.
//
// Work distribution: a device work-list of (pair, first row) tiles of 64, 16
// or 4 rows per half-step covers ragged pairs in one launch; a workgroup of
// 256 lanes stages the pair's column points and h_j in LDS and splits each
// row's sum over 256 / rows lanes.  The list is rebuilt for the active pairs
// as pairs converge.
#include <algorithm>
#include <cstring>
#include <vector>

#include "tm_common.h"

namespace tmh {

constexpr float kL2E = 1.4426950408889634f;  // log2(e)
constexpr int kRowsPerTile = 64;
constexpr int kMaxCols = 4096;

struct SkPair {
  int64_t xoff, yoff;  // first row of the pair in the concatenated x / y
  int32_t n, m;
  float eps;
  int32_t gbase;       // first 4-row error group of the pair (see sk_half_kernel)
};

// Per-pair centre (mean of the y cloud) subtracted from both clouds when
// points are loaded: the cost is invariant to a common translation, and
// centred coordinates keep the |x|^2 + |y|^2 - 2 x.y expansion accurate.
// Computed by a fixed-order tree reduction, so results are deterministic.
constexpr int kCtrStride = 4;

struct SkTile {
  int32_t pair;
  int32_t r0;
};

template <int D>
__device__ __forceinline__ float dotp(const float* a, const float* b) {
  float s = 0.f;
#pragma unroll
  for (int d = 0; d < D; ++d) s = fmaf(a[d], b[d], s);
  return s;
}

// Convergence check fused into the g-update of iteration k (1-based): the
// error of the state after iteration k-1 (f still in fold, g in the g
// buffer) is the sum of the f-update's 4-row group partials of the pair, by
// a fixed tree; every g-tile of the pair forms the same sum.  A converged
// pair (or one at max_iter) updates no g: its first tile copies that state
// out, sets its iteration count and error and deactivates it.
struct SkCheck {
  const float* fpart;         // f-update 4-row group partials of the L1 row-marginal error
  const int* iter_base;       // iterations before this chunk
  int k_in_chunk;
  int max_iter;
  float tol;
  const float* fold;          // f of the state the error belongs to
  float* fout;
  float* gout;
  int32_t* iters;
  float* err;
  int* active;
  int* n_active;
};

// One half-step for every tile of the work-list.
//   rows/cols: concatenated points (N x D); lwc: log2 of the column weights;
//   pc: column potentials; pin/pout: row potentials (old/new, may alias);
//   wr: plain row weights (for the marginal error).
// ROWS_ARE_X selects which offsets of SkPair index rows; CHECK (g-updates
// of the iteration loop) runs the fused convergence check first.
//
// A tile is RPT rows (64, 16 or 4) with 256 / RPT lanes per row.  The
// arithmetic does not depend on RPT, so the schedule can change tile shapes
// between iterations (more, smaller tiles for a few straggling pairs) with
// identical results: a row's sum runs over 64 interleaved column groups
// (group g holds the columns j = g mod 64, summed in column order; columns
// are padded to a multiple of 64 with terms that are exactly 0), and the 64
// group sums combine by one fixed pairwise tree (g + 32 into g, then + 16,
// ..., + 1), in registers where the partners share a lane and in LDS where
// they do not.  The row-marginal error is written per group of 4 rows
// (summed in row order) and a pair's error is summed over its groups by a
// fixed tree as well.
// The pair's L1 row-marginal error over its ceil(n / 4) error groups, called
// by lanes 0..63: lane l sums groups l, l + 64, ... in order, then a fixed
// tree over the lanes; the result is valid in lane 0.
__device__ __forceinline__ float sk_pair_error(const float* __restrict__ fpart, const SkPair& pr) {
  const int ng = (pr.n + 3) >> 2;
  float e = 0.f;
  for (int t = threadIdx.x; t < ng; t += 64) e += fpart[pr.gbase + t];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) e += __shfl_down(e, off, 64);
  return e;
}

// One half-step tile: rows tile.r0 .. tile.r0 + RPT - 1 of the pair (every
// thread of the workgroup calls it; barriers inside).
template <int D, bool ROWS_ARE_X, int RPT>
__device__ __forceinline__ void sk_tile(const SkTile tile, const SkPair& pr,
                                        const float* __restrict__ rows,
                                        const float* __restrict__ cols,
                                        const float* __restrict__ lwc,
                                        const float* __restrict__ pc,
                                        const float* __restrict__ pin, float* __restrict__ pout,
                                        const float* __restrict__ wr, float* __restrict__ errpart,
                                        int maxc, const float* __restrict__ centre) {
  static_assert(RPT == 64 || RPT == 16 || RPT == 4, "tile rows");
  constexpr int NQ = 256 / RPT;  // lanes per row
  constexpr int GPL = 64 / NQ;   // column groups per lane
  // column j in LDS as one float4: (y_0, .., y_{D-1}, h_j) for D <= 3, so a
  // term is one LDS read, three FMAs (h_j as the addend), one subtraction,
  // one v_exp_f32 and one add; D = 4 keeps h_j in its own array
  constexpr bool PACK = D <= 3;
  extern __shared__ float4 dyn4[];
  const int maxcp = (maxc + 63) & ~63;
  float4* s4 = dyn4;                    // [maxcp] points (+ h_j when PACK)
  float* shx = (float*)(dyn4 + maxcp);  // [maxcp] h_j when !PACK
  __shared__ float ssum[NQ][RPT];
  __shared__ float serr4[RPT];
  __shared__ int sbad;
  const int64_t roff = ROWS_ARE_X ? pr.xoff : pr.yoff;
  const int64_t coff = ROWS_ARE_X ? pr.yoff : pr.xoff;
  const int nr = ROWS_ARE_X ? pr.n : pr.m;
  const int nc = ROWS_ARE_X ? pr.m : pr.n;
  const int ncp = (nc + 63) & ~63;
  const float k1 = kL2E / pr.eps;  // L / eps
  float ctr[D];
#pragma unroll
  for (int d = 0; d < D; ++d) ctr[d] = centre[tile.pair * kCtrStride + d];
  // stage the pair's columns; padding columns have h = -inf (terms exactly 0)
  for (int j = threadIdx.x; j < ncp; j += blockDim.x) {
    float y[4] = {0.f, 0.f, 0.f, 0.f};
    float h = -INFINITY;
    if (j < nc) {
      float yy = 0.f;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        y[d] = cols[(coff + j) * D + d] - ctr[d];
        yy = fmaf(y[d], y[d], yy);
      }
      h = fmaf(k1, pc[coff + j] - yy, lwc[coff + j]);
    }
    if (PACK) {
      s4[j] = make_float4(y[0], y[1], y[2], h);
    } else {
      s4[j] = make_float4(y[0], y[1], y[2], y[3]);
      shx[j] = h;
    }
  }
  if (threadIdx.x == 0) sbad = 0;
  __syncthreads();
  const int r = threadIdx.x % RPT;
  const int q = threadIdx.x / RPT;  // this lane's first column group
  const int i = tile.r0 + r;
  const bool live = i < nr;
  float xt[4] = {0.f, 0.f, 0.f, 0.f};
  float ci = 0.f, s = 0.f, fold = 0.f;
  if (live) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float x = rows[(roff + i) * D + d] - ctr[d];
      xt[d] = 2.f * k1 * x;
      ci = fmaf(x, x, ci);
    }
    ci *= k1;
    fold = pin[roff + i];
    s = ci - k1 * fold;
  }
  // t_j = h_j + xt . y_j
  auto term = [&](int j) -> float {
    const float4 v = s4[j];
    float t = PACK ? v.w : shx[j];
    t = fmaf(xt[0], v.x, t);
    if (D >= 2) t = fmaf(xt[1], v.y, t);
    if (D >= 3) t = fmaf(xt[2], v.z, t);
    if (D >= 4) t = fmaf(xt[3], v.w, t);
    return t;
  };
  // canonical row sum of 2^(t_j - shift); the result is valid in the q == 0
  // lanes.  Every thread of the workgroup calls it (barriers inside).
  auto row_sum = [&](float shift) -> float {
    float acc[GPL];
#pragma unroll
    for (int k = 0; k < GPL; ++k) acc[k] = 0.f;
    if (live)
      for (int j0 = 0; j0 < ncp; j0 += 64) {
#pragma unroll
        for (int k = 0; k < GPL; ++k) acc[k] += __builtin_amdgcn_exp2f(term(j0 + q + NQ * k) - shift);
      }
#pragma unroll
    for (int dk = GPL / 2; dk >= 1; dk >>= 1)
#pragma unroll
      for (int k = 0; k < dk; ++k) acc[k] += acc[k + dk];
    ssum[q][r] = acc[0];
    __syncthreads();
#pragma unroll
    for (int dq = NQ / 2; dq >= 1; dq >>= 1) {
      if (q < dq) ssum[q][r] += ssum[q + dq][r];
      __syncthreads();
    }
    const float v = ssum[0][r];
    __syncthreads();  // ssum is reused by the next call
    return v;
  };
  float S = row_sum(s);
  const bool bad = live && q == 0 && !(S >= 0x1p-60f && S <= 0x1p60f);
  if (bad) sbad = 1;
  __syncthreads();
  if (sbad) {
    // exact path for the whole workgroup: the row maximum (any order), then
    // the canonical sum relative to it
    float mx = -INFINITY;
    if (live)
      for (int j = q; j < nc; j += NQ) mx = fmaxf(mx, term(j));
    ssum[q][r] = mx;
    __syncthreads();
    if (q == 0)
      for (int k = 1; k < NQ; ++k) mx = fmaxf(mx, ssum[k][r]);
    __syncthreads();
    if (q == 0) ssum[0][r] = mx;
    __syncthreads();
    mx = ssum[0][r];
    __syncthreads();
    const float S2 = row_sum(mx);
    // rebase to the stale shift: log2 S = mx - s + log2 S2
    if (live && q == 0) {
      const float l2 = mx - s + __builtin_amdgcn_logf(S2);
      S = exp2f(l2);  // may be inf/0 for the error only
      pout[roff + i] = fold - l2 / k1;
    }
  } else if (live && q == 0) {
    pout[roff + i] = fold - __builtin_amdgcn_logf(S) / k1;
  }
  if (errpart) {
    // row errors a_i |S_i - 1| summed per 4-row group in row order
    if (q == 0) serr4[r] = live ? wr[roff + i] * fabsf(S - 1.f) : 0.f;
    __syncthreads();
    if (threadIdx.x < RPT / 4 && tile.r0 + 4 * (int)threadIdx.x < nr) {
      const int b = 4 * threadIdx.x;
      errpart[pr.gbase + ((tile.r0 + b) >> 2)] =
          ((serr4[b] + serr4[b + 1]) + serr4[b + 2]) + serr4[b + 3];
    }
  }
}

template <int D, bool ROWS_ARE_X, bool CHECK, int RPT>
__global__ __launch_bounds__(256) void sk_half_kernel(
    const float* __restrict__ rows, const float* __restrict__ cols,
    const float* __restrict__ lwc, const float* __restrict__ pc, const float* __restrict__ pin,
    float* __restrict__ pout, const float* __restrict__ wr, const SkPair* __restrict__ pairs,
    const SkTile* __restrict__ tiles, const int* active, float* __restrict__ errpart,
    int maxc, const float* __restrict__ centre, SkCheck ck) {
  const SkTile tile = tiles[blockIdx.x];
  if (!active[tile.pair]) return;  // uniform per workgroup
  const SkPair pr = pairs[tile.pair];
  if constexpr (CHECK) {
    __shared__ int sdone;
    __shared__ float serr;
    __shared__ int sk;
    if (threadIdx.x < 64) {
      const float e = sk_pair_error(ck.fpart, pr);
      if (threadIdx.x == 0) {
        const int k = *ck.iter_base + ck.k_in_chunk;
        serr = e;
        sk = k;
        sdone = (e <= ck.tol || k - 1 >= ck.max_iter) ? 1 : 0;
      }
    }
    __syncthreads();
    if (sdone) {
      if (tile.r0 == 0) {
        for (int i = threadIdx.x; i < pr.n; i += blockDim.x)
          ck.fout[pr.xoff + i] = ck.fold[pr.xoff + i];
        for (int j = threadIdx.x; j < pr.m; j += blockDim.x)
          ck.gout[pr.yoff + j] = pin[pr.yoff + j];
        __syncthreads();
        if (threadIdx.x == 0) {
          ck.iters[tile.pair] = sk - 1;
          ck.err[tile.pair] = serr;
          ck.active[tile.pair] = 0;
          atomicSub(ck.n_active, 1);
        }
      }
      return;
    }
    if (tile.r0 == 0 && threadIdx.x == 0) ck.err[tile.pair] = serr;
  }
  sk_tile<D, ROWS_ARE_X, RPT>(tile, pr, rows, cols, lwc, pc, pin, pout, wr, errpart, maxc,
                              centre);
}



// Gradient of the OT value <f, a> + <g, b> w.r.t. the row points at the
// solution (envelope theorem): grad_i = 2 (S0_i x_i - S1_i) with
// S0_i = sum_j P_ij, S1_i = sum_j P_ij y_j (centred coordinates; the
// difference is translation invariant), P_ij = a_i 2^(e_ij - s_i).
template <int D, bool ROWS_ARE_X>
__global__ __launch_bounds__(256) void sk_grad_kernel(
    const float* __restrict__ rows, const float* __restrict__ cols,
    const float* __restrict__ lwc, const float* __restrict__ pc, const float* __restrict__ prow,
    const float* __restrict__ wr, const SkPair* __restrict__ pairs,
    const SkTile* __restrict__ tiles, int maxc, const float* __restrict__ centre,
    float* __restrict__ grad) {
  extern __shared__ float dyn[];
  float* sy = dyn;
  float* sh = dyn + maxc * D;
  __shared__ float s0[4][kRowsPerTile];
  __shared__ float s1[4][kRowsPerTile][D];
  const SkTile tile = tiles[blockIdx.x];
  const SkPair pr = pairs[tile.pair];
  const int64_t roff = ROWS_ARE_X ? pr.xoff : pr.yoff;
  const int64_t coff = ROWS_ARE_X ? pr.yoff : pr.xoff;
  const int nr = ROWS_ARE_X ? pr.n : pr.m;
  const int nc = ROWS_ARE_X ? pr.m : pr.n;
  const float k1 = kL2E / pr.eps;
  float ctr[D];
#pragma unroll
  for (int d = 0; d < D; ++d) ctr[d] = centre[tile.pair * kCtrStride + d];
  for (int j = threadIdx.x; j < nc; j += blockDim.x) {
    float yy = 0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float y = cols[(coff + j) * D + d] - ctr[d];
      sy[j * D + d] = y;
      yy = fmaf(y, y, yy);
    }
    sh[j] = fmaf(k1, pc[coff + j] - yy, lwc[coff + j]);
  }
  __syncthreads();
  const int r = threadIdx.x & (kRowsPerTile - 1);
  const int q = threadIdx.x >> 6;
  const int i = tile.r0 + r;
  const bool live = i < nr;
  float x[D], xt[D];
  float ci = 0.f, s = 0.f;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    x[d] = live ? rows[(roff + i) * D + d] - ctr[d] : 0.f;
    xt[d] = 2.f * k1 * x[d];
    ci = fmaf(x[d], x[d], ci);
  }
  if (live) s = k1 * ci - k1 * prow[roff + i];
  float a0 = 0.f, a1[D];
#pragma unroll
  for (int d = 0; d < D; ++d) a1[d] = 0.f;
  if (live) {
    for (int j = q; j < nc; j += 4) {
      const float w = __builtin_amdgcn_exp2f(sh[j] + dotp<D>(xt, &sy[j * D]) - s);
      a0 += w;
#pragma unroll
      for (int d = 0; d < D; ++d) a1[d] = fmaf(w, sy[j * D + d], a1[d]);
    }
  }
  s0[q][r] = a0;
#pragma unroll
  for (int d = 0; d < D; ++d) s1[q][r][d] = a1[d];
  __syncthreads();
  if (live && q == 0) {
    const float wi = wr[roff + i];
    const float S0 = (s0[0][r] + s0[1][r] + s0[2][r] + s0[3][r]) * wi;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float S1 = (s1[0][r][d] + s1[1][r][d] + s1[2][r][d] + s1[3][r][d]) * wi;
      grad[(roff + i) * D + d] = 2.f * (S0 * x[d] - S1);
    }
  }
}

// Products with the plan of converged potentials, formed on the fly:
// out[i, k] = sum_j P_ij V[j, k] for the tile's rows (K vectors), with
// P_ij = a_i 2^(e_ij - s_i) as in sk_grad_kernel.  Rows are x (P V) or y
// (P^T V).  Four lanes per row take the columns j = q mod 4; their partial
// sums add in lane order (deterministic).  V is read from global memory
// (every lane of a wavefront reads the same V[j], one broadcast load).
template <int D, bool ROWS_ARE_X, int K>
__global__ __launch_bounds__(256) void sk_planmat_kernel(
    const float* __restrict__ rows, const float* __restrict__ cols,
    const float* __restrict__ lwc, const float* __restrict__ pc, const float* __restrict__ prow,
    const float* __restrict__ wr, const SkPair* __restrict__ pairs,
    const SkTile* __restrict__ tiles, int maxc, const float* __restrict__ centre,
    const float* __restrict__ V, float* __restrict__ out) {
  extern __shared__ float dyn[];
  float* sy = dyn;
  float* sh = dyn + maxc * D;
  __shared__ float part[4][kRowsPerTile][K];
  const SkTile tile = tiles[blockIdx.x];
  const SkPair pr = pairs[tile.pair];
  const int64_t roff = ROWS_ARE_X ? pr.xoff : pr.yoff;
  const int64_t coff = ROWS_ARE_X ? pr.yoff : pr.xoff;
  const int nr = ROWS_ARE_X ? pr.n : pr.m;
  const int nc = ROWS_ARE_X ? pr.m : pr.n;
  const float k1 = kL2E / pr.eps;
  float ctr[D];
#pragma unroll
  for (int d = 0; d < D; ++d) ctr[d] = centre[tile.pair * kCtrStride + d];
  for (int j = threadIdx.x; j < nc; j += blockDim.x) {
    float yy = 0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float y = cols[(coff + j) * D + d] - ctr[d];
      sy[j * D + d] = y;
      yy = fmaf(y, y, yy);
    }
    sh[j] = fmaf(k1, pc[coff + j] - yy, lwc[coff + j]);
  }
  __syncthreads();
  const int r = threadIdx.x & (kRowsPerTile - 1);
  const int q = threadIdx.x >> 6;
  const int i = tile.r0 + r;
  const bool live = i < nr;
  float xt[D];
  float ci = 0.f, s = 0.f;
#pragma unroll
  for (int d = 0; d < D; ++d) {
    const float x = live ? rows[(roff + i) * D + d] - ctr[d] : 0.f;
    xt[d] = 2.f * k1 * x;
    ci = fmaf(x, x, ci);
  }
  if (live) s = k1 * ci - k1 * prow[roff + i];
  float acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0.f;
  if (live) {
    for (int j = q; j < nc; j += 4) {
      const float w = __builtin_amdgcn_exp2f(sh[j] + dotp<D>(xt, &sy[j * D]) - s);
      const float* vj = V + (coff + j) * K;
#pragma unroll
      for (int k = 0; k < K; ++k) acc[k] = fmaf(w, vj[k], acc[k]);
    }
  }
#pragma unroll
  for (int k = 0; k < K; ++k) part[q][r][k] = acc[k];
  __syncthreads();
  if (live && q == 0) {
    const float wi = wr[roff + i];
#pragma unroll
    for (int k = 0; k < K; ++k)
      out[(roff + i) * K + k] =
          (((part[0][r][k] + part[1][r][k]) + part[2][r][k]) + part[3][r][k]) * wi;
  }
}

// One workgroup per pair: centre[p] = mean of the pair's y points.
__global__ __launch_bounds__(256) void sk_centre_kernel(const float* __restrict__ y, int dim,
                                                        const SkPair* __restrict__ pairs,
                                                        float* __restrict__ centre) {
  __shared__ float red[256];
  const SkPair pr = pairs[blockIdx.x];
  for (int d = 0; d < kCtrStride; ++d) {
    float s = 0.f;
    if (d < dim)
      for (int j = threadIdx.x; j < pr.m; j += blockDim.x) s += y[(pr.yoff + j) * dim + d];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) centre[blockIdx.x * kCtrStride + d] = d < dim ? red[0] / (float)pr.m : 0.f;
    __syncthreads();
  }
}

__global__ void sk_init_kernel(int* active, int npairs, int* n_active, int* iter_base) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < npairs; p += gridDim.x * blockDim.x)
    active[p] = 1;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *n_active = npairs;
    *iter_base = 0;
  }
}

__global__ void sk_log2_kernel(const float* __restrict__ w, float* __restrict__ lw, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) lw[i] = __builtin_amdgcn_logf(w[i]);
}

// End of a chunk of k iterations: advances the iteration base and writes
// the active count and the active flags of every pair to host memory
// (pinned, mapped, fine-grained), where the host reads them once the chunk's
// event has completed.
__global__ void sk_chunk_end_kernel(int* iter_base, int k, const int* __restrict__ n_active,
                                    const int* __restrict__ active, int npairs,
                                    int* __restrict__ host) {
  if (threadIdx.x == 0) {
    *iter_base += k;
    host[0] = *n_active;
  }
  for (int p = threadIdx.x; p < npairs; p += blockDim.x) host[1 + p] = active[p];
}

__global__ void sk_copy_kernel(const float* __restrict__ src, float* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Host read-back of the iteration loop, one per host thread: (active count,
// active flags per pair) in pinned, device-mapped, coherent host memory that
// every chunk's last kernel overwrites, and the events of the last two
// chunks.  The host reads it after a chunk's event, possibly while the next
// chunk writes it: the count and the flags only decrease, so any mix of two
// consecutive chunks' values is still a valid (conservative) state.
struct SkReadback {
  int* host = nullptr;
  int* dev = nullptr;  // the same memory as the device addresses it
  size_t cap = 0;      // ints
  hipEvent_t ev[2] = {nullptr, nullptr};
};

inline int sk_readback_acquire(SkReadback& rb, size_t ints) {
  // a previous call ended with a stream synchronisation: nothing is pending
  if (rb.cap < ints) {
    if (rb.host) TM_CHECK_HIP(hipHostFree(rb.host));
    rb.host = rb.dev = nullptr;
    rb.cap = 0;
    TM_CHECK_HIP(hipHostMalloc((void**)&rb.host, ints * sizeof(int),
                               hipHostMallocMapped | hipHostMallocCoherent));
    TM_CHECK_HIP(hipHostGetDevicePointer((void**)&rb.dev, rb.host, 0));
    rb.cap = ints;
  }
  for (hipEvent_t& e : rb.ev)
    if (!e) TM_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  return TM_OK;
}

static thread_local SkReadback g_sk_readback;
static thread_local PinnedStage g_sk_tiles[2];

}  // namespace tmh

// ------------------------------------------------------------ C ABI: points --

extern "C" size_t tm_sinkhorn_points_workspace_bytes(int npairs, const int32_t* n,
                                                     const int32_t* m) {
  // tiles sized for the smallest tile shape (4 rows), error groups of 4 rows
  int64_t sx = 0, sy = 0, gx = 0, gy = 0;
  for (int p = 0; p < npairs; ++p) {
    sx += n[p];
    sy += m[p];
    gx += (n[p] + 3) / 4;
    gy += (m[p] + 3) / 4;
  }
  size_t b = 0;
  auto add = [&](size_t x) { b += (x + 255) & ~(size_t)255; };
  add(sizeof(tmh::SkPair) * npairs);
  add(sizeof(tmh::SkTile) * (gx + gy));
  add(sizeof(int32_t) * (npairs + 1));
  add(sizeof(float) * gx);        // errpart: one per 4-row group of x
  add(sizeof(int) * npairs);      // active
  add(sizeof(int) * 4);           // n_active, iter_base
  add(sizeof(float) * (sx + sy)); // log2 weights
  add(sizeof(float) * sx * 2);    // f ping-pong
  add(sizeof(float) * sy);        // g working copy
  add(sizeof(float) * tmh::kCtrStride * npairs);  // centres
  return b;
}

// Solves every pair to `tol` (L1 row-marginal error after a g-step) or
// `max_iter`.  f/g hold the initial potentials on entry (zeros = cold start;
// a warm start passes predicted f and any g, which is recomputed from f
// before the first iteration) and the solution on return.
extern "C" int tm_sinkhorn_points(const float* x, const float* y, const float* a, const float* b,
                                  int dim, int npairs, const int32_t* n, const int32_t* m,
                                  const float* eps, float tol, int max_iter, int chunk, float* f,
                                  float* g, int32_t* iters, float* err, void* workspace,
                                  size_t workspace_bytes, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(x && y && a && b && f && g && iters && err, "tm_sinkhorn_points: NULL argument");
  TM_CHECK_ARG(dim >= 1 && dim <= 4, "tm_sinkhorn_points: dim must be 1..4, got %d", dim);
  TM_CHECK_ARG(npairs >= 1, "tm_sinkhorn_points: npairs must be >= 1");
  TM_CHECK_ARG(max_iter >= 1 && chunk >= 1, "tm_sinkhorn_points: max_iter and chunk must be >= 1");
  hipStream_t st = (hipStream_t)stream;
  std::vector<SkPair> hp(npairs);
  std::vector<SkTile> htile;
  std::vector<int32_t> hfirst(npairs + 1);
  int64_t sx = 0, sy = 0, gx = 0;
  int maxc = 1;
  for (int p = 0; p < npairs; ++p) {
    TM_CHECK_ARG(n[p] >= 1 && m[p] >= 1, "tm_sinkhorn_points: empty cloud in pair %d", p);
    maxc = n[p] > maxc ? n[p] : maxc;
    maxc = m[p] > maxc ? m[p] : maxc;
    TM_CHECK_ARG(n[p] <= kMaxCols && m[p] <= kMaxCols,
                 "tm_sinkhorn_points: clouds are limited to %d points (pair %d)", kMaxCols, p);
    TM_CHECK_ARG(eps[p] > 0.f, "tm_sinkhorn_points: eps must be > 0 (pair %d)", p);
    TM_CHECK_ARG(gx + (n[p] + 3) / 4 < ((int64_t)1 << 31),
                 "tm_sinkhorn_points: more than 2^33 source points");
    hp[p] = SkPair{sx, sy, n[p], m[p], eps[p], (int32_t)gx};
    sx += n[p];
    sy += m[p];
    gx += (n[p] + 3) / 4;
  }
  for (int p = 0; p < npairs; ++p) {
    hfirst[p] = (int32_t)htile.size();
    for (int r = 0; r < n[p]; r += kRowsPerTile) htile.push_back(SkTile{p, r});
  }
  hfirst[npairs] = (int32_t)htile.size();
  const int64_t tx = (int64_t)htile.size();
  for (int p = 0; p < npairs; ++p)
    for (int r = 0; r < m[p]; r += kRowsPerTile) htile.push_back(SkTile{p, r});
  const int64_t ty = (int64_t)htile.size() - tx;
  const size_t need = tm_sinkhorn_points_workspace_bytes(npairs, n, m);
  if (!workspace || workspace_bytes < need) {
    set_error("tm_sinkhorn_points: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, need);
    return TM_ERR_WORKSPACE;
  }
  char* w = (char*)workspace;
  auto take = [&](size_t bytes) {
    char* p = w;
    w += (bytes + 255) & ~(size_t)255;
    return (void*)p;
  };
  SkPair* dpairs = (SkPair*)take(sizeof(SkPair) * npairs);
  int64_t gy = 0;
  for (int p = 0; p < npairs; ++p) gy += (m[p] + 3) / 4;
  SkTile* dtiles = (SkTile*)take(sizeof(SkTile) * (gx + gy));
  int32_t* dfirst = (int32_t*)take(sizeof(int32_t) * (npairs + 1));
  float* errpart = (float*)take(sizeof(float) * gx);
  int* active = (int*)take(sizeof(int) * npairs);
  int* ctr = (int*)take(sizeof(int) * 4);
  int* n_active = ctr;
  int* iter_base = ctr + 1;
  float* lw = (float*)take(sizeof(float) * (sx + sy));
  float* la = lw;
  float* lb = lw + sx;
  float* fbuf = (float*)take(sizeof(float) * sx * 2);
  float* gw = (float*)take(sizeof(float) * sy);
  float* centre = (float*)take(sizeof(float) * kCtrStride * npairs);
  // the small tables travel with the stream (pageable copies are staged
  // before this call returns)
  TM_CHECK_HIP(hipMemcpyAsync(dpairs, hp.data(), sizeof(SkPair) * npairs, hipMemcpyHostToDevice, st));
  TM_CHECK_HIP(hipMemcpyAsync(dtiles, htile.data(), sizeof(SkTile) * (tx + ty),
                              hipMemcpyHostToDevice, st));
  TM_CHECK_HIP(hipMemcpyAsync(dfirst, hfirst.data(), sizeof(int32_t) * (npairs + 1),
                              hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(sk_centre_kernel, dim3(npairs), dim3(256), 0, st, y, dim, dpairs, centre);
  hipLaunchKernelGGL(sk_init_kernel, dim3(grid_for(npairs, 256)), dim3(256), 0, st, active, npairs,
                     n_active, iter_base);
  hipLaunchKernelGGL(sk_log2_kernel, dim3(grid_for(sx, 256)), dim3(256), 0, st, a, la, sx);
  hipLaunchKernelGGL(sk_log2_kernel, dim3(grid_for(sy, 256)), dim3(256), 0, st, b, lb, sy);
  hipLaunchKernelGGL(sk_copy_kernel, dim3(grid_for(sx, 256)), dim3(256), 0, st, f, fbuf, sx);
  TM_LAUNCH_CHECK();
  float* fA = fbuf;
  float* fB = fbuf + sx;
  hipStream_t lst = st;  // launch stream: a private stream while capturing

  // tile counts and tile rows of the current work-list (rebuilt as pairs
  // converge)
  int64_t ntx = tx, nty = ty;
  int rpt = kRowsPerTile;
  SkCheck nock{};
  const size_t half_lds = (size_t)((maxc + 63) & ~63) * 20;
#define SK_HALF_R(D, RX, CH, R, ROWS, COLS, LWC, PC, PIN, POUT, WR, TILES, NT, ERR, CK)         \
  hipLaunchKernelGGL((sk_half_kernel<D, RX, CH, R>), dim3((unsigned)(NT)), dim3(256), half_lds, \
                     lst, ROWS, COLS, LWC, PC, PIN, POUT, WR, dpairs, TILES, active, ERR, maxc,  \
                     centre, CK)
#define SK_HALF(D, RX, CH, ...)                                                              \
  do {                                                                                       \
    if (rpt == 64) SK_HALF_R(D, RX, CH, 64, __VA_ARGS__);                                    \
    else if (rpt == 16) SK_HALF_R(D, RX, CH, 16, __VA_ARGS__);                               \
    else SK_HALF_R(D, RX, CH, 4, __VA_ARGS__);                                               \
  } while (0)
  // g-update from fcur; with `check`, the fused convergence check of the
  // state (fold, g) of iteration k_in_chunk - 1 runs first
  auto gstep = [&](const float* fcur, const float* fold, int k_in_chunk, bool check) -> int {
    if (nty == 0) return TM_OK;
    SkCheck ck{errpart, iter_base, k_in_chunk, max_iter, tol, fold, f, g, iters, err, active,
               n_active};
    const SkTile* gt = dtiles + ntx;
    if (check) {
      switch (dim) {
        case 1: SK_HALF(1, false, true, y, x, la, fcur, gw, gw, b, gt, nty, nullptr, ck); break;
        case 2: SK_HALF(2, false, true, y, x, la, fcur, gw, gw, b, gt, nty, nullptr, ck); break;
        case 3: SK_HALF(3, false, true, y, x, la, fcur, gw, gw, b, gt, nty, nullptr, ck); break;
        default: SK_HALF(4, false, true, y, x, la, fcur, gw, gw, b, gt, nty, nullptr, ck); break;
      }
    } else {
      switch (dim) {
        case 1: SK_HALF(1, false, false, y, x, la, fcur, gw, gw, b, gt, nty, nullptr, nock); break;
        case 2: SK_HALF(2, false, false, y, x, la, fcur, gw, gw, b, gt, nty, nullptr, nock); break;
        case 3: SK_HALF(3, false, false, y, x, la, fcur, gw, gw, b, gt, nty, nullptr, nock); break;
        default: SK_HALF(4, false, false, y, x, la, fcur, gw, gw, b, gt, nty, nullptr, nock); break;
      }
    }
    TM_LAUNCH_CHECK();
    return TM_OK;
  };
  auto fstep = [&](const float* fin, float* fo) -> int {
    if (ntx == 0) return TM_OK;
    switch (dim) {
      case 1: SK_HALF(1, true, false, x, y, lb, gw, fin, fo, a, dtiles, ntx, errpart, nock); break;
      case 2: SK_HALF(2, true, false, x, y, lb, gw, fin, fo, a, dtiles, ntx, errpart, nock); break;
      case 3: SK_HALF(3, true, false, x, y, lb, gw, fin, fo, a, dtiles, ntx, errpart, nock); break;
      default: SK_HALF(4, true, false, x, y, lb, gw, fin, fo, a, dtiles, ntx, errpart, nock); break;
    }
    TM_LAUNCH_CHECK();
    return TM_OK;
  };
#undef SK_HALF
#undef SK_HALF_R
  // state 0: g from the initial f (g's own initial values only seed the shift)
  hipLaunchKernelGGL(sk_copy_kernel, dim3(grid_for(sy, 256)), dim3(256), 0, st, g, gw, sy);
  TM_LAUNCH_CHECK();
  if (gstep(fA, nullptr, 0, false) != TM_OK) return TM_ERR_HIP;

  // one chunk = `chunk` iterations (chunk even keeps the f ping-pong aligned);
  // an iteration is an f-update and a g-update with the fused check
  const int K = chunk % 2 ? chunk + 1 : chunk;
  if (sk_readback_acquire(g_sk_readback, (size_t)npairs + 1) != TM_OK) return TM_ERR_HIP;
  auto enqueue_chunk = [&]() -> int {
    float* fo = fA;
    float* fn = fB;
    for (int k = 1; k <= K; ++k) {
      if (fstep(fo, fn) != TM_OK) return TM_ERR_HIP;
      if (gstep(fn, fo, k, true) != TM_OK) return TM_ERR_HIP;
      float* t = fo;
      fo = fn;
      fn = t;
    }
    hipLaunchKernelGGL(sk_chunk_end_kernel, dim3(1), dim3(256), 0, lst, iter_base, K, n_active,
                       active, npairs, g_sk_readback.dev);
    TM_LAUNCH_CHECK();
    return TM_OK;
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  const bool outer_capture = hipStreamIsCapturing(st, &cs) == hipSuccess &&
                             cs != hipStreamCaptureStatusNone;
  TM_CHECK_ARG(!outer_capture, "tm_sinkhorn_points synchronises on convergence and cannot be "
                               "captured; capture tm_sinkhorn_points_iterate instead");
  // capture one chunk on a private stream (the caller's may be the null
  // stream, which cannot capture), replay it on the caller's stream; captured
  // again whenever the work-list shrinks
  hipStream_t cap = nullptr;
  bool use_graph = getenv("TORCHMORPH_AMD_NO_GRAPH") == nullptr &&
                   hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) == hipSuccess;
  // executable graphs replaced by a compaction may still have a replay in
  // flight: they are destroyed after the final synchronisation
  std::vector<hipGraphExec_t> retired;
  auto build_graph = [&]() {
    if (exec) retired.push_back(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    bool ok = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
    if (ok) {
      lst = cap;
      const int s1 = enqueue_chunk();
      lst = st;
      ok = hipStreamEndCapture(cap, &graph) == hipSuccess && s1 == TM_OK;
      if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
    }
    if (!ok) {
      if (graph) (void)hipGraphDestroy(graph);
      graph = nullptr;
      exec = nullptr;
      (void)hipGetLastError();
    }
    return ok;
  };
  if (use_graph) use_graph = build_graph();
  // Work-list compaction: once the active pairs fall below 3/4 of the pairs
  // in the list, the tile tables are rebuilt for the active pairs only, with
  // the largest tile shape (64, 16 or 4 rows) that still gives at least
  // kTailTiles tiles per launch, or else the smallest shape whose tiles per
  // launch the GPU holds at once (256 CUs, each with up to 8 workgroups of
  // half_lds bytes of column staging in its 160 KB of LDS): the long tail
  // of a batch launches only the tiles of its stragglers, spread over many
  // small tiles (the results do not depend on the tile shape).
  // The new tables go up through pinned stages (two, alternating), ordered
  // on the stream after the chunk in flight, which still uses the old ones.
  constexpr int64_t kTailTiles = 1024;
  const int64_t resident =
      (int64_t)256 * std::max<int64_t>(1, std::min<int64_t>(8, (160 << 10) / (int64_t)half_lds));
  int listed = npairs;
  int nstage = 0;
  auto upload = [&](const void* tiles, size_t tb, const void* firsts, size_t fb) -> int {
    PinnedStage& ps = g_sk_tiles[nstage++ & 1];
    int rc = stage_acquire(ps, align256(tb) + fb);
    if (rc != TM_OK) return rc;
    memcpy(ps.host, tiles, tb);
    memcpy((char*)ps.host + align256(tb), firsts, fb);
    if (tb) TM_CHECK_HIP(hipMemcpyAsync(dtiles, ps.host, tb, hipMemcpyHostToDevice, st));
    TM_CHECK_HIP(hipMemcpyAsync(dfirst, (char*)ps.host + align256(tb), fb, hipMemcpyHostToDevice,
                                st));
    TM_CHECK_HIP(hipEventRecord(ps.ev, st));
    ps.pending = true;
    return TM_OK;
  };
  auto compact = [&](const int* hact) -> int {
    // tiles per launch at each shape: the larger of the f and g lists
    int64_t tf[3] = {0, 0, 0}, tg[3] = {0, 0, 0};
    const int shapes[3] = {64, 16, 4};
    for (int p = 0; p < npairs; ++p)
      if (hact[p])
        for (int k = 0; k < 3; ++k) {
          tf[k] += (n[p] + shapes[k] - 1) / shapes[k];
          tg[k] += (m[p] + shapes[k] - 1) / shapes[k];
        }
    const int64_t T64 = std::max(tf[0], tg[0]), T16 = std::max(tf[1], tg[1]),
                  T4 = std::max(tf[2], tg[2]);
    if (T64 >= kTailTiles) rpt = 64;
    else if (T16 >= kTailTiles) rpt = T16 <= resident ? 16 : 64;
    else if (T4 <= resident) rpt = 4;
    else rpt = T16 <= resident ? 16 : 64;
    htile.clear();
    int kept = 0;
    for (int p = 0; p < npairs; ++p) {
      hfirst[p] = (int32_t)htile.size();
      if (!hact[p]) continue;
      ++kept;
      for (int r = 0; r < n[p]; r += rpt) htile.push_back(SkTile{p, r});
    }
    hfirst[npairs] = (int32_t)htile.size();
    ntx = (int64_t)htile.size();
    for (int p = 0; p < npairs; ++p)
      if (hact[p])
        for (int r = 0; r < m[p]; r += rpt) htile.push_back(SkTile{p, r});
    nty = (int64_t)htile.size() - ntx;
    const int rc = upload(htile.data(), sizeof(SkTile) * htile.size(), hfirst.data(),
                          sizeof(int32_t) * (npairs + 1));
    if (rc != TM_OK) return rc;
    listed = kept;
    if (use_graph) use_graph = build_graph();
    return TM_OK;
  };
  // Pipelined replays: chunk c + 1 is enqueued before the host waits for
  // chunk c's active count, so the GPU does not idle through the host's
  // turnaround (or a compaction's capture).  The active flags come back with
  // the count, so a compaction works from flags one chunk old: pairs that
  // converged since stay listed until the next compaction and skip all work.
  // A chunk enqueued after the last pair converged only launches kernels
  // that return at once.
  SkReadback& rb = g_sk_readback;
  int status = TM_OK;
  const int64_t max_chunks = ((int64_t)max_iter + K) / K + ((max_iter + K) % K != 0);
  auto launch_chunk = [&](int slot) -> int {
    if (use_graph) {
      if (hipGraphLaunch(exec, st) != hipSuccess) return TM_ERR_HIP;
    } else if (enqueue_chunk() != TM_OK) {
      return TM_ERR_HIP;
    }
    return hipEventRecord(rb.ev[slot], st) == hipSuccess ? TM_OK : TM_ERR_HIP;
  };
  int64_t launched = 0;
  status = launch_chunk(0);
  if (status == TM_OK) launched = 1;
  for (int64_t c = 0; status == TM_OK; ++c) {
    if (launched < max_chunks) {
      if ((status = launch_chunk((int)(launched & 1))) != TM_OK) break;
      ++launched;
    }
    if (hipEventSynchronize(rb.ev[c & 1]) != hipSuccess) {
      status = TM_ERR_HIP;
      break;
    }
    const int* h = rb.host;
    const int host_active = h[0];
    if (host_active == 0 || c + 1 >= launched) break;
    if (4 * host_active < 3 * listed && (status = compact(h + 1)) != TM_OK) break;
  }
  // the replays still in flight return at once; the graphs they use are
  // released after them
  if (hipStreamSynchronize(st) != hipSuccess && status == TM_OK) status = TM_ERR_HIP;
  for (hipGraphExec_t e : retired) (void)hipGraphExecDestroy(e);
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  if (cap) (void)hipStreamDestroy(cap);
  if (status != TM_OK) {
    set_error("tm_sinkhorn_points: HIP failure in the iteration loop");
    return status;
  }
  return TM_OK;
}


extern "C" size_t tm_sinkhorn_points_grad_workspace_bytes(int npairs, const int32_t* n,
                                                          const int32_t* m) {
  return tm_sinkhorn_points_workspace_bytes(npairs, n, m);
}

// Gradients of the entropic OT value <f, a> + <g, b> of every pair w.r.t.
// its points, from converged potentials f, g (as returned by
// tm_sinkhorn_points): gx (sum n, dim), gy (sum m, dim).
extern "C" int tm_sinkhorn_points_grad(const float* x, const float* y, const float* a,
                                       const float* b, int dim, int npairs, const int32_t* n,
                                       const int32_t* m, const float* eps, const float* f,
                                       const float* g, float* gx, float* gy, void* workspace,
                                       size_t workspace_bytes, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(x && y && a && b && f && g && gx && gy, "tm_sinkhorn_points_grad: NULL argument");
  TM_CHECK_ARG(dim >= 1 && dim <= 4, "tm_sinkhorn_points_grad: dim must be 1..4");
  TM_CHECK_ARG(npairs >= 1, "tm_sinkhorn_points_grad: npairs must be >= 1");
  hipStream_t st = (hipStream_t)stream;
  std::vector<SkPair> hp(npairs);
  std::vector<SkTile> htile;
  int64_t sx = 0, sy = 0;
  int maxc = 1;
  for (int p = 0; p < npairs; ++p) {
    TM_CHECK_ARG(n[p] >= 1 && m[p] >= 1 && n[p] <= kMaxCols && m[p] <= kMaxCols,
                 "tm_sinkhorn_points_grad: cloud sizes must be in 1..%d", kMaxCols);
    TM_CHECK_ARG(eps[p] > 0.f, "tm_sinkhorn_points_grad: eps must be > 0");
    hp[p] = SkPair{sx, sy, n[p], m[p], eps[p], 0};
    sx += n[p];
    sy += m[p];
    maxc = n[p] > maxc ? n[p] : maxc;
    maxc = m[p] > maxc ? m[p] : maxc;
  }
  for (int p = 0; p < npairs; ++p)
    for (int r = 0; r < n[p]; r += kRowsPerTile) htile.push_back(SkTile{p, r});
  const int64_t tx = (int64_t)htile.size();
  for (int p = 0; p < npairs; ++p)
    for (int r = 0; r < m[p]; r += kRowsPerTile) htile.push_back(SkTile{p, r});
  const int64_t ty = (int64_t)htile.size() - tx;
  const size_t need = tm_sinkhorn_points_grad_workspace_bytes(npairs, n, m);
  if (!workspace || workspace_bytes < need) {
    set_error("tm_sinkhorn_points_grad: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, need);
    return TM_ERR_WORKSPACE;
  }
  char* w = (char*)workspace;
  auto take = [&](size_t bytes) {
    char* p = w;
    w += (bytes + 255) & ~(size_t)255;
    return (void*)p;
  };
  SkPair* dpairs = (SkPair*)take(sizeof(SkPair) * npairs);
  SkTile* dtiles = (SkTile*)take(sizeof(SkTile) * (tx + ty));
  float* lw = (float*)take(sizeof(float) * (sx + sy));
  float* centre = (float*)take(sizeof(float) * kCtrStride * npairs);
  float* la = lw;
  float* lb = lw + sx;
  TM_CHECK_HIP(hipMemcpyAsync(dpairs, hp.data(), sizeof(SkPair) * npairs, hipMemcpyHostToDevice, st));
  TM_CHECK_HIP(hipMemcpyAsync(dtiles, htile.data(), sizeof(SkTile) * (tx + ty),
                              hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(sk_centre_kernel, dim3(npairs), dim3(256), 0, st, y, dim, dpairs, centre);
  hipLaunchKernelGGL(sk_log2_kernel, dim3(grid_for(sx, 256)), dim3(256), 0, st, a, la, sx);
  hipLaunchKernelGGL(sk_log2_kernel, dim3(grid_for(sy, 256)), dim3(256), 0, st, b, lb, sy);
  TM_LAUNCH_CHECK();
  const size_t lds = (size_t)maxc * (dim + 1) * sizeof(float);
#define SK_GRAD(D)                                                                                \
  hipLaunchKernelGGL((sk_grad_kernel<D, true>), dim3((unsigned)tx), dim3(256), lds, st, x, y, lb, \
                     g, f, a, dpairs, dtiles, maxc, centre, gx);                                  \
  hipLaunchKernelGGL((sk_grad_kernel<D, false>), dim3((unsigned)ty), dim3(256), lds, st, y, x,    \
                     la, f, g, b, dpairs, dtiles + tx, maxc, centre, gy)
  switch (dim) {
    case 1: SK_GRAD(1); break;
    case 2: SK_GRAD(2); break;
    case 3: SK_GRAD(3); break;
    default: SK_GRAD(4); break;
  }
#undef SK_GRAD
  TM_LAUNCH_CHECK();
  return TM_OK;
}

// out = P V (transpose = 0: V is (sum m, k), out (sum n, k)) or P^T V
// (transpose = 1: V (sum n, k), out (sum m, k)) for every pair, with the plan
// of the potentials f, g formed on the fly; k in 1..8.  Workspace as
// tm_sinkhorn_points_grad.
extern "C" int tm_sinkhorn_points_planmat(const float* x, const float* y, const float* a,
                                          const float* b, int dim, int npairs, const int32_t* n,
                                          const int32_t* m, const float* eps, const float* f,
                                          const float* g, int transpose, int k, const float* V,
                                          float* out, void* workspace, size_t workspace_bytes,
                                          void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(x && y && a && b && f && g && V && out, "tm_sinkhorn_points_planmat: NULL argument");
  TM_CHECK_ARG(dim >= 1 && dim <= 4, "tm_sinkhorn_points_planmat: dim must be 1..4");
  TM_CHECK_ARG(npairs >= 1, "tm_sinkhorn_points_planmat: npairs must be >= 1");
  TM_CHECK_ARG(k >= 1 && k <= 8, "tm_sinkhorn_points_planmat: k must be 1..8, got %d", k);
  hipStream_t st = (hipStream_t)stream;
  std::vector<SkPair> hp(npairs);
  std::vector<SkTile> htile;
  int64_t sx = 0, sy = 0;
  int maxc = 1;
  for (int p = 0; p < npairs; ++p) {
    TM_CHECK_ARG(n[p] >= 1 && m[p] >= 1 && n[p] <= kMaxCols && m[p] <= kMaxCols,
                 "tm_sinkhorn_points_planmat: cloud sizes must be in 1..%d", kMaxCols);
    TM_CHECK_ARG(eps[p] > 0.f, "tm_sinkhorn_points_planmat: eps must be > 0");
    hp[p] = SkPair{sx, sy, n[p], m[p], eps[p], 0};
    sx += n[p];
    sy += m[p];
    maxc = n[p] > maxc ? n[p] : maxc;
    maxc = m[p] > maxc ? m[p] : maxc;
  }
  for (int p = 0; p < npairs; ++p) {
    const int rows = transpose ? m[p] : n[p];
    for (int r = 0; r < rows; r += kRowsPerTile) htile.push_back(SkTile{p, r});
  }
  const int64_t nt = (int64_t)htile.size();
  const size_t need = tm_sinkhorn_points_grad_workspace_bytes(npairs, n, m);
  if (!workspace || workspace_bytes < need) {
    set_error("tm_sinkhorn_points_planmat: workspace of %zu bytes is smaller than the %zu "
              "required", workspace_bytes, need);
    return TM_ERR_WORKSPACE;
  }
  char* w = (char*)workspace;
  auto take = [&](size_t bytes) {
    char* p = w;
    w += (bytes + 255) & ~(size_t)255;
    return (void*)p;
  };
  SkPair* dpairs = (SkPair*)take(sizeof(SkPair) * npairs);
  SkTile* dtiles = (SkTile*)take(sizeof(SkTile) * nt);
  float* lw = (float*)take(sizeof(float) * (sx + sy));
  float* centre = (float*)take(sizeof(float) * kCtrStride * npairs);
  float* la = lw;
  float* lb = lw + sx;
  TM_CHECK_HIP(hipMemcpyAsync(dpairs, hp.data(), sizeof(SkPair) * npairs, hipMemcpyHostToDevice, st));
  TM_CHECK_HIP(hipMemcpyAsync(dtiles, htile.data(), sizeof(SkTile) * nt, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(sk_centre_kernel, dim3(npairs), dim3(256), 0, st, y, dim, dpairs, centre);
  hipLaunchKernelGGL(sk_log2_kernel, dim3(grid_for(sx, 256)), dim3(256), 0, st, a, la, sx);
  hipLaunchKernelGGL(sk_log2_kernel, dim3(grid_for(sy, 256)), dim3(256), 0, st, b, lb, sy);
  TM_LAUNCH_CHECK();
  const size_t lds = (size_t)maxc * (dim + 1) * sizeof(float);
#define SK_PM(D, K)                                                                            \
  if (transpose)                                                                               \
    hipLaunchKernelGGL((sk_planmat_kernel<D, false, K>), dim3((unsigned)nt), dim3(256), lds, st, \
                       y, x, la, f, g, b, dpairs, dtiles, maxc, centre, V, out);               \
  else                                                                                         \
    hipLaunchKernelGGL((sk_planmat_kernel<D, true, K>), dim3((unsigned)nt), dim3(256), lds, st, \
                       x, y, lb, g, f, a, dpairs, dtiles, maxc, centre, V, out)
#define SK_PM_K(D)                                                                             \
  switch (k) {                                                                                 \
    case 1: SK_PM(D, 1); break;                                                                \
    case 2: SK_PM(D, 2); break;                                                                \
    case 3: SK_PM(D, 3); break;                                                                \
    case 4: SK_PM(D, 4); break;                                                                \
    case 5: SK_PM(D, 5); break;                                                                \
    case 6: SK_PM(D, 6); break;                                                                \
    case 7: SK_PM(D, 7); break;                                                                \
    default: SK_PM(D, 8); break;                                                               \
  }
  switch (dim) {
    case 1: SK_PM_K(1); break;
    case 2: SK_PM_K(2); break;
    case 3: SK_PM_K(3); break;
    default: SK_PM_K(4); break;
  }
#undef SK_PM_K
#undef SK_PM
  TM_LAUNCH_CHECK();
  return TM_OK;
}
