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
// so each exponent is a dot product of (y_j, h_j) with (xt_i, 1), a matrix
// product over blocks of rows and columns, less s_i, then one v_exp_f32; clouds
// are centred by the caller, which keeps the expansion accurate.  The sum for row i is taken
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
// Work distribution: a device work-list of (pair, first row) tiles of 128, 64
// or 32 rows per half-step covers ragged pairs in one launch; a workgroup of
// four wavefronts stages the pair's column points and h_j in LDS and forms
// the exponents of 32 x 32 blocks on the matrix cores (three FMAs and a
// subtraction per term move off the vector ALU, which keeps the v_exp_f32
// and the add).  The list is rebuilt for the active pairs as pairs converge.
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

// A half-step work-list entry: a tile of rows with its pair's fields, so a
// workgroup reads one entry instead of a tile and then its pair.
struct SkTask {
  SkPair pr;
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
// A tile is RPT rows (128, 64 or 32) of one pair, computed by four wavefronts
// on matrix cores: a wavefront forms the exponents of 32 rows x 32 columns
// at a time with v_mfma_f32_32x32x2f32 (see SkMf) and keeps only the
// exponentials and the sums on the vector ALU.  The arithmetic does not
// depend on RPT, so the schedule can change tile shapes between iterations
// (more, smaller tiles for a few straggling pairs) with identical results.
// Columns are padded to a multiple of 128 with terms that are exactly 0 and
// cut into 32-column blocks; block b belongs to phase b mod 4.  A row's sum
// runs over 128 interleaved column groups (group g holds the columns
// j = g mod 128, summed in column order); the 32 groups of a phase combine by
// one fixed tree (the 16 groups a lane holds pairwise, r + 8 into r, then
// + 4, + 2, + 1; then the two lane halves), and the four phase sums as
// (P0 + P1) + (P2 + P3).  A wavefront covers 32 rows and RPT / 32 phases:
// all four in 128-row tiles, two in 64-row tiles (two wavefronts per 32
// rows), one in 32-row tiles (four wavefronts per 32 rows); phase sums of
// different wavefronts meet in LDS.  The row-marginal error is written per
// group of 4 rows (summed in row order) and a pair's error is summed over its
// groups by a fixed tree as well.

typedef float sk_f32x16 __attribute__((ext_vector_type(16)));

// Exponents as a matrix product.  Column j has the components
//   c = (y_0 .. y_{D-1}, h_j, 0 ..)
// and row i the components
//   r = (xt_0 .. xt_{D-1}, 1, 0 ..),
// so c . r = h_j + xt_i . y_j over K = 2 NK >= D + 1 entries.  One 32 x 32
// block is NK chained v_mfma_f32_32x32x2f32 whose accumulator starts at
// -shift_i (a lane holds one row, so that is one value per lane): MFMA m
// takes the component pair (2m, 2m + 1): lane l supplies component
// 2m + l / 32 of column l mod 32 (the A operand) and of row l mod 32 (the B
// operand), and receives the 16 results of row l mod 32 (columns
// 8 (k / 4) + 4 (l / 32) + k mod 4 of the block, k = 0..15).  Column
// components 0 .. D are staged in LDS (half h of the wavefront reads
// components h and 2 + h as one float2, and 4 + h as a float when D = 4);
// the rest are 0.  Matrix products and vector instructions do not overlap
// on this part, so the two MFMAs of D = 3 (K = 4) cost about as much as the
// v_exp_f32 and the add of the 16 terms a lane receives.
template <int D>
struct SkMf {
  static constexpr int NK = (D + 2) / 2;
  static constexpr int BYTES = D == 4 ? 24 : 16;  // LDS bytes per column
};

// The component pair of MFMA m for lane half hf, from the staged float2 p2
// and the extra float p1 (D = 4).
template <int D>
__device__ __forceinline__ float sk_col_comp(int m, float2 p2, float p1) {
  if (m == 0) return p2.x;
  if (m == 1) return p2.y;
  return p1;  // component 4 + hf
}

// One half-step tile: rows tile.r0 .. tile.r0 + RPT - 1 of the pair (every
// thread of the workgroup calls it; barriers inside).
template <int D, bool ROWS_ARE_X, bool CHECK, int RPT>
__device__ __forceinline__ void sk_tile(const SkTask& tk, const float* __restrict__ rows,
                                        const float* __restrict__ cols,
                                        const float* __restrict__ lwc,
                                        const float* __restrict__ pc,
                                        const float* __restrict__ pin, float* __restrict__ pout,
                                        const float* __restrict__ wr, const int* active,
                                        float* __restrict__ errpart, int maxc,
                                        const float* __restrict__ centre, const SkCheck& ck) {
  static_assert(RPT == 128 || RPT == 64 || RPT == 32, "tile rows");
  constexpr int NK = SkMf<D>::NK;
  constexpr int RB = RPT / 32;  // 32-row blocks of the tile
  constexpr int WPR = 4 / RB;   // wavefronts per row block
  constexpr int NPH = RB;       // phases per wavefront
  extern __shared__ float2 dyn2[];
  const int maxcp = (maxc + 127) & ~127;
  float2* sp2 = dyn2;                        // [2][maxcp]: components (h, 2 + h)
  float* sp1 = (float*)(dyn2 + 2 * maxcp);   // [2][maxcp]: component 4 + h (D = 4)
  __shared__ float sq[RB][WPR][32];
  __shared__ float serr4[RPT];
  __shared__ int sbad;
  const SkPair& pr = tk.pr;
  const int64_t roff = ROWS_ARE_X ? pr.xoff : pr.yoff;
  const int64_t coff = ROWS_ARE_X ? pr.yoff : pr.xoff;
  const int nr = ROWS_ARE_X ? pr.n : pr.m;
  const int nc = ROWS_ARE_X ? pr.m : pr.n;
  const int ncp = (nc + 127) & ~127;
  const float k1 = kL2E / pr.eps;  // L / eps
  // Everything below the task entry is issued before anything is waited
  // for: the active flag, the pair's error partials (CHECK), this lane's row
  // and the column staging; a converged pair's workgroup returns after the
  // staging barrier.
  const int act = active[tk.pair];
  constexpr int kEp = kMaxCols / 4 / 64;  // error groups per lane, at most
  // CHECK: error groups lane, lane + 64, ... of the pair, loaded by every
  // wavefront (the first one uses them) unconditionally at clamped indices
  // (group 0 exists for every pair) and masked where they are summed: no
  // load waits behind a branch
  const int ng = (pr.n + 3) >> 2;
  float ep[kEp];
  if (CHECK) {
#pragma unroll
    for (int u = 0; u < kEp; ++u) {
      const int t = (int)(threadIdx.x & 63) + 64 * u;
      ep[u] = ck.fpart[pr.gbase + (t < ng ? t : 0)];
    }
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int hf = lane >> 5;
  const int rb = wave / WPR, wq = wave % WPR;
  const int i = tk.r0 + rb * 32 + (lane & 31);
  const bool live = i < nr;
  const bool block_live = tk.r0 + rb * 32 < nr;  // uniform per wavefront
  float xr[D], fold = 0.f, wi = 0.f;
#pragma unroll
  for (int d = 0; d < D; ++d) xr[d] = live ? rows[(roff + i) * D + d] : 0.f;
  if (live) {
    fold = pin[roff + i];
    if (errpart) wi = wr[roff + i];
  }
  float ctr[D];
#pragma unroll
  for (int d = 0; d < D; ++d) ctr[d] = centre[tk.pair * kCtrStride + d];
  // stage the pair's columns; padding columns (and weights of 0) have
  // h = -2^100, whose terms are exactly 0
  for (int j = threadIdx.x; j < ncp; j += blockDim.x) {
    float c[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float h = -0x1p100f;
    if (j < nc) {
      float yy = 0.f;
#pragma unroll
      for (int d = 0; d < D; ++d) {
        c[d] = cols[(coff + j) * D + d] - ctr[d];
        yy = fmaf(c[d], c[d], yy);
      }
      h = fmaf(k1, pc[coff + j] - yy, lwc[coff + j]);
      h = h < -0x1p100f ? -0x1p100f : h;  // NaN stays NaN
    }
    c[D] = h;
    sp2[j] = make_float2(c[0], c[2]);
    sp2[maxcp + j] = make_float2(c[1], c[3]);
    if (D == 4) {
      sp1[j] = c[4];
      sp1[maxcp + j] = c[5];
    }
  }
  if (threadIdx.x == 0) sbad = 0;
  __syncthreads();
  if (!act) return;  // uniform per workgroup
  if constexpr (CHECK) {
    // the pair's L1 row-marginal error over its ceil(n / 4) error groups:
    // lane l sums groups l, l + 64, ... in order, then a fixed tree over the
    // lanes (every g-tile of the pair forms the same sum)
    __shared__ int sdone;
    __shared__ float serr;
    __shared__ int sk;
    if (threadIdx.x < 64) {
      float e = 0.f;
#pragma unroll
      for (int u = 0; u < kEp; ++u) e += (int)threadIdx.x + 64 * u < ng ? ep[u] : 0.f;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) e += __shfl_down(e, off, 64);
      if (threadIdx.x == 0) {
        const int k = *ck.iter_base + ck.k_in_chunk;
        serr = e;
        sk = k;
        sdone = (e <= ck.tol || k - 1 >= ck.max_iter) ? 1 : 0;
      }
    }
    __syncthreads();
    if (sdone) {
      // converged (or at max_iter): the first tile copies the state the
      // error belongs to out and deactivates the pair
      if (tk.r0 == 0) {
        for (int r = threadIdx.x; r < pr.n; r += blockDim.x)
          ck.fout[pr.xoff + r] = ck.fold[pr.xoff + r];
        for (int c = threadIdx.x; c < pr.m; c += blockDim.x)
          ck.gout[pr.yoff + c] = pin[pr.yoff + c];
        __syncthreads();
        if (threadIdx.x == 0) {
          ck.iters[tk.pair] = sk - 1;
          ck.err[tk.pair] = serr;
          ck.active[tk.pair] = 0;
          atomicSub(ck.n_active, 1);
        }
      }
      return;
    }
    if (tk.r0 == 0 && threadIdx.x == 0) ck.err[tk.pair] = serr;
  }
  // row components; rows past the pair are all 0 (their results are unused)
  float rc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float s = 0.f;
  if (live) {
    float ci = 0.f;
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const float x = xr[d] - ctr[d];
      rc[d] = 2.f * k1 * x;
      ci = fmaf(x, x, ci);
    }
    ci *= k1;
    s = ci - k1 * fold;
    rc[D] = 1.f;
  }
  const float2* cp2 = sp2 + hf * maxcp + (lane & 31);
  const float* cp1 = sp1 + hf * maxcp + (lane & 31);
  // B operands of this lane (row components 2m + hf)
  // (both values pinned to registers first: a select of two array entries
  // would otherwise become a select of addresses into scratch)
  float bo[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    float lo = rc[2 * m], hi = rc[2 * m + 1];
    asm volatile("" : "+v"(lo), "+v"(hi));
    bo[m] = hf ? hi : lo;
  }
  // the exponents t_j - shift of block b for this lane's row (c0 = -shift
  // in all 16 entries)
  auto block = [&](int b, const sk_f32x16& c0) -> sk_f32x16 {
    const float2 p2 = cp2[b * 32];
    const float p1 = D == 4 ? cp1[b * 32] : 0.f;
    sk_f32x16 acc =
        __builtin_amdgcn_mfma_f32_32x32x2f32(sk_col_comp<D>(0, p2, p1), bo[0], c0, 0, 0, 0);
    if (NK >= 2)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sk_col_comp<D>(1, p2, p1), bo[1], acc, 0, 0, 0);
    if (NK == 3)
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(sk_col_comp<D>(2, p2, p1), bo[2], acc, 0, 0, 0);
    return acc;
  };
  auto splat = [&](float v) {
    sk_f32x16 c0;
#pragma unroll
    for (int e = 0; e < 16; ++e) c0[e] = v;
    return c0;
  };
  // canonical row sum of 2^(t_j - shift), valid in every lane of the row
  auto row_sum = [&](float shift) -> float {
    const sk_f32x16 c0 = splat(live ? -shift : 0.f);
    float acc[NPH][16];
#pragma unroll
    for (int k = 0; k < NPH; ++k)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[k][e] = 0.f;
    if (block_live)
      for (int c = 0; c < ncp; c += 128) {
#pragma unroll
        for (int k = 0; k < NPH; ++k) {
          const sk_f32x16 t = block(c / 32 + wq * NPH + k, c0);
#pragma unroll
          for (int e = 0; e < 16; ++e) acc[k][e] += __builtin_amdgcn_exp2f(t[e]);
        }
      }
    float ph[NPH];
#pragma unroll
    for (int k = 0; k < NPH; ++k) {
#pragma unroll
      for (int dk = 8; dk >= 1; dk >>= 1)
#pragma unroll
        for (int e = 0; e < dk; ++e) acc[k][e] += acc[k][e + dk];
      ph[k] = acc[k][0] + __shfl_xor(acc[k][0], 32, 64);
    }
    float q = ph[0];
    if (NPH == 2) q = ph[0] + ph[1];
    if (NPH == 4) q = (ph[0] + ph[1]) + (ph[2] + ph[3]);
    if (WPR == 1) return q;
    if (hf == 0) sq[rb][wq][lane] = q;
    __syncthreads();
    const int l = lane & 31;
    const float v = WPR == 2 ? sq[rb][0][l] + sq[rb][1][l]
                             : (sq[rb][0][l] + sq[rb][1][l]) + (sq[rb][2][l] + sq[rb][3][l]);
    __syncthreads();  // sq is reused by the next call
    return v;
  };
  float S = row_sum(s);
  const bool writer = live && hf == 0 && wq == 0;  // one lane per row
  if (writer && !(S >= 0x1p-60f && S <= 0x1p60f)) sbad = 1;
  __syncthreads();
  if (sbad) {
    // exact path for the whole workgroup: the row maximum (any order), then
    // the canonical sum relative to it
    const sk_f32x16 c0 = splat(0.f);
    float mx = -INFINITY;
    if (block_live)
      for (int c = 0; c < ncp; c += 128) {
#pragma unroll
        for (int k = 0; k < NPH; ++k) {
          const sk_f32x16 t = block(c / 32 + wq * NPH + k, c0);
#pragma unroll
          for (int e = 0; e < 16; ++e) mx = fmaxf(mx, t[e]);
        }
      }
    mx = fmaxf(mx, __shfl_xor(mx, 32, 64));
    if (WPR > 1) {
      if (hf == 0) sq[rb][wq][lane] = mx;
      __syncthreads();
#pragma unroll
      for (int k = 0; k < WPR; ++k) mx = fmaxf(mx, sq[rb][k][lane & 31]);
      __syncthreads();
    }
    const float S2 = row_sum(mx);
    // rebase to the stale shift: log2 S = mx - s + log2 S2
    if (writer) {
      const float l2 = mx - s + __builtin_amdgcn_logf(S2);
      S = exp2f(l2);  // may be inf/0 for the error only
      pout[roff + i] = fold - l2 / k1;
    }
  } else if (writer) {
    pout[roff + i] = fold - __builtin_amdgcn_logf(S) / k1;
  }
  if (errpart) {
    // row errors a_i |S_i - 1| summed per 4-row group in row order
    if (hf == 0 && wq == 0) serr4[rb * 32 + lane] = live ? wi * fabsf(S - 1.f) : 0.f;
    __syncthreads();
    if (threadIdx.x < RPT / 4 && tk.r0 + 4 * (int)threadIdx.x < nr) {
      const int b = 4 * threadIdx.x;
      errpart[pr.gbase + ((tk.r0 + b) >> 2)] =
          ((serr4[b] + serr4[b + 1]) + serr4[b + 2]) + serr4[b + 3];
    }
  }
}

// At least four wavefronts per SIMD: within that register budget the
// matrix products write their results straight to VGPRs (no AGPR copies).
template <int D, bool ROWS_ARE_X, bool CHECK, int RPT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void sk_half_kernel(
    const float* __restrict__ rows, const float* __restrict__ cols,
    const float* __restrict__ lwc, const float* __restrict__ pc, const float* __restrict__ pin,
    float* __restrict__ pout, const float* __restrict__ wr, const SkTask* __restrict__ tasks,
    const int* active, float* __restrict__ errpart, int maxc, const float* __restrict__ centre,
    SkCheck ck) {
  const SkTask tk = tasks[blockIdx.x];
  sk_tile<D, ROWS_ARE_X, CHECK, RPT>(tk, rows, cols, lwc, pc, pin, pout, wr, active, errpart,
                                     maxc, centre, ck);
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
  // tile tables sized for one tile per 4 rows (more than any tile shape
  // needs), error groups of 4 rows
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
  add(sizeof(tmh::SkTask) * (gx + gy));
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
  std::vector<SkTask> htile;
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
  // Tile shape of a work-list: the largest (128, 64 or 32 rows) that still
  // gives at least kTailTiles tiles per launch, or else the smallest shape
  // whose tiles per launch the GPU holds at once (256 CUs, each with up to 4
  // workgroups within its 160 KB of LDS): the long tail of a batch launches
  // only the tiles of its stragglers, spread over many small tiles (the
  // results do not depend on the tile shape).
  const size_t half_lds = (size_t)((maxc + 127) & ~127) * (dim == 4 ? 24 : 16);
  constexpr int64_t kTailTiles = 1024;
  const int64_t resident =
      (int64_t)256 * std::max<int64_t>(1, std::min<int64_t>(4, (160 << 10) / (int64_t)half_lds));
  auto choose_rpt = [&](const int* act) {
    int64_t tf[3] = {0, 0, 0}, tg[3] = {0, 0, 0};
    const int shapes[3] = {128, 64, 32};
    for (int p = 0; p < npairs; ++p)
      if (!act || act[p])
        for (int k = 0; k < 3; ++k) {
          tf[k] += (n[p] + shapes[k] - 1) / shapes[k];
          tg[k] += (m[p] + shapes[k] - 1) / shapes[k];
        }
    const int64_t T128 = std::max(tf[0], tg[0]), T64 = std::max(tf[1], tg[1]),
                  T32 = std::max(tf[2], tg[2]);
    if (T128 >= kTailTiles) return 128;
    if (T64 >= kTailTiles) return T64 <= resident ? 64 : 128;
    if (T32 <= resident) return 32;
    return T64 <= resident ? 64 : 128;
  };
  int rpt = choose_rpt(nullptr);
  for (int p = 0; p < npairs; ++p) {
    hfirst[p] = (int32_t)htile.size();
    for (int r = 0; r < n[p]; r += rpt) htile.push_back(SkTask{hp[p], p, r});
  }
  hfirst[npairs] = (int32_t)htile.size();
  const int64_t tx = (int64_t)htile.size();
  for (int p = 0; p < npairs; ++p)
    for (int r = 0; r < m[p]; r += rpt) htile.push_back(SkTask{hp[p], p, r});
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
  SkTask* dtiles = (SkTask*)take(sizeof(SkTask) * (gx + gy));
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
  TM_CHECK_HIP(hipMemcpyAsync(dtiles, htile.data(), sizeof(SkTask) * (tx + ty),
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
  SkCheck nock{};
#define SK_HALF_R(D, RX, CH, R, ROWS, COLS, LWC, PC, PIN, POUT, WR, TILES, NT, ERR, CK)         \
  hipLaunchKernelGGL((sk_half_kernel<D, RX, CH, R>), dim3((unsigned)(NT)), dim3(256), half_lds, \
                     lst, ROWS, COLS, LWC, PC, PIN, POUT, WR, TILES, active, ERR, maxc, centre, CK)
#define SK_HALF(D, RX, CH, ...)                                                              \
  do {                                                                                       \
    if (rpt == 128) SK_HALF_R(D, RX, CH, 128, __VA_ARGS__);                                  \
    else if (rpt == 64) SK_HALF_R(D, RX, CH, 64, __VA_ARGS__);                               \
    else SK_HALF_R(D, RX, CH, 32, __VA_ARGS__);                                              \
  } while (0)
  // g-update from fcur; with `check`, the fused convergence check of the
  // state (fold, g) of iteration k_in_chunk - 1 runs first
  auto gstep = [&](const float* fcur, const float* fold, int k_in_chunk, bool check) -> int {
    if (nty == 0) return TM_OK;
    SkCheck ck{errpart, iter_base, k_in_chunk, max_iter, tol, fold, f, g, iters, err, active,
               n_active};
    const SkTask* gt = dtiles + ntx;
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
                               "captured into a graph");
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
  // A (re)capture is deferred to the next chunk launch, which enqueues that
  // chunk directly and then captures while the GPU runs it: the GPU does not
  // idle through the capture and instantiation.
  bool need_capture = use_graph;
  // Work-list compaction: once the active pairs fall below 3/4 of the pairs
  // in the list, the tile tables are rebuilt for the active pairs only, with
  // the tile shape of choose_rpt.  The new tables go up through pinned stages
  // (two, alternating), ordered on the stream after the chunk in flight, which
  // still uses the old ones.
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
    rpt = choose_rpt(hact);
    htile.clear();
    int kept = 0;
    for (int p = 0; p < npairs; ++p) {
      hfirst[p] = (int32_t)htile.size();
      if (!hact[p]) continue;
      ++kept;
      for (int r = 0; r < n[p]; r += rpt) htile.push_back(SkTask{hp[p], p, r});
    }
    hfirst[npairs] = (int32_t)htile.size();
    ntx = (int64_t)htile.size();
    for (int p = 0; p < npairs; ++p)
      if (hact[p])
        for (int r = 0; r < m[p]; r += rpt) htile.push_back(SkTask{hp[p], p, r});
    nty = (int64_t)htile.size() - ntx;
    const int rc = upload(htile.data(), sizeof(SkTask) * htile.size(), hfirst.data(),
                          sizeof(int32_t) * (npairs + 1));
    if (rc != TM_OK) return rc;
    listed = kept;
    need_capture = use_graph;
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
    if (use_graph && !need_capture) {
      if (hipGraphLaunch(exec, st) != hipSuccess) return TM_ERR_HIP;
    } else {
      if (enqueue_chunk() != TM_OK) return TM_ERR_HIP;
      if (need_capture) {
        use_graph = build_graph();
        need_capture = false;
      }
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
