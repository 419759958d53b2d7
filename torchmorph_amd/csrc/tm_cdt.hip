// Chamfer distance transform (taxicab / chessboard) for gfx950.
//
// Semantics are SciPy's distance_transform_cdt, which the reference tests pin
// (reference test/test_distance_transform.py:32-69, :289-320): int32 values
// start at 0 on background and -1 on foreground; a forward raster pass and a
// pass over the axis-reversed array each relax every nonzero voxel from the
// causal half of the 3^n metric structure, visited in structure order, taking a
// neighbour only when it strictly improves (the voxel's own value first, the
// in-row predecessor last).  The reference returns the values as float32 with
// -1 for unreachable voxels (cdt_kernel.cu:47-55); indices are the coordinates
// of the feature voxel.
//
// Parallel schedule with identical results: a row (fixed leading coordinates,
// the contiguous axis free) depends only on rows that precede it in raster
// order through the non-row neighbours, and on its own predecessor element.
//  * Rows are scheduled in wavefront steps t = sum_j 2^(m-1-j) c_j over the m
//    leading axes (2-D: t = y; 3-D: t = 2z + y), which every causal offset
//    strictly decreases, so rows of one step are independent; a workgroup
//    barrier separates steps.  ndim >= 4 runs rows one per step.
//  * Inside a row the recurrence v(x) = a(x) unless v(x-1)+1 < a(x) (a = the
//    value after the non-row neighbours) is a min-plus prefix scan of
//    w(x) = a(x) - x where the later position wins ties; one wavefront scans a
//    row in 64-element slots (element x in lane x % 64), so loads and stores
//    are coalesced along the contiguous axis.
// Pass 2 runs the same code on reversed coordinates, in place, as SciPy does.
//
// Without indices the values are computed separably instead (every line of
// an axis in parallel; tm_dt_lines.h shows why the result is the same).  Two
// launch variants, chosen by the autotuner, both exact:
//  * 0: one lane per line; the float output holds the int32 intermediates in
//    place, the chessboard envelope keeps its stack in the workspace;
//  * 1: chessboard axes >= 1 per voxel by expanding windows (cheb_window,
//    alternating between the workspace and the output), and the taxicab
//    pass along the contiguous last axis row-staged through LDS
//    (tm_dt_chunk.h).
#include "tm_cdt_dev.h"
#include "tm_dt_chunk.h"
#include "tm_dt_lines.h"
#include "tm_tune.h"

namespace tmh {

// One workgroup per item of a dense batch.
template <typename TIn>
__global__ __launch_bounds__(256) void cdt_pass_kernel(const TIn* __restrict__ in,
                                                       int32_t* __restrict__ dt,
                                                       int32_t* __restrict__ ft, Shape sh,
                                                       int mode, int metric, int pass2,
                                                       CdtOut o) {
  const int64_t base = (int64_t)blockIdx.x * sh.numel;
  CdtOut oi = o;
  if (oi.dist) oi.dist += base;
  // index planes keep the batch-wide plane stride; offset the item in them
  if (oi.idx) oi.idx = (char*)oi.idx + base * (oi.idx64 ? 8 : 4);
  cdt_item_pass(in + base, dt + base, ft + base, sh, mode, metric, pass2, oi);
}

// Separable distances, axis 0 from the mask.
template <typename TIn, bool FINAL>
__global__ __launch_bounds__(256) void cdt_sep0_kernel(const TIn* __restrict__ in,
                                                       int32_t* __restrict__ w, Shape sh,
                                                       int64_t nlines) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlines) return;
  int64_t c[kMaxNdim], base, b;
  decode_line(sh, 0, t, c, &base, &b);
  cdt_line_axis0<TIn, FINAL>(in, w, base, sh.stride[0], sh.size[0]);
}

// Separable distances, axis d >= 1 (metric 0 chessboard, 1 taxicab).  The
// chessboard stack entry k of line t sits at k * nlines + t.
template <int METRIC, bool FINAL>
__global__ __launch_bounds__(256) void cdt_sepd_kernel(int32_t* __restrict__ w, Shape sh, int d,
                                                       int64_t nlines, uint16_t* __restrict__ S,
                                                       uint16_t* __restrict__ T,
                                                       int32_t* __restrict__ G) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlines) return;
  int64_t c[kMaxNdim], base, b;
  decode_line(sh, d, t, c, &base, &b);
  if (METRIC == 1)
    cdt_line_taxicab<FINAL>(w, base, sh.stride[d], sh.size[d]);
  else
    cdt_line_chessboard<FINAL>(w, base, sh.stride[d], sh.size[d], S + t, T + t, G + t, nlines);
}

template <typename TIn>
static int cdt_separable_lanes(const TIn* in, int64_t batch, const Shape& sh, int metric,
                               float* dist, void* ws, hipStream_t st);

// Chessboard axis d >= 1, one voxel per thread (cheb_window), src -> dst.
template <bool FINAL>
__global__ __launch_bounds__(256) void cdt_window_kernel(const int32_t* __restrict__ src,
                                                         int32_t* __restrict__ dst, Shape sh,
                                                         int d, int64_t n) {
  const int64_t v = (int64_t)xcd_swizzle(blockIdx.x, gridDim.x) * 256 + threadIdx.x;
  if (v >= n) return;
  const int64_t step = sh.stride[d];
  const int L = (int)sh.size[d];
  const int i = (int)((v / step) % L);
  const int32_t r = cheb_window(src, v - (int64_t)i * step, step, L, i);
  dst[v] = FINAL ? cdt_final_bits(r) : r;
}

// Taxicab along the contiguous last axis (final pass), 256 rows per block.
__global__ __launch_bounds__(kCLanes) void cdt_rows_kernel(int32_t* __restrict__ w, Shape sh,
                                                           int64_t nrows_all) {
  extern __shared__ int32_t lds_rows[];
  const int64_t row0 = (int64_t)xcd_swizzle(blockIdx.x, gridDim.x) * kCLanes;
  const int nrows = (int)(nrows_all - row0 < kCLanes ? nrows_all - row0 : kCLanes);
  cdt_rows_taxicab<32>(w, sh, row0, nrows, lds_rows);
}

template <typename TIn>
static int cdt_separable(const TIn* in, int64_t batch, const Shape& sh, int metric, float* dist,
                         void* ws, hipStream_t st) {
  const int64_t N = batch * sh.numel;
  auto run = [&](int v) -> int {
    if (v == 1) {
      // chessboard: passes alternate so that the last one lands in the output
      int32_t* out = (int32_t*)dist;
      int32_t* buf = (int32_t*)ws;
      auto dst_of = [&](int d) { return ((sh.ndim - 1 - d) & 1) ? buf : out; };
      int32_t* w0 = metric == 0 ? dst_of(0) : out;
      const int64_t nlines = batch * (sh.numel / sh.size[0]);
      if (sh.ndim == 1)
        hipLaunchKernelGGL((cdt_sep0_kernel<TIn, true>), dim3(grid_for(nlines, 256)), dim3(256),
                           0, st, in, w0, sh, nlines);
      else
        hipLaunchKernelGGL((cdt_sep0_kernel<TIn, false>), dim3(grid_for(nlines, 256)), dim3(256),
                           0, st, in, w0, sh, nlines);
      TM_LAUNCH_CHECK();
      for (int d = 1; d < sh.ndim; ++d) {
        const bool fin = d == sh.ndim - 1;
        if (metric == 0) {
          const dim3 g(grid_for(N, 256));
          if (fin)
            hipLaunchKernelGGL((cdt_window_kernel<true>), g, dim3(256), 0, st, dst_of(d - 1),
                               dst_of(d), sh, d, N);
          else
            hipLaunchKernelGGL((cdt_window_kernel<false>), g, dim3(256), 0, st, dst_of(d - 1),
                               dst_of(d), sh, d, N);
        } else if (fin) {
          const int64_t nrows = N / sh.size[d];
          hipLaunchKernelGGL(cdt_rows_kernel, dim3(grid_for(nrows, kCLanes)), dim3(kCLanes),
                             (size_t)kCLanes * 33 * sizeof(int32_t), st, out, sh, nrows);
        } else {
          const int64_t nlines = batch * (sh.numel / sh.size[d]);
          hipLaunchKernelGGL((cdt_sepd_kernel<1, false>), dim3(grid_for(nlines, 256)), dim3(256),
                             0, st, out, sh, d, nlines, nullptr, nullptr, nullptr);
        }
        TM_LAUNCH_CHECK();
      }
      return TM_OK;
    }
    return cdt_separable_lanes(in, batch, sh, metric, dist, ws, st);
  };
  auto p2 = [](int64_t x) { int64_t q = 1; while (q < x) q <<= 1; return q; };
  int64_t key[kMaxNdim + 3];
  key[0] = metric;
  key[1] = sh.ndim;
  key[2] = p2(batch);
  for (int j = 0; j < sh.ndim; ++j) key[3 + j] = p2(sh.size[j]);
  bool ran = false;
  const int v = tune_pick("cdtsep", key, 3 + sh.ndim, 2, 1, run, st, &ran);
  return ran ? TM_OK : run(v);
}

template <typename TIn>
static int cdt_separable_lanes(const TIn* in, int64_t batch, const Shape& sh, int metric,
                               float* dist, void* ws, hipStream_t st) {
  const int64_t N = batch * sh.numel;
  int32_t* w = (int32_t*)dist;
  uint16_t* S = (uint16_t*)ws;
  uint16_t* T = S + N;
  int32_t* G = (int32_t*)(T + N);
  {
    const int64_t nlines = batch * (sh.numel / sh.size[0]);
    if (sh.ndim == 1)
      hipLaunchKernelGGL((cdt_sep0_kernel<TIn, true>), dim3(grid_for(nlines, 256)), dim3(256), 0,
                         st, in, w, sh, nlines);
    else
      hipLaunchKernelGGL((cdt_sep0_kernel<TIn, false>), dim3(grid_for(nlines, 256)), dim3(256),
                         0, st, in, w, sh, nlines);
    TM_LAUNCH_CHECK();
  }
  for (int d = 1; d < sh.ndim; ++d) {
    const int64_t nlines = batch * (sh.numel / sh.size[d]);
    const dim3 g(grid_for(nlines, 256)), blk(256);
    const bool fin = d == sh.ndim - 1;
    if (metric == 1) {
      if (fin) hipLaunchKernelGGL((cdt_sepd_kernel<1, true>), g, blk, 0, st, w, sh, d, nlines, S, T, G);
      else hipLaunchKernelGGL((cdt_sepd_kernel<1, false>), g, blk, 0, st, w, sh, d, nlines, S, T, G);
    } else {
      if (fin) hipLaunchKernelGGL((cdt_sepd_kernel<0, true>), g, blk, 0, st, w, sh, d, nlines, S, T, G);
      else hipLaunchKernelGGL((cdt_sepd_kernel<0, false>), g, blk, 0, st, w, sh, d, nlines, S, T, G);
    }
    TM_LAUNCH_CHECK();
  }
  return TM_OK;
}

template <typename TIn>
static int cdt_dispatch(const TIn* in, int64_t batch, int ndim, const int64_t* shape, int metric,
                        CdtOut o, void* ws, size_t ws_bytes, hipStream_t st) {
  Shape sh = make_shape(ndim, shape);
  o.plane = batch * sh.numel;
  const size_t need = tm_cdt_workspace_bytes(batch, ndim, shape);
  if (!ws || ws_bytes < need) {
    set_error("tm_cdt: workspace of %zu bytes is smaller than the %zu required", ws_bytes, need);
    return TM_ERR_WORKSPACE;
  }
  bool short_axes = true;
  for (int j = 0; j < ndim; ++j) short_axes = short_axes && shape[j] <= 65535;
  if (!o.idx && short_axes) return cdt_separable(in, batch, sh, metric, o.dist, ws, st);
  int32_t* dt = (int32_t*)ws;
  int32_t* ft = dt + o.plane;
  const int mode = ndim == 3 ? 1 : 0;
  const int threads = ndim == 3 ? 256 : 64;
  for (int pass = 0; pass < 2; ++pass) {
    hipLaunchKernelGGL((cdt_pass_kernel<TIn>), dim3((unsigned)batch), dim3(threads), 0, st, in, dt,
                       ft, sh, mode, metric, pass, o);
    TM_LAUNCH_CHECK();
  }
  return TM_OK;
}

}  // namespace tmh

extern "C" size_t tm_cdt_workspace_bytes(int64_t batch, int ndim, const int64_t* shape) {
  if (ndim < 1 || ndim > tmh::kMaxNdim || batch < 1) return 0;
  int64_t n = batch;
  for (int j = 0; j < ndim; ++j) n *= shape[j];
  return (size_t)n * 8;
}

extern "C" int tm_cdt(const void* input, int input_dtype, int64_t batch, int ndim,
                      const int64_t* shape, int metric, float* distances, void* indices,
                      int index_dtype, void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(input != nullptr, "tm_cdt: input is NULL");
  TM_CHECK_ARG(ndim >= 1 && ndim <= kMaxNdim, "tm_cdt: ndim must be in 1..8, got %d", ndim);
  TM_CHECK_ARG(batch >= 1 && batch <= 2147483647, "tm_cdt: batch out of range");
  TM_CHECK_ARG(metric == 0 || metric == 1, "tm_cdt: metric must be 0 (chessboard) or 1 (taxicab)");
  int64_t n = 1;
  for (int j = 0; j < ndim; ++j) {
    TM_CHECK_ARG(shape[j] >= 1, "tm_cdt: empty axis");
    n *= shape[j];
  }
  TM_CHECK_ARG(n < (int64_t)1 << 31, "tm_cdt: items above 2^31 voxels are not supported");
  TM_CHECK_ARG(distances || indices, "tm_cdt: at least one output must be requested");
  TM_CHECK_ARG(!indices || index_dtype == TM_I32 || index_dtype == TM_I64,
               "tm_cdt: index dtype must be int32 or int64");
  CdtOut o{distances, indices, index_dtype == TM_I64 ? 1 : 0, 0};
  hipStream_t st = (hipStream_t)stream;
  if (input_dtype == TM_U8)
    return cdt_dispatch((const uint8_t*)input, batch, ndim, shape, metric, o, workspace,
                        workspace_bytes, st);
  if (input_dtype == TM_F32)
    return cdt_dispatch((const float*)input, batch, ndim, shape, metric, o, workspace,
                        workspace_bytes, st);
  set_error("tm_cdt: input dtype must be uint8/bool or float32");
  return TM_ERR_ARG;
}
