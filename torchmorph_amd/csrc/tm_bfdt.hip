// Brute-force distance transform for gfx950: every foreground voxel takes the
// nearest background voxel of its item by scanning all of them (reference
// torchmorph/csrc/bfdt_kernel.cu, torchmorph.cpp:49-55; Python wrapper
// distance_transform.py:141-175).  O(voxels^2) per item: a checker for large
// fixtures and the reference's API, not a fast path.
//
// Semantics (the reference's, with its FP32 arithmetic replaced by FP64 so
// that values are those of scipy.ndimage.distance_transform_bf rounded to
// float32):
//  * metric 0 euclidean: sqrt(sum_d ((c_d - y_d) * s_d)^2), the terms summed
//    in axis order (the same operations as the EDT's distance from its
//    feature); 1 taxicab: sum_d |c_d - y_d| * s_d; 2 chessboard:
//    max_d |c_d - y_d| * s_d.  Sampling applies to every metric, as in the
//    reference (bfdt_kernel.cu:40-45).
//  * the nearest background is the first one in row-major order among those
//    at the smallest distance (bfdt_kernel.cu:92-96 keeps a strict minimum
//    over the background voxels in flat order);
//  * background voxels: distance 0, index = own coordinates;
//  * an item without background: euclidean +inf, taxicab / chessboard
//    4294967295 (bfdt_kernel.cu:208-218), indices 0.
//
// Layout: one lane per voxel of an item, grid (voxel tiles, items).  Each
// workgroup walks its item in tiles of 256 voxels: the tile's background flags
// and integer coordinates are decoded once into LDS, then every foreground
// lane reads them as broadcasts.  Workgroups without foreground skip the walk.
#include "tm_dt_lines.h"

namespace tmh {

constexpr int kBfThreads = 256;

template <int METRIC>
__device__ __forceinline__ double bf_term(int32_t delta, double s) {
  const double t = (double)delta * s;
  if (METRIC == 0) return t * t;
  return t < 0.0 ? -t : t;
}

template <typename TIn, int METRIC>
__global__ __launch_bounds__(kBfThreads) void bfdt_kernel(const TIn* __restrict__ in_all,
                                                          Shape sh, Samp sp,
                                                          float* __restrict__ dist,
                                                          void* __restrict__ idx, int idx64,
                                                          int64_t plane) {
  __shared__ int32_t tc[kMaxNdim][kBfThreads];
  __shared__ uint8_t tbg[kBfThreads];
  const int nd = sh.ndim;
  const int64_t n = sh.numel;
  const int64_t item = blockIdx.y;
  const TIn* in = in_all + item * n;
  const int64_t v = (int64_t)blockIdx.x * kBfThreads + threadIdx.x;
  const bool live = v < n;
  int32_t c[kMaxNdim];
  {
    int64_t r = live ? v : 0;
    for (int d = nd - 1; d >= 0; --d) {
      c[d] = (int32_t)(r % sh.size[d]);
      r /= sh.size[d];
    }
  }
  const bool fg = live && in[v] != TIn(0);
  double best = 0.0;
  int64_t bi = -1;
  if (__syncthreads_or(fg ? 1 : 0)) {
    for (int64_t t0 = 0; t0 < n; t0 += kBfThreads) {
      const int64_t u = t0 + threadIdx.x;
      if (u < n) {
        int64_t r = u;
        for (int d = nd - 1; d >= 0; --d) {
          tc[d][threadIdx.x] = (int32_t)(r % sh.size[d]);
          r /= sh.size[d];
        }
        tbg[threadIdx.x] = in[u] == TIn(0) ? 1 : 0;
      }
      __syncthreads();
      if (fg) {
        const int m = n - t0 < kBfThreads ? (int)(n - t0) : kBfThreads;
        for (int j = 0; j < m; ++j) {
          if (!tbg[j]) continue;
          double acc = bf_term<METRIC>(c[0] - tc[0][j], sp.s[0]);
          for (int d = 1; d < nd; ++d) {
            const double t = bf_term<METRIC>(c[d] - tc[d][j], sp.s[d]);
            if (METRIC == 2) acc = t > acc ? t : acc;
            else acc = acc + t;
          }
          if (bi < 0 || acc < best) {
            best = acc;
            bi = t0 + j;
          }
        }
      }
      __syncthreads();  // the tile is rewritten next
    }
  }
  if (!live) return;
  const int64_t g = item * n + v;
  int32_t f[kMaxNdim];
  float out;
  if (!fg) {  // background: itself
    out = 0.0f;
    for (int d = 0; d < nd; ++d) f[d] = c[d];
  } else if (bi < 0) {  // no background in the item
    out = METRIC == 0 ? __int_as_float(0x7f800000) : 4294967295.0f;
    for (int d = 0; d < nd; ++d) f[d] = 0;
  } else {
    out = METRIC == 0 ? (float)sqrt(best) : (float)best;
    int64_t r = bi;
    for (int d = nd - 1; d >= 0; --d) {
      f[d] = (int32_t)(r % sh.size[d]);
      r /= sh.size[d];
    }
  }
  if (dist) dist[g] = out;
  if (idx) {
    if (idx64) {
      for (int d = 0; d < nd; ++d) ((int64_t*)idx)[d * plane + g] = f[d];
    } else {
      for (int d = 0; d < nd; ++d) ((int32_t*)idx)[d * plane + g] = f[d];
    }
  }
}

template <typename TIn>
static int bfdt_launch(const TIn* in, int64_t batch, const Shape& sh, const Samp& sp, int metric,
                       float* dist, void* idx, int idx64, hipStream_t st) {
  const int64_t tiles = (sh.numel + kBfThreads - 1) / kBfThreads;
  if (tiles > 2147483647 || batch > 65535) {
    set_error("tm_bfdt: grid too large (%lld tiles x %lld items)", (long long)tiles,
              (long long)batch);
    return TM_ERR_ARG;
  }
  const dim3 g((unsigned)tiles, (unsigned)batch), blk(kBfThreads);
  const int64_t plane = batch * sh.numel;
  if (metric == 0)
    hipLaunchKernelGGL((bfdt_kernel<TIn, 0>), g, blk, 0, st, in, sh, sp, dist, idx, idx64, plane);
  else if (metric == 1)
    hipLaunchKernelGGL((bfdt_kernel<TIn, 1>), g, blk, 0, st, in, sh, sp, dist, idx, idx64, plane);
  else
    hipLaunchKernelGGL((bfdt_kernel<TIn, 2>), g, blk, 0, st, in, sh, sp, dist, idx, idx64, plane);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

}  // namespace tmh

extern "C" int tm_bfdt(const void* input, int input_dtype, int64_t batch, int ndim,
                       const int64_t* shape, const double* sampling, int metric,
                       float* distances, void* indices, int index_dtype, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(input != nullptr, "tm_bfdt: input is NULL");
  TM_CHECK_ARG(ndim >= 1 && ndim <= kMaxNdim, "tm_bfdt: ndim must be in 1..8, got %d", ndim);
  TM_CHECK_ARG(batch >= 1, "tm_bfdt: batch must be >= 1");
  TM_CHECK_ARG(metric >= 0 && metric <= 2,
               "tm_bfdt: metric must be 0 (euclidean), 1 (taxicab) or 2 (chessboard)");
  TM_CHECK_ARG(shape != nullptr, "tm_bfdt: shape is NULL");
  int64_t n = 1;
  for (int j = 0; j < ndim; ++j) {
    TM_CHECK_ARG(shape[j] >= 1 && shape[j] <= 2147483647, "tm_bfdt: axis %d size out of range",
                 j);
    n *= shape[j];
    TM_CHECK_ARG(n < (int64_t)1 << 31, "tm_bfdt: items above 2^31 voxels are not supported");
  }
  TM_CHECK_ARG(distances || indices, "tm_bfdt: at least one output must be requested");
  TM_CHECK_ARG(!indices || index_dtype == TM_I32 || index_dtype == TM_I64,
               "tm_bfdt: index dtype must be TM_I32 or TM_I64");
  Samp sp;
  for (int j = 0; j < kMaxNdim; ++j) sp.s[j] = 1.0;
  if (sampling)
    for (int j = 0; j < ndim; ++j) {
      TM_CHECK_ARG(sampling[j] > 0.0 && sampling[j] < 1e300,
                   "tm_bfdt: sampling must be finite and > 0");
      sp.s[j] = sampling[j];
    }
  const Shape sh = make_shape(ndim, shape);
  hipStream_t st = (hipStream_t)stream;
  const int idx64 = index_dtype == TM_I64 ? 1 : 0;
  if (input_dtype == TM_U8)
    return bfdt_launch((const uint8_t*)input, batch, sh, sp, metric, distances, indices, idx64,
                       st);
  if (input_dtype == TM_F32)
    return bfdt_launch((const float*)input, batch, sh, sp, metric, distances, indices, idx64, st);
  set_error("tm_bfdt: input dtype must be uint8/bool or float32");
  return TM_ERR_ARG;
}
