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
#include "tm_common.h"

namespace tmh {

constexpr int32_t kInf = 1 << 29;  // "unreached" inside the scan

struct CdtOut {
  float* dist;
  void* idx;
  int idx64;
  int64_t plane;
};

__device__ __forceinline__ int64_t phys(const Shape& sh, const int64_t* lc, bool rev) {
  int64_t p = 0;
  for (int j = 0; j < sh.ndim; ++j) {
    const int64_t c = rev ? sh.size[j] - 1 - lc[j] : lc[j];
    p += c * sh.stride[j];
  }
  return p;
}

// Processes one row (leading logical coordinates lc[0..ndim-2]) with one wave.
template <typename TIn>
__device__ void cdt_row(const TIn* __restrict__ in, int32_t* __restrict__ dt,
                        int32_t* __restrict__ ft, const Shape& sh, int64_t item, int64_t* lc,
                        bool rev, bool pass2, int metric, CdtOut o) {
  const int lane = threadIdx.x & 63;
  const int nd = sh.ndim;
  const int64_t W = sh.size[nd - 1];
  const int64_t ibase = item * sh.numel;
  // number of structure elements strictly before the centre, minus the row
  // predecessor (the last of them)
  int half = 1;
  for (int j = 0; j < nd; ++j) half *= 3;
  half = half / 2 - 1;

  int32_t carry_w = kInf;
  int32_t carry_s = -1;
  for (int64_t x0 = 0; x0 < W; x0 += 64) {
    const int64_t x = x0 + lane;
    const bool live = x < W;
    int32_t a = -1, src = -1;
    int64_t p = 0;
    if (live) {
      lc[nd - 1] = x;
      p = phys(sh, lc, rev);
      int32_t own;
      if (pass2) {
        own = dt[ibase + p];
        src = ft[ibase + p];
      } else {
        own = is_fg(in, ibase + p) ? -1 : 0;
        src = (int32_t)p;
      }
      a = own;
      if (own != 0) {
        for (int e = 0; e < half; ++e) {
          // decode structure element e (raster order over {-1,0,1}^nd)
          int rem = e, nz = 0;
          int d[kMaxNdim];
          for (int j = nd - 1; j >= 0; --j) {
            d[j] = rem % 3 - 1;
            rem /= 3;
            nz += d[j] != 0;
          }
          if (metric == 1 && nz > 1) continue;
          bool inb = true;
          int64_t q = 0;
          for (int j = 0; j < nd; ++j) {
            const int64_t c = lc[j] + d[j];
            if (c < 0 || c >= sh.size[j]) { inb = false; break; }
            q += (rev ? sh.size[j] - 1 - c : c) * sh.stride[j];
          }
          if (!inb) continue;
          const int32_t tt = dt[ibase + q];
          if (tt >= 0 && (a < 0 || tt + 1 < a)) {
            a = tt + 1;
            src = ft[ibase + q];
          }
        }
      }
    }
    // inclusive min-plus scan over the slot; the later element wins ties
    int32_t w = (live && a >= 0) ? a - (int32_t)x : kInf;
    int32_t s = src;
#pragma unroll
    for (int dlt = 1; dlt < 64; dlt <<= 1) {
      const int32_t ow = __shfl_up(w, dlt);
      const int32_t os = __shfl_up(s, dlt);
      if (lane >= dlt && ow < w) {
        w = ow;
        s = os;
      }
    }
    if (carry_w < w) {
      w = carry_w;
      s = carry_s;
    }
    carry_w = __shfl(w, 63);
    carry_s = __shfl(s, 63);
    if (live) {
      const int32_t v = (w >= kInf / 2) ? -1 : w + (int32_t)x;
      const int32_t f = (w >= kInf / 2) ? src : s;
      dt[ibase + p] = v;
      ft[ibase + p] = f;
      if (pass2) {
        const int64_t g = ibase + p;
        if (o.dist) o.dist[g] = (float)v;
        if (o.idx) {
          int64_t rem = f;
          for (int j = nd - 1; j >= 0; --j) {
            const int64_t c = rem % sh.size[j];
            rem /= sh.size[j];
            if (o.idx64) ((int64_t*)o.idx)[j * o.plane + g] = c;
            else ((int32_t*)o.idx)[j * o.plane + g] = (int32_t)c;
          }
        }
      }
    }
  }
}

// One workgroup per item.  mode 0: rows one per step (raster order);
// mode 1: 3-D wavefront t = 2z + y.
template <typename TIn>
__global__ __launch_bounds__(256) void cdt_pass_kernel(const TIn* __restrict__ in,
                                                       int32_t* __restrict__ dt,
                                                       int32_t* __restrict__ ft, Shape sh,
                                                       int mode, int metric, int pass2,
                                                       CdtOut o) {
  const int64_t item = blockIdx.x;
  const int wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  const int nd = sh.ndim;
  const bool rev = pass2 != 0;
  int64_t lc[kMaxNdim];
  if (mode == 1) {
    const int64_t D = sh.size[0], H = sh.size[1];
    const int64_t T = 2 * (D - 1) + (H - 1);
    for (int64_t t = 0; t <= T; ++t) {
      int64_t zlo = t - (H - 1) > 0 ? (t - (H - 1) + 1) / 2 : 0;
      int64_t zhi = t / 2 < D - 1 ? t / 2 : D - 1;
      for (int64_t z = zlo + wave; z <= zhi; z += nwaves) {
        lc[0] = z;
        lc[1] = t - 2 * z;
        cdt_row(in, dt, ft, sh, item, lc, rev, rev, metric, o);
      }
      __syncthreads();
    }
  } else {
    const int64_t W = sh.size[nd - 1];
    const int64_t nrows = sh.numel / W;
    for (int64_t r = 0; r < nrows; ++r) {
      if (wave == 0) {
        int64_t rem = r;
        for (int j = nd - 2; j >= 0; --j) {
          lc[j] = rem % sh.size[j];
          rem /= sh.size[j];
        }
        cdt_row(in, dt, ft, sh, item, lc, rev, rev, metric, o);
      }
      __syncthreads();
    }
  }
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
