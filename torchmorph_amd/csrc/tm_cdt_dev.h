// Device code of the chamfer DT shared by tm_cdt.hip (dense batches) and
// tm_ragged.hip (ragged work-lists).  Semantics and schedule are described in
// tm_cdt.hip.
#pragma once
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

// One item's pass by one workgroup (pointers already offset to the item).
// mode 0: rows one per step (raster order); mode 1: 3-D wavefront t = 2z + y.
template <typename TIn>
__device__ void cdt_item_pass(const TIn* __restrict__ in, int32_t* __restrict__ dt,
                              int32_t* __restrict__ ft, const Shape& sh, int mode, int metric,
                              int pass2, CdtOut o) {
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
        cdt_row(in, dt, ft, sh, 0, lc, rev, rev, metric, o);
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
        cdt_row(in, dt, ft, sh, 0, lc, rev, rev, metric, o);
      }
      __syncthreads();
    }
  }
}

}  // namespace tmh
