// Passes along the contiguous last axis, one workgroup of kCLanes lanes per
// kCLanes consecutive lines (rows) of one item, lane = line.
//
// One lane per row makes every step of every lane touch its own cache line;
// with a full GPU of lanes in flight those lines do not survive in the caches
// between steps, so each 4-byte access costs a line from HBM.  Here the rows
// are walked in chunks of CH voxels: a chunk of all the workgroup's rows is
// loaded with consecutive lanes on consecutive addresses (whole lines), staged
// in LDS ([row][CH + 1], conflict-free for both access patterns), consumed
// serially per lane, and outputs leave through LDS the same way.  The next
// chunk's loads are issued before the current chunk is consumed.
//
// The per-voxel arithmetic is that of tm_dt_fast.h / tm_dt_lines.h, in the
// same order, so results are bit-identical.  Every barrier sits at the top
// level of loops whose trip counts depend only on the row length, which is
// the same for every lane of the workgroup (lanes past the item's last row
// take part in loads, stores and barriers and compute nothing).
#pragma once
#include "tm_dt_fast.h"

namespace tmh {

constexpr int kCLanes = 256;

// Thread tid's share of one chunk: element q of its registers is row
// (q * kCLanes + tid) / CH, column (q * kCLanes + tid) % CH of the chunk.
template <int CH>
__device__ __forceinline__ void chunk_fetch(int32_t (&r)[CH], const int32_t* __restrict__ src,
                                            int64_t row0, int L, int nrows, int i0) {
#pragma unroll
  for (int q = 0; q < CH; ++q) {
    const int e = q * kCLanes + (int)threadIdx.x;
    const int row = e / CH, k = e % CH;
    r[q] = (row < nrows && i0 + k < L) ? src[(row0 + row) * L + i0 + k] : 0;
  }
}

template <int CH>
__device__ __forceinline__ void chunk_put(const int32_t (&r)[CH], int32_t* lds) {
#pragma unroll
  for (int q = 0; q < CH; ++q) {
    const int e = q * kCLanes + (int)threadIdx.x;
    lds[(e / CH) * (CH + 1) + e % CH] = r[q];
  }
}

template <int CH>
__device__ __forceinline__ void chunk_flush(const int32_t* lds, int32_t* __restrict__ dst,
                                            int64_t row0, int L, int nrows, int i0) {
#pragma unroll
  for (int q = 0; q < CH; ++q) {
    const int e = q * kCLanes + (int)threadIdx.x;
    const int row = e / CH, k = e % CH;
    if (row < nrows && i0 + k < L) dst[(row0 + row) * L + i0 + k] = lds[row * (CH + 1) + k];
  }
}

// The float distance edt_store writes for features f at coordinates c.
__device__ __forceinline__ float edt_dist_of(int ndim, const int64_t* c, const int32_t* f,
                                             const Samp& sp) {
  double acc = 0.0;
  for (int j = 0; j < ndim; ++j) {
    int32_t delta = f[j] - (int32_t)c[j];
    double t = (double)delta * sp.s[j];
    t = t * t;
    acc = (j == 0) ? t : acc + t;
  }
  return (float)sqrt(acc);
}

// EDT, last pass (d = ndim - 1, NP = d feature planes, 1 or 2), distances
// only.  fin: the item's feature planes (plane stride `plane`); rows
// row0 .. row0 + nrows - 1; the compact stack of edt_line_stack_c for line t
// at E/F1 + t, stride nl_axis.  lds: (NP + 1) * kCLanes * (CH + 1) ints.
template <int NP, int CH>
__device__ void edt_rows_final(const int32_t* __restrict__ fin, int64_t plane, const Shape& sh,
                               int64_t row0, int nrows, const Samp& sp, uint32_t* __restrict__ Eall,
                               uint16_t* __restrict__ Fall, int64_t nl_axis,
                               float* __restrict__ dist, int32_t* lds) {
  constexpr int RS = CH + 1;
  const int d = sh.ndim - 1;
  const int L = (int)sh.size[d];
  const int lane = threadIdx.x;
  const bool active = lane < nrows;
  const int64_t t = row0 + lane;
  int64_t c[kMaxNdim], base = 0, b = 0;
  for (int j = 0; j < kMaxNdim; ++j) c[j] = 0;
  if (active) decode_line(sh, d, t, c, &base, &b);
  const double sd = sp.s[d];
  const double s0 = sp.s[0], s1 = sp.s[1];
  const int32_t c0 = (int32_t)c[0], c1 = (int32_t)c[1];
  int32_t* in0 = lds;
  int32_t* in1 = lds + kCLanes * RS;
  int32_t* ob = lds + NP * kCLanes * RS;
  uint32_t* E = Eall + t;
  uint16_t* F1 = Fall + t;
  const int64_t ss = nl_axis;
  const int nch = (L + CH - 1) / CH;

  // build: q1 = entry top, q0 = entry top - 1; memory holds 0 .. top - 2
  int top = -1;
  EdtSite q1{0, 0, 0, 0.0}, q0{0, 0, 0, 0.0};
  int32_t r0[CH], r1[CH];
  chunk_fetch<CH>(r0, fin, row0, L, nrows, 0);
  if (NP == 2) chunk_fetch<CH>(r1, fin + plane, row0, L, nrows, 0);
  for (int q = 0; q < nch; ++q) {
    __syncthreads();  // the previous chunk has been consumed
    chunk_put<CH>(r0, in0);
    if (NP == 2) chunk_put<CH>(r1, in1);
    __syncthreads();
    if (q + 1 < nch) {
      chunk_fetch<CH>(r0, fin, row0, L, nrows, (q + 1) * CH);
      if (NP == 2) chunk_fetch<CH>(r1, fin + plane, row0, L, nrows, (q + 1) * CH);
    }
    if (active) {
      for (int k = 0; k < CH; ++k) {
        const int i = q * CH + k;
        if (i >= L) break;
        const int32_t f0 = in0[lane * RS + k];
        if (f0 < 0) continue;
        const int32_t f1 = NP == 2 ? in1[lane * RS + k] : 0;
        const double wR = edt_partial<NP>(f0, f1, c0, c1, s0, s1);
        while (top >= 1) {
          const double a = (double)(q1.p - q0.p) * sd;
          const double bb = (double)(i - q1.p) * sd;
          const double cc = a + bb;
          if (cc * q1.h - bb * q0.h - a * wR - a * bb * cc <= 0.0) break;
          --top;
          q1 = q0;
          if (top >= 1) q0 = edt_site_get<NP>(E, F1, top - 1, ss, c0, c1, s0, s1);
        }
        if (top >= 1) edt_site_put<NP>(E, F1, top - 1, ss, q0);
        ++top;
        q0 = q1;
        q1 = EdtSite{i, f0, f1, wR};
      }
    }
  }
  if (active && top >= 1) edt_site_put<NP>(E, F1, top - 1, ss, q0);
  if (active && top >= 0) edt_site_put<NP>(E, F1, top, ss, q1);

  // query: cur = site l, nx = site l + 1, nn = raw entry l + 2 (loaded ahead)
  int l = 0;
  EdtSite cur{0, 0, 0, 0.0}, nx{0, 0, 0, 0.0};
  uint32_t nnE = 0;
  uint16_t nnF = 0;
  if (active && top >= 0) {
    cur = edt_site_get<NP>(E, F1, 0, ss, c0, c1, s0, s1);
    if (top >= 1) nx = edt_site_get<NP>(E, F1, 1, ss, c0, c1, s0, s1);
    if (top >= 2) {
      nnE = E[2 * ss];
      if (NP == 2) nnF = F1[2 * ss];
    }
  }
  int32_t fnone[kMaxNdim];  // no site on the line: feature (-1, 0, ..., 0)
  for (int j = 0; j < kMaxNdim; ++j) fnone[j] = 0;
  fnone[0] = -1;
  for (int q = 0; q < nch; ++q) {
    if (active) {
      for (int k = 0; k < CH; ++k) {
        const int i = q * CH + k;
        if (i >= L) break;
        float out;
        if (top < 0) {
          c[d] = i;
          out = edt_dist_of(sh.ndim, c, fnone, sp);
        } else {
          double tt = (double)(cur.p - i) * sd;
          double d1 = cur.h + tt * tt;
          while (l < top) {
            double t2 = (double)(nx.p - i) * sd;
            double d2 = nx.h + t2 * t2;
            if (d1 <= d2) break;
            d1 = d2;
            ++l;
            cur = nx;
            if (l < top) {
              nx.p = (int32_t)(nnE & 0xffffu);
              nx.f0 = (int32_t)(nnE >> 16);
              nx.f1 = (int32_t)nnF;
              nx.h = edt_partial<NP>(nx.f0, nx.f1, c0, c1, s0, s1);
              if (l + 2 <= top) {
                nnE = E[(int64_t)(l + 2) * ss];
                if (NP == 2) nnF = F1[(int64_t)(l + 2) * ss];
              }
            }
          }
          out = (float)sqrt(d1);
        }
        ob[lane * RS + k] = __float_as_int(out);
      }
    }
    __syncthreads();
    chunk_flush<CH>(ob, (int32_t*)dist, row0, L, nrows, q * CH);
    __syncthreads();  // ob is rewritten by the next chunk
  }
}

// Taxicab chamfer, last pass, in place on w (int32 in, float bits out):
// forward sweep g'(i) = min(g(i), g'(i-1) + 1), then backward.
template <int CH>
__device__ void cdt_rows_taxicab(int32_t* __restrict__ w, const Shape& sh, int64_t row0,
                                 int nrows, int32_t* lds) {
  constexpr int RS = CH + 1;
  const int L = (int)sh.size[sh.ndim - 1];
  const int lane = threadIdx.x;
  const bool active = lane < nrows;
  const int nch = (L + CH - 1) / CH;
  int32_t r[CH];
  int32_t prev = kCdtInf;
  chunk_fetch<CH>(r, w, row0, L, nrows, 0);
  for (int q = 0; q < nch; ++q) {
    __syncthreads();
    chunk_put<CH>(r, lds);
    __syncthreads();
    if (q + 1 < nch) chunk_fetch<CH>(r, w, row0, L, nrows, (q + 1) * CH);
    if (active) {
      for (int k = 0; k < CH; ++k) {
        if (q * CH + k >= L) break;
        int32_t a = lds[lane * RS + k];
        if (prev + 1 < a) a = prev + 1;
        lds[lane * RS + k] = a;
        prev = a;
      }
    }
    __syncthreads();
    chunk_flush<CH>(lds, w, row0, L, nrows, q * CH);
  }
  // the forward values written above are read back by other lanes below
  __syncthreads();
  int32_t nxt = kCdtInf;
  chunk_fetch<CH>(r, w, row0, L, nrows, (nch - 1) * CH);
  for (int q = nch - 1; q >= 0; --q) {
    __syncthreads();
    chunk_put<CH>(r, lds);
    __syncthreads();
    if (q > 0) chunk_fetch<CH>(r, w, row0, L, nrows, (q - 1) * CH);
    if (active) {
      for (int k = CH - 1; k >= 0; --k) {
        if (q * CH + k >= L) continue;
        int32_t a = lds[lane * RS + k];
        if (nxt + 1 < a) a = nxt + 1;
        nxt = a;
        lds[lane * RS + k] = cdt_final_bits(a);
      }
    }
    __syncthreads();
    chunk_flush<CH>(lds, w, row0, L, nrows, q * CH);
  }
}

// Chessboard pass along an axis, one voxel per lane, exact: the lower
// envelope of max(|i - k|, g(k)) at i is the smallest r >= 0 for which some
// k with |i - k| <= r has g(k) <= r (the window minimum m(r) does not grow
// with r, so max(r, m(r)) is smallest where m(r) first drops to r or below),
// or the minimum of the whole line when no such r reaches the line's ends
// (kCdtInf: no background on the line).  Equal to cdt_line_chessboard's
// values; the cost per voxel is its distance.
__device__ __forceinline__ int32_t cheb_window(const int32_t* __restrict__ g, int64_t base,
                                               int64_t step, int L, int i) {
  int32_t m = g[base + (int64_t)i * step];
  if (m <= 0) return m;
  const int rfull = i > L - 1 - i ? i : L - 1 - i;
  for (int r = 1; r <= rfull; ++r) {
    if (i - r >= 0) {
      const int32_t v = g[base + (int64_t)(i - r) * step];
      m = v < m ? v : m;
    }
    if (i + r < L) {
      const int32_t v = g[base + (int64_t)(i + r) * step];
      m = v < m ? v : m;
    }
    if (m <= r) return r;
  }
  return m;
}

}  // namespace tmh
