// Per-line device code of the separable distance transforms, shared by the
// dense batched kernels (tm_edt.hip, tm_cdt.hip) and the ragged work-list
// kernel (tm_ragged.hip).
//
// A "line" is the set of voxels that differ only in coordinate d; one lane
// owns one line and walks it serially.  Lines are numbered so that adjacent
// lanes own lines adjacent along the innermost free axis, which makes the
// per-step loads coalesced for every axis but the last.
//
// EDT: SciPy's Maurer-Qi-Raghavan envelope in double precision (semantics in
// tm_edt.hip).
//
// CDT without indices: SciPy's two raster passes over the causal half of the
// 3^n structure produce exactly the taxicab (cross structure) or chessboard
// (full structure) distance to the nearest background voxel: every shortest
// path can be ordered so that its raster-forward steps come first and its
// raster-backward steps last, and it stays inside the item's bounding box.
// Both metrics are separable, so the values come from one pass per axis:
//   axis 0:  1-D distance to the nearest background voxel on the line;
//   taxicab: g'(i) = min_k |i-k| + g(k)       (two sweeps);
//   chessboard: g'(i) = min_k max(|i-k|, g(k)) (lower envelope of the
//            functions max(|x-k|, g(k)), Meijster, Roerdink and Hesselink,
//            "A general algorithm for computing distance transforms in linear
//            time", 2000, with its chessboard separator).
// All quantities are exact integers; kCdtInf marks "no background in this
// slice" and becomes -1 in the float output, SciPy's value when the item has
// no background at all.
#pragma once
#include "tm_common.h"

namespace tmh {

struct EdtOut {
  float* dist;
  void* idx;
  int idx64;
  int64_t plane;  // elements per index plane
};

struct Samp {
  double s[kMaxNdim];
};

__device__ __forceinline__ void edt_store(const EdtOut& o, int ndim, int64_t g, const int64_t* c,
                                          const int32_t* f, const Samp& sp) {
  if (o.dist) {
    double acc = 0.0;
    for (int j = 0; j < ndim; ++j) {
      int32_t delta = f[j] - (int32_t)c[j];
      double t = (double)delta * sp.s[j];
      t = t * t;
      acc = (j == 0) ? t : acc + t;
    }
    o.dist[g] = (float)sqrt(acc);
  }
  if (o.idx) {
    if (o.idx64) {
      int64_t* p = (int64_t*)o.idx;
      for (int j = 0; j < ndim; ++j) p[j * o.plane + g] = f[j];
    } else {
      int32_t* p = (int32_t*)o.idx;
      for (int j = 0; j < ndim; ++j) p[j * o.plane + g] = f[j];
    }
  }
}

// Line t of axis d: t enumerates (batch, c_0..c_{d-1}, c_{d+1}..c_{n-1}) with
// the inner axes fastest.  Returns the coordinates (c[d] = 0), the element
// offset of the line's first voxel and the batch index.
__device__ __forceinline__ void decode_line(const Shape& sh, int d, int64_t t, int64_t* c,
                                            int64_t* base, int64_t* b) {
  const int64_t inner = sh.stride[d];
  int64_t ob = t / inner;
  int64_t in = t - ob * inner;
  int64_t rem = in;
  for (int j = sh.ndim - 1; j > d; --j) {
    c[j] = rem % sh.size[j];
    rem /= sh.size[j];
  }
  rem = ob;
  for (int j = d - 1; j >= 0; --j) {
    c[j] = rem % sh.size[j];
    rem /= sh.size[j];
  }
  *b = rem;
  c[d] = 0;
  *base = ob * sh.size[d] * inner + in;
}

// ------------------------------------------------------------------ EDT --

// Axis 0 from the mask: nearest background along the line, lower index on
// ties (with every partial distance zero each site stays on the envelope, so
// SciPy's query reduces to this).  Writes the feature row into f0, or the
// final outputs when FINAL (ndim == 1).
template <typename TIn, bool FINAL>
__device__ __forceinline__ void edt_line_axis0(const TIn* __restrict__ in, int32_t* __restrict__ f0,
                                               const Shape& sh, int64_t t, const Samp& sp,
                                               const EdtOut& o) {
  int64_t c[kMaxNdim], base, b;
  decode_line(sh, 0, t, c, &base, &b);
  const int64_t step = sh.stride[0];
  const int64_t L = sh.size[0];
  int32_t last = -1;
  for (int64_t i = 0; i < L; ++i) {
    const int64_t v = base + i * step;
    if (!is_fg(in, v)) last = (int32_t)i;
    f0[v] = last;
  }
  int32_t nx = -1;
  for (int64_t i = L - 1; i >= 0; --i) {
    const int64_t v = base + i * step;
    if (!is_fg(in, v)) nx = (int32_t)i;
    const int32_t pv = f0[v];
    int32_t r;
    if (pv < 0) r = nx;
    else if (nx < 0) r = pv;
    else r = ((int32_t)i - pv <= nx - (int32_t)i) ? pv : nx;
    if (FINAL) {
      c[0] = i;
      int32_t f[kMaxNdim];
      f[0] = r;
      for (int j = 1; j < sh.ndim; ++j) f[j] = r >= 0 ? (int32_t)c[j] : 0;
      edt_store(o, sh.ndim, v, c, f, sp);
    } else {
      f0[v] = r;
    }
  }
}

// Axis d >= 1: envelope of the sites on line t.  fin holds the d feature
// planes of the previous pass (plane stride `plane`), fout receives d+1
// planes, or the final outputs when FINAL.  Stack entry k of this line lives
// at pos[k * ss] / hh[k * ss] (pointers already offset to the line's slot 0;
// LDS or global memory).
template <bool FINAL, typename PT>
__device__ __forceinline__ void edt_line_stack(const int32_t* __restrict__ fin,
                                               int32_t* __restrict__ fout, int64_t plane,
                                               const Shape& sh, int d, int64_t t, const Samp& sp,
                                               PT* pos, double* hh, int64_t ss, const EdtOut& o) {
  int64_t c[kMaxNdim], base, b;
  decode_line(sh, d, t, c, &base, &b);
  const int64_t step = sh.stride[d];
  const int64_t L = sh.size[d];
  const double sd = sp.s[d];

  int64_t top = -1;
  int32_t nf0 = fin[base];
  for (int64_t i = 0; i < L; ++i) {
    const int64_t v = base + i * step;
    const int32_t f0 = nf0;
    if (i + 1 < L) nf0 = fin[v + step];  // independent load for the next step
    if (f0 < 0) continue;
    double wR = 0.0;
    {
      double tt = (double)(f0 - (int32_t)c[0]) * sp.s[0];
      wR += tt * tt;
      for (int j = 1; j < d; ++j) {
        double u = (double)(fin[j * plane + v] - (int32_t)c[j]) * sp.s[j];
        wR += u * u;
      }
    }
    while (top >= 1) {
      const int32_t pv = (int32_t)pos[top * ss], pu = (int32_t)pos[(top - 1) * ss];
      const double a = (double)(pv - pu) * sd;
      const double bb = (double)((int32_t)i - pv) * sd;
      const double cc = a + bb;
      const double uR = hh[(top - 1) * ss], vR = hh[top * ss];
      if (cc * vR - bb * uR - a * wR - a * bb * cc <= 0.0) break;
      --top;
    }
    ++top;
    pos[top * ss] = (PT)i;
    hh[top * ss] = wR;
  }

  int32_t f[kMaxNdim];
  if (top < 0) {
    // no site on this line: every voxel keeps the feature (-1, 0, ..., 0)
    for (int j = 0; j < sh.ndim; ++j) f[j] = 0;
    f[0] = -1;
    for (int64_t i = 0; i < L; ++i) {
      const int64_t v = base + i * step;
      if (FINAL) {
        c[d] = i;
        edt_store(o, sh.ndim, v, c, f, sp);
      } else {
        for (int j = 0; j <= d; ++j) fout[j * plane + v] = f[j];
      }
    }
    return;
  }
  int64_t l = 0;
  int32_t pl = (int32_t)pos[0];
  double hl = hh[0];
  for (int64_t i = 0; i < L; ++i) {
    double tt = (double)(pl - (int32_t)i) * sd;
    double d1 = hl + tt * tt;
    while (l < top) {
      const int32_t pn = (int32_t)pos[(l + 1) * ss];
      const double hn = hh[(l + 1) * ss];
      double t2 = (double)(pn - (int32_t)i) * sd;
      double d2 = hn + t2 * t2;
      if (d1 <= d2) break;
      d1 = d2;
      ++l;
      pl = pn;
      hl = hn;
    }
    const int64_t site = base + (int64_t)pl * step;
    for (int j = 0; j < d; ++j) f[j] = fin[j * plane + site];
    f[d] = pl;
    const int64_t v = base + i * step;
    if (FINAL) {
      for (int j = d + 1; j < sh.ndim; ++j) f[j] = (int32_t)c[j];
      c[d] = i;
      edt_store(o, sh.ndim, v, c, f, sp);
    } else {
      for (int j = 0; j <= d; ++j) fout[j * plane + v] = f[j];
    }
  }
}

// ------------------------------------------------- CDT (distances only) --

constexpr int32_t kCdtInf = 1 << 28;

// Final value as float bits: -1 when no background was reachable.
__device__ __forceinline__ int32_t cdt_final_bits(int32_t a) {
  return __float_as_int(a >= kCdtInf ? -1.0f : (float)a);
}

// Axis 0 from the mask.  w is int32 storage of the item (the float output
// itself is used, reinterpreted, so no workspace is needed); FINAL writes the
// float bits.
template <typename TIn, bool FINAL>
__device__ __forceinline__ void cdt_line_axis0(const TIn* __restrict__ in, int32_t* __restrict__ w,
                                               int64_t base, int64_t step, int64_t L) {
  int64_t last = -1;
  for (int64_t i = 0; i < L; ++i) {
    const int64_t v = base + i * step;
    if (!is_fg(in, v)) last = i;
    w[v] = last >= 0 ? (int32_t)(i - last) : kCdtInf;
  }
  int64_t nx = -1;
  for (int64_t i = L - 1; i >= 0; --i) {
    const int64_t v = base + i * step;
    int32_t a = w[v];
    if (a == 0) nx = i;
    else if (nx >= 0 && (int32_t)(nx - i) < a) a = (int32_t)(nx - i);
    w[v] = FINAL ? cdt_final_bits(a) : a;
  }
}

// Taxicab, axis d >= 1, in place: g'(i) = min_k |i-k| + g(k) by a forward and
// a backward sweep.
template <bool FINAL>
__device__ __forceinline__ void cdt_line_taxicab(int32_t* __restrict__ w, int64_t base,
                                                 int64_t step, int64_t L) {
  int32_t prev = kCdtInf;
  for (int64_t i = 0; i < L; ++i) {
    const int64_t v = base + i * step;
    int32_t a = w[v];
    if (prev + 1 < a) a = prev + 1;
    w[v] = a;
    prev = a;
  }
  int32_t nxt = kCdtInf;
  for (int64_t i = L - 1; i >= 0; --i) {
    const int64_t v = base + i * step;
    int32_t a = w[v];
    if (nxt + 1 < a) a = nxt + 1;
    nxt = a;
    w[v] = FINAL ? cdt_final_bits(a) : a;
  }
}

__device__ __forceinline__ int32_t cheb(int32_t x, int32_t s, int32_t g) {
  const int32_t dx = x > s ? x - s : s - x;
  return dx > g ? dx : g;
}

// Chessboard, axis d >= 1, in place: lower envelope of max(|x-k|, g(k)) over
// the sites k with finite g.  Stack entry k: site S[k*ss], region start
// T[k*ss] (PT holds any coordinate of the line), value G[k*ss].  A site is popped when the new one is at least as
// good at the start of its region; for these functions the new site is then
// at least as good on the whole rest of the line, so only tied minima are
// dropped and every value is exact.
template <bool FINAL, typename PT>
__device__ __forceinline__ void cdt_line_chessboard(int32_t* __restrict__ w, int64_t base,
                                                    int64_t step, int64_t L, PT* S, PT* T,
                                                    int32_t* G, int64_t ss) {
  int64_t q = -1;
  int32_t sq = 0, tq = 0, gq = 0;  // top of the stack, kept in registers
  for (int64_t u = 0; u < L; ++u) {
    const int32_t g = w[base + u * step];
    if (g >= kCdtInf) continue;
    const int32_t ui = (int32_t)u;
    while (q >= 0 && cheb(tq, sq, gq) >= cheb(tq, ui, g)) {
      --q;
      if (q >= 0) {
        sq = S[q * ss];
        tq = T[q * ss];
        gq = G[q * ss];
      }
    }
    if (q < 0) {
      q = 0;
      sq = ui;
      tq = 0;
      gq = g;
      S[0] = (PT)sq;
      T[0] = (PT)tq;
      G[0] = gq;
    } else {
      const int32_t mid = (sq + ui) / 2;
      int32_t sep;
      if (gq <= g) sep = sq + g > mid ? sq + g : mid;
      else sep = ui - gq < mid ? ui - gq : mid;
      const int64_t start = (int64_t)sep + 1;
      if (start < L) {
        ++q;
        sq = ui;
        tq = (int32_t)start;
        gq = g;
        S[q * ss] = (PT)sq;
        T[q * ss] = (PT)tq;
        G[q * ss] = gq;
      }
    }
  }
  for (int64_t u = L - 1; u >= 0; --u) {
    const int32_t a = q < 0 ? kCdtInf : cheb((int32_t)u, sq, gq);
    w[base + u * step] = FINAL ? cdt_final_bits(a) : a;
    if (q >= 0 && u == tq) {
      --q;
      if (q >= 0) {
        sq = S[q * ss];
        tq = T[q * ss];
        gq = G[q * ss];
      }
    }
  }
}

}  // namespace tmh
