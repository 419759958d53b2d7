// EDT envelope pass (axis d >= 1) for the dense N-D kernels, with less
// memory latency in each line's serial chain than edt_line_stack
// (tm_dt_lines.h), and the same double operations in the same order, so the
// same bits:
//  * the feature planes of the next PF voxels are loaded before they
//    are used (independent loads in flight per lane);
//  * the two stack entries that the next removal test reads (build) and the
//    current and next envelope sites (query) are held in registers, so the
//    stack memory is touched once per push and once per pop or advance;
//  * the features of the current site are loaded once per site, not once
//    per voxel;
//  * the final pass without indices stores sqrt of the query's own squared
//    sum: the site's partial sum of the axes below d plus the square of its
//    offset along d is, term for term and in order, the sum that edt_store
//    forms from the site's features.
#pragma once
#include "tm_dt_lines.h"

namespace tmh {

// Features of planes 0 and 1 of the next PF voxels of a line (planes above 1
// are read where they are used: items of four or more axes).
template <int PF>
struct LineChunk {
  int32_t f0[PF];
  int32_t f1[PF];
};

template <int PF>
__device__ __forceinline__ void line_chunk_load(LineChunk<PF>& ch,
                                                const int32_t* __restrict__ fin, int64_t plane,
                                                int64_t base, int64_t step, int i0, int L,
                                                int d) {
#pragma unroll
  for (int k = 0; k < PF; ++k)
    ch.f0[k] = (i0 + k < L) ? fin[base + (int64_t)(i0 + k) * step] : -1;
  if (d >= 2) {
#pragma unroll
    for (int k = 0; k < PF; ++k)
      ch.f1[k] = (i0 + k < L) ? fin[plane + base + (int64_t)(i0 + k) * step] : 0;
  }
}

template <bool FINAL, int PF, typename PT>
__device__ __forceinline__ void edt_line_stack_fast(const int32_t* __restrict__ fin,
                                                    int32_t* __restrict__ fout, int64_t plane,
                                                    const Shape& sh, int d, int64_t t,
                                                    const Samp& sp, PT* pos, double* hh,
                                                    int64_t ss, const EdtOut& o) {
  int64_t c[kMaxNdim], base, b;
  decode_line(sh, d, t, c, &base, &b);
  const int64_t step = sh.stride[d];
  const int L = (int)sh.size[d];
  const double sd = sp.s[d];
  const double s0 = sp.s[0], s1 = sp.s[1];
  const int32_t c0 = (int32_t)c[0], c1 = (int32_t)c[1];

  // build: (p1, h1) = top entry, (p0, h0) = the entry below it
  int top = -1;
  int32_t p1 = 0, p0 = 0;
  double h1 = 0.0, h0 = 0.0;
  LineChunk<PF> ch;
  line_chunk_load(ch, fin, plane, base, step, 0, L, d);
  for (int i0 = 0; i0 < L; i0 += PF) {
    LineChunk<PF> cur = ch;
    if (i0 + PF < L) line_chunk_load(ch, fin, plane, base, step, i0 + PF, L, d);
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int32_t f0 = cur.f0[k];
      if (f0 < 0) continue;  // also every voxel past the line's end
      const int i = i0 + k;
      double wR = 0.0;
      {
        double tt = (double)(f0 - c0) * s0;
        wR += tt * tt;
        if (d >= 2) {
          double u = (double)(cur.f1[k] - c1) * s1;
          wR += u * u;
          for (int j = 2; j < d; ++j) {
            const int32_t fj = fin[j * plane + base + (int64_t)i * step];
            double w = (double)(fj - (int32_t)c[j]) * sp.s[j];
            wR += w * w;
          }
        }
      }
      while (top >= 1) {
        const double a = (double)(p1 - p0) * sd;
        const double bb = (double)(i - p1) * sd;
        const double cc = a + bb;
        if (cc * h1 - bb * h0 - a * wR - a * bb * cc <= 0.0) break;
        --top;
        p1 = p0;
        h1 = h0;
        if (top >= 1) {
          p0 = (int32_t)pos[(top - 1) * ss];
          h0 = hh[(top - 1) * ss];
        }
      }
      ++top;
      p0 = p1;
      h0 = h1;
      p1 = i;
      h1 = wR;
      pos[top * ss] = (PT)i;
      hh[top * ss] = wR;
    }
  }

  int32_t f[kMaxNdim];
  if (top < 0) {
    // no site on this line: every voxel keeps the feature (-1, 0, ..., 0)
    for (int j = 0; j < sh.ndim; ++j) f[j] = 0;
    f[0] = -1;
    for (int i = 0; i < L; ++i) {
      const int64_t v = base + (int64_t)i * step;
      if (FINAL) {
        c[d] = i;
        edt_store(o, sh.ndim, v, c, f, sp);
      } else {
        for (int j = 0; j <= d; ++j) fout[j * plane + v] = f[j];
      }
    }
    return;
  }

  // query: (pl, hl) = current site, (pn, hn) = the next one (l < top)
  const bool dist_only = FINAL && !o.idx;
  int l = 0;
  int32_t pl = (int32_t)pos[0];
  double hl = hh[0];
  int32_t pn = 0;
  double hn = 0.0;
  if (top >= 1) {
    pn = (int32_t)pos[ss];
    hn = hh[ss];
  }
  int32_t site_loaded = -1;  // features f[0..d-1] belong to this site
  for (int i = 0; i < L; ++i) {
    double tt = (double)(pl - i) * sd;
    double d1 = hl + tt * tt;
    while (l < top) {
      double t2 = (double)(pn - i) * sd;
      double d2 = hn + t2 * t2;
      if (d1 <= d2) break;
      d1 = d2;
      ++l;
      pl = pn;
      hl = hn;
      if (l < top) {
        pn = (int32_t)pos[(l + 1) * ss];
        hn = hh[(l + 1) * ss];
      }
    }
    const int64_t v = base + (int64_t)i * step;
    if (dist_only) {
      o.dist[v] = (float)sqrt(d1);
      continue;
    }
    if (pl != site_loaded) {
      const int64_t site = base + (int64_t)pl * step;
      for (int j = 0; j < d; ++j) f[j] = fin[j * plane + site];
      site_loaded = pl;
    }
    f[d] = pl;
    if (FINAL) {
      for (int j = d + 1; j < sh.ndim; ++j) f[j] = (int32_t)c[j];
      c[d] = i;
      edt_store(o, sh.ndim, v, c, f, sp);
    } else {
      for (int j = 0; j <= d; ++j) fout[j * plane + v] = f[j];
    }
  }
}

// Compact-stack form for d = 1, 2 (items of up to 3 axes).  A stack entry is
// the site and its features on the earlier axes -- E[k * ss] = site |
// f0 << 16, F1[k * ss] = f1 (d = 2) -- 4 or 6 bytes instead of the site and
// its double partial sum: the partial sum is recomputed from f0 (and f1)
// with the same operations as at the push, so it has the same bits.  Only
// entries buried two deep are written: the top two live in registers, and a
// site pushed and popped again never touches memory.  The query reads the
// site's features from its entry (no gather of the feature planes).
struct EdtSite {
  int32_t p, f0, f1;
  double h;
};

template <int D>
__device__ __forceinline__ double edt_partial(int32_t f0, int32_t f1, int32_t c0, int32_t c1,
                                              double s0, double s1) {
  double wR = 0.0;
  double tt = (double)(f0 - c0) * s0;
  wR += tt * tt;
  if (D == 2) {
    double u = (double)(f1 - c1) * s1;
    wR += u * u;
  }
  return wR;
}

template <int D>
__device__ __forceinline__ void edt_site_put(uint32_t* E, uint16_t* F1, int64_t k, int64_t ss,
                                             const EdtSite& e) {
  E[k * ss] = (uint32_t)(uint16_t)e.p | ((uint32_t)(uint16_t)e.f0 << 16);
  if (D == 2) F1[k * ss] = (uint16_t)e.f1;
}

template <int D>
__device__ __forceinline__ EdtSite edt_site_get(const uint32_t* E, const uint16_t* F1, int64_t k,
                                                int64_t ss, int32_t c0, int32_t c1, double s0,
                                                double s1) {
  EdtSite e;
  const uint32_t w = E[k * ss];
  e.p = (int32_t)(w & 0xffffu);
  e.f0 = (int32_t)(w >> 16);
  e.f1 = D == 2 ? (int32_t)F1[k * ss] : 0;
  e.h = edt_partial<D>(e.f0, e.f1, c0, c1, s0, s1);
  return e;
}

template <bool FINAL, int D, int PF>
__device__ __forceinline__ void edt_line_stack_c(const int32_t* __restrict__ fin,
                                                 int32_t* __restrict__ fout, int64_t plane,
                                                 const Shape& sh, int64_t t, const Samp& sp,
                                                 uint32_t* E, uint16_t* F1, int64_t ss,
                                                 const EdtOut& o) {
  const int d = D;
  int64_t c[kMaxNdim], base, b;
  decode_line(sh, d, t, c, &base, &b);
  const int64_t step = sh.stride[d];
  const int L = (int)sh.size[d];
  const double sd = sp.s[d];
  const double s0 = sp.s[0], s1 = sp.s[1];
  const int32_t c0 = (int32_t)c[0], c1 = (int32_t)c[1];

  // build: q1 = entry top, q0 = entry top - 1 (registers); memory holds
  // entries 0 .. top - 2
  int top = -1;
  EdtSite q1{0, 0, 0, 0.0}, q0{0, 0, 0, 0.0};
  LineChunk<PF> ch;
  line_chunk_load(ch, fin, plane, base, step, 0, L, d);
  for (int i0 = 0; i0 < L; i0 += PF) {
    LineChunk<PF> cur = ch;
    if (i0 + PF < L) line_chunk_load(ch, fin, plane, base, step, i0 + PF, L, d);
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int32_t f0 = cur.f0[k];
      if (f0 < 0) continue;  // also every voxel past the line's end
      const int i = i0 + k;
      const int32_t f1 = D == 2 ? cur.f1[k] : 0;
      const double wR = edt_partial<D>(f0, f1, c0, c1, s0, s1);
      while (top >= 1) {
        const double a = (double)(q1.p - q0.p) * sd;
        const double bb = (double)(i - q1.p) * sd;
        const double cc = a + bb;
        if (cc * q1.h - bb * q0.h - a * wR - a * bb * cc <= 0.0) break;
        --top;
        q1 = q0;
        if (top >= 1) q0 = edt_site_get<D>(E, F1, top - 1, ss, c0, c1, s0, s1);
      }
      if (top >= 1) edt_site_put<D>(E, F1, top - 1, ss, q0);
      ++top;
      q0 = q1;
      q1 = EdtSite{i, f0, f1, wR};
    }
  }
  if (top >= 1) edt_site_put<D>(E, F1, top - 1, ss, q0);
  if (top >= 0) edt_site_put<D>(E, F1, top, ss, q1);

  int32_t f[kMaxNdim];
  if (top < 0) {
    // no site on this line: every voxel keeps the feature (-1, 0, ..., 0)
    for (int j = 0; j < sh.ndim; ++j) f[j] = 0;
    f[0] = -1;
    for (int i = 0; i < L; ++i) {
      const int64_t v = base + (int64_t)i * step;
      if (FINAL) {
        c[d] = i;
        edt_store(o, sh.ndim, v, c, f, sp);
      } else {
        for (int j = 0; j <= d; ++j) fout[j * plane + v] = f[j];
      }
    }
    return;
  }

  // query: cur = site l, nx = site l + 1, nn = raw entry l + 2 (loaded ahead)
  const bool dist_only = FINAL && !o.idx;
  int l = 0;
  EdtSite cur = edt_site_get<D>(E, F1, 0, ss, c0, c1, s0, s1);
  EdtSite nx = cur;
  if (top >= 1) nx = edt_site_get<D>(E, F1, 1, ss, c0, c1, s0, s1);
  uint32_t nnE = 0;
  uint16_t nnF = 0;
  if (top >= 2) {
    nnE = E[2 * ss];
    if (D == 2) nnF = F1[2 * ss];
  }
  for (int i = 0; i < L; ++i) {
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
        nx.h = edt_partial<D>(nx.f0, nx.f1, c0, c1, s0, s1);
        if (l + 2 <= top) {
          nnE = E[(int64_t)(l + 2) * ss];
          if (D == 2) nnF = F1[(int64_t)(l + 2) * ss];
        }
      }
    }
    const int64_t v = base + (int64_t)i * step;
    if (dist_only) {
      o.dist[v] = (float)sqrt(d1);
      continue;
    }
    f[0] = cur.f0;
    if (D == 2) f[1] = cur.f1;
    f[d] = cur.p;
    if (FINAL) {
      c[d] = i;
      edt_store(o, sh.ndim, v, c, f, sp);
    } else {
      for (int j = 0; j <= d; ++j) fout[j * plane + v] = f[j];
    }
  }
}

}  // namespace tmh
