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
// are read where they are used: items of four or more axes).  The loads are
// unconditional (past the line's end they read its last voxel, which the
// consumer skips) and nothing touches the registers until the chunk is
// consumed, so they stay in flight meanwhile.
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
    ch.f0[k] = fin[base + (int64_t)(i0 + k < L ? i0 + k : L - 1) * step];
  if (d >= 2) {
#pragma unroll
    for (int k = 0; k < PF; ++k)
      ch.f1[k] = fin[plane + base + (int64_t)(i0 + k < L ? i0 + k : L - 1) * step];
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
      if (i0 + k >= L || f0 < 0) continue;
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
// the site and its features on the earlier axes -- E = site | f0 << 16,
// F1 = f1 (d = 2) -- 4 or 6 bytes instead of the site and its double partial
// sum: the partial sum is recomputed from f0 (and f1) with the same
// operations as at the push, so it has the same bits.  The top two entries
// live in registers and only entries buried two deep are written, once; the
// last four buried ones also stay in registers, so a pop (rarely more than
// two deep) reads memory only past those.  The query walks the stack with
// the block of eight entries holding its next site in registers and the
// block after it already in flight, and reads the site's features from its
// entry (no gather of the feature planes).
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

// Where a line's stack lives.  BLK: blocked, entry k of line t at slot
// ((k >> 3) * nl + t) * 8 + (k & 7) of E and F1, so eight consecutive
// entries of a line share 32 bytes of E (16 of F1) -- global memory.
// Otherwise entry k at E[k * ss] / F1[k * ss] with E / F1 already offset to
// the line -- LDS, lanes interleaved.
template <bool BLK>
struct StkAddr {
  uint32_t* E;
  uint16_t* F1;
  int64_t ss;       // !BLK
  int64_t nl8, t8;  // BLK: 8 * lines, 8 * line
  __device__ __forceinline__ int64_t at(int k) const {
    return BLK ? (int64_t)(k >> 3) * nl8 + t8 + (k & 7) : (int64_t)k * ss;
  }
};

// Bytes of the blocked stacks of nl lines of length L: E holds
// rows_stack_slots(L, nl) slots from a 256-byte boundary, F1 as many from the
// next 256-byte boundary after E.
__host__ __device__ inline int64_t rows_stack_slots(int64_t L, int64_t nl) {
  return (L + 7) / 8 * 8 * nl;
}
__host__ __device__ inline int64_t rows_stack_bytes(int64_t L, int64_t nl) {
  return 256 + (rows_stack_slots(L, nl) * 4 + 255) / 256 * 256 + rows_stack_slots(L, nl) * 2;
}
// E and F1 for a stack area starting at `area` (any 4-byte boundary); plain
// pointer offsets, so the global address space carries over.
__host__ __device__ inline uint32_t* rows_stack_e(void* area) {
  return (uint32_t*)((char*)area + ((256 - ((uintptr_t)area & 255)) & 255));
}
__host__ __device__ inline uint16_t* rows_stack_f1(void* area, int64_t L, int64_t nl) {
  char* e = (char*)rows_stack_e(area) + rows_stack_slots(L, nl) * 4;
  return (uint16_t*)(e + ((256 - ((uintptr_t)e & 255)) & 255));
}

// One block of eight stack entries of a line in registers: e* from E, f*
// from F1 (two entries per word).  Named scalars rather than arrays, so that
// picking entry j stays a select tree on registers.
struct StkBlk {
  uint32_t e0, e1, e2, e3, e4, e5, e6, e7, f0, f1, f2, f3;
};

// Block b (entries 8b .. 8b + 7) of a stack whose top entry is `top`:
// blocked, two 16-byte loads (and one of F1); interleaved, eight loads with
// the entries past the top reading the top one.
template <int NP, bool BLK>
__device__ __forceinline__ void stk_fetch(StkBlk& w, const StkAddr<BLK>& A, int b, int top) {
  if (BLK) {
    const int64_t s = A.at(8 * b);
    const uint4 x = *(const uint4*)(A.E + s);
    const uint4 y = *(const uint4*)(A.E + s + 4);
    w.e0 = x.x; w.e1 = x.y; w.e2 = x.z; w.e3 = x.w;
    w.e4 = y.x; w.e5 = y.y; w.e6 = y.z; w.e7 = y.w;
    if (NP == 2) {
      const uint4 z = *(const uint4*)(A.F1 + s);
      w.f0 = z.x; w.f1 = z.y; w.f2 = z.z; w.f3 = z.w;
    }
  } else {
    int64_t s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = A.at(8 * b + j < top ? 8 * b + j : top);
    w.e0 = A.E[s[0]]; w.e1 = A.E[s[1]]; w.e2 = A.E[s[2]]; w.e3 = A.E[s[3]];
    w.e4 = A.E[s[4]]; w.e5 = A.E[s[5]]; w.e6 = A.E[s[6]]; w.e7 = A.E[s[7]];
    if (NP == 2) {
      w.f0 = (uint32_t)A.F1[s[0]] | ((uint32_t)A.F1[s[1]] << 16);
      w.f1 = (uint32_t)A.F1[s[2]] | ((uint32_t)A.F1[s[3]] << 16);
      w.f2 = (uint32_t)A.F1[s[4]] | ((uint32_t)A.F1[s[5]] << 16);
      w.f3 = (uint32_t)A.F1[s[6]] | ((uint32_t)A.F1[s[7]] << 16);
    }
  }
}

__device__ __forceinline__ uint32_t stk_pick_e(const StkBlk& b, int j) {
  const bool s0 = (j & 1) != 0, s1 = (j & 2) != 0, s2 = (j & 4) != 0;
  const uint32_t x0 = s0 ? b.e1 : b.e0, x1 = s0 ? b.e3 : b.e2;
  const uint32_t x2 = s0 ? b.e5 : b.e4, x3 = s0 ? b.e7 : b.e6;
  const uint32_t y0 = s1 ? x1 : x0, y1 = s1 ? x3 : x2;
  return s2 ? y1 : y0;
}

__device__ __forceinline__ uint32_t stk_pick_f(const StkBlk& b, int j) {
  const bool s1 = (j & 2) != 0, s2 = (j & 4) != 0;
  const uint32_t y0 = s1 ? b.f1 : b.f0, y1 = s1 ? b.f3 : b.f2;
  const uint32_t v = s2 ? y1 : y0;
  return (j & 1) ? v >> 16 : v & 0xffffu;
}

// The envelope of one line along axis d (NP = d = 1 or 2 feature planes
// below it): add() the sites in order, close(), then walk().
template <int NP, bool BLK>
struct Envelope {
  StkAddr<BLK> A;
  int32_t c0, c1;
  double s0, s1, sd;
  int top = -1, nv = 0;
  EdtSite q1{0, 0, 0, 0.0}, q0{0, 0, 0, 0.0};  // entries top, top - 1
  uint32_t g0 = 0, g1 = 0, g2 = 0, g3 = 0;    // entries top - 2, top - 3, ... (raw)
  uint32_t h0 = 0, h1 = 0, h2 = 0, h3 = 0;    // their f1

  __device__ __forceinline__ EdtSite site(uint32_t w, uint32_t f) const {
    EdtSite e;
    e.p = (int32_t)(w & 0xffffu);
    e.f0 = (int32_t)(w >> 16);
    e.f1 = NP == 2 ? (int32_t)f : 0;
    e.h = edt_partial<NP>(e.f0, e.f1, c0, c1, s0, s1);
    return e;
  }
  __device__ __forceinline__ void put(int k, const EdtSite& e) {
    const int64_t s = A.at(k);
    A.E[s] = (uint32_t)(uint16_t)e.p | ((uint32_t)(uint16_t)e.f0 << 16);
    if (NP == 2) A.F1[s] = (uint16_t)e.f1;
  }
  // Site i with features f0, f1 and partial sum wR (edt_partial): pops the
  // entries it hides (the removal test of the reference's envelope), pushes.
  __device__ __forceinline__ void add(int i, int32_t f0, int32_t f1, double wR) {
    while (top >= 1) {
      const double a = (double)(q1.p - q0.p) * sd;
      const double bb = (double)(i - q1.p) * sd;
      const double cc = a + bb;
      if (cc * q1.h - bb * q0.h - a * wR - a * bb * cc <= 0.0) break;
      --top;
      q1 = q0;
      if (top >= 1) {
        uint32_t w, wf = 0;
        if (BLK && nv > 0) {
          w = g0;
          wf = h0;
          g0 = g1; g1 = g2; g2 = g3;
          h0 = h1; h1 = h2; h2 = h3;
          --nv;
        } else {
          const int64_t s = A.at(top - 1);
          w = A.E[s];
          if (NP == 2) wf = A.F1[s];
          // wait for these loads here, not wherever w is next written
          asm volatile("" : "+v"(w), "+v"(wf));
        }
        q0 = site(w, wf);
      }
    }
    if (top >= 1) {
      put(top - 1, q0);
      if (BLK) {  // an LDS stack is read back at LDS latency: no copies kept
        g3 = g2; g2 = g1; g1 = g0;
        h3 = h2; h2 = h1; h1 = h0;
        g0 = (uint32_t)(uint16_t)q0.p | ((uint32_t)(uint16_t)q0.f0 << 16);
        h0 = (uint32_t)(uint16_t)q0.f1;
        nv = nv < 4 ? nv + 1 : 4;
      }
    }
    ++top;
    q0 = q1;
    q1 = EdtSite{i, f0, f1, wR};
  }
  // Writes the two register entries: memory then holds entries 0 .. top.
  __device__ __forceinline__ void close() {
    if (top >= 1) put(top - 1, q0);
    if (top >= 0) put(top, q1);
  }
};

// The query over a closed envelope (top >= 0): for voxels i = 0, 1, ... in
// order, next(i) returns the squared distance to the envelope's site for i
// (the walk of the reference's query: advance while the next site is
// strictly nearer) and leaves that site in cur.  Blocked stacks are read a
// block of eight ahead; LDS stacks one entry ahead.
template <int NP, bool BLK>
struct EnvWalk {
  EdtSite cur{0, 0, 0, 0.0}, nx{0, 0, 0, 0.0};  // entries l, l + 1
  int l = 0, n = 2;                              // n = l + 2: next entry to decode
  StkBlk wa{}, wb{};                             // BLK: blocks n >> 3 and the one after
  uint32_t nnE = 0, nnF = 0;                     // !BLK: entry n

  __device__ __forceinline__ void start(const Envelope<NP, BLK>& e) {
    if (BLK) {
      stk_fetch<NP, BLK>(wa, e.A, 0, e.top);
      if (e.top >= 8) stk_fetch<NP, BLK>(wb, e.A, 1, e.top);
      cur = e.site(wa.e0, wa.f0 & 0xffffu);
      if (e.top >= 1) nx = e.site(wa.e1, wa.f0 >> 16);
    } else {
      cur = e.site(e.A.E[0], NP == 2 ? e.A.F1[0] : 0u);
      if (e.top >= 1) nx = e.site(e.A.E[e.A.at(1)], NP == 2 ? e.A.F1[e.A.at(1)] : 0u);
      if (e.top >= 2) {
        nnE = e.A.E[e.A.at(2)];
        if (NP == 2) nnF = e.A.F1[e.A.at(2)];
      }
    }
  }
  __device__ __forceinline__ double next(const Envelope<NP, BLK>& e, int i) {
    const int top = e.top;
    double tt = (double)(cur.p - i) * e.sd;
    double d1 = cur.h + tt * tt;
    while (l < top) {
      double t2 = (double)(nx.p - i) * e.sd;
      double d2 = nx.h + t2 * t2;
      if (d1 <= d2) break;
      d1 = d2;
      ++l;
      cur = nx;
      if (l < top) {
        if (BLK) {
          const int j = n & 7;
          nx = e.site(stk_pick_e(wa, j), NP == 2 ? stk_pick_f(wa, j) : 0u);
          ++n;
          if ((n & 7) == 0) {
            wa = wb;
            if ((n >> 3) + 1 <= (top >> 3)) stk_fetch<NP, BLK>(wb, e.A, (n >> 3) + 1, top);
          }
        } else {
          nx = e.site(nnE, nnF);
          ++n;
          if (n <= top) {
            nnE = e.A.E[e.A.at(n)];
            if (NP == 2) nnF = e.A.F1[e.A.at(n)];
          }
        }
      }
    }
    return d1;
  }
};

template <bool FINAL, int D, int PF, bool BLK>
__device__ __forceinline__ void edt_line_stack_c(const int32_t* __restrict__ fin,
                                                 int32_t* __restrict__ fout, int64_t plane,
                                                 const Shape& sh, int64_t t, const Samp& sp,
                                                 const StkAddr<BLK>& stk, const EdtOut& o) {
  const int d = D;
  int64_t c[kMaxNdim], base, b;
  decode_line(sh, d, t, c, &base, &b);
  const int64_t step = sh.stride[d];
  const int L = (int)sh.size[d];
  Envelope<D, BLK> env;
  env.A = stk;
  env.sd = sp.s[d];
  env.s0 = sp.s[0];
  env.s1 = sp.s[1];
  env.c0 = (int32_t)c[0];
  env.c1 = (int32_t)c[1];

  LineChunk<PF> ch;
  line_chunk_load(ch, fin, plane, base, step, 0, L, d);
  for (int i0 = 0; i0 < L; i0 += PF) {
    LineChunk<PF> cur = ch;
    if (i0 + PF < L) line_chunk_load(ch, fin, plane, base, step, i0 + PF, L, d);
#pragma unroll
    for (int k = 0; k < PF; ++k) {
      const int32_t f0 = cur.f0[k];
      if (i0 + k >= L || f0 < 0) continue;
      const int32_t f1 = D == 2 ? cur.f1[k] : 0;
      env.add(i0 + k, f0, f1, edt_partial<D>(f0, f1, env.c0, env.c1, env.s0, env.s1));
    }
  }
  env.close();

  int32_t f[kMaxNdim];
  if (env.top < 0) {
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

  const bool dist_only = FINAL && !o.idx;
  EnvWalk<D, BLK> q;
  q.start(env);
  for (int i = 0; i < L; ++i) {
    const double d1 = q.next(env, i);
    const int64_t v = base + (int64_t)i * step;
    if (dist_only) {
      o.dist[v] = (float)sqrt(d1);
      continue;
    }
    f[0] = q.cur.f0;
    if (D == 2) f[1] = q.cur.f1;
    f[d] = q.cur.p;
    if (FINAL) {
      c[d] = i;
      edt_store(o, sh.ndim, v, c, f, sp);
    } else {
      for (int j = 0; j <= d; ++j) fout[j * plane + v] = f[j];
    }
  }
}

}  // namespace tmh
