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
// q * (kCLanes / CH) + tid / CH, column tid % CH of the chunk.  Addresses are
// a workgroup-uniform base plus 32-bit lane byte offsets (a block of rows
// spans at most kCLanes * 65535 elements), recomputed per chunk rather than
// held in CH registers.  Slots outside the rows or the row length load the
// chunk's first element, which no lane reads: the loaded registers are left
// untouched until chunk_put, so the loads stay in flight meanwhile.
__device__ __forceinline__ uint32_t chunk_lane_off(int L, int CH) {
  uint32_t o = (uint32_t)(((int)threadIdx.x / CH) * L + (int)threadIdx.x % CH) * 4u;
  asm volatile("" : "+v"(o));  // keeps the per-slot offsets out of loop-invariant registers
  return o;
}

template <int CH>
__device__ __forceinline__ void chunk_fetch(int32_t (&r)[CH], const int32_t* __restrict__ src,
                                            int64_t row0, int L, int nrows, int i0) {
  constexpr int RPQ = kCLanes / CH;
  const int tr = (int)threadIdx.x / CH, tk = (int)threadIdx.x % CH;
  const char* blk = (const char*)(src + row0 * L + i0);
  const bool colok = i0 + tk < L;
  const uint32_t off0 = chunk_lane_off(L, CH), step = (uint32_t)(RPQ * L) * 4u;
#pragma unroll
  for (int q = 0; q < CH; ++q) {
    const bool ok = colok && q * RPQ + tr < nrows;
    r[q] = *(const int32_t*)(blk + (ok ? off0 + (uint32_t)q * step : 0u));
  }
}

template <int CH>
__device__ __forceinline__ void chunk_put(const int32_t (&r)[CH], int32_t* lds) {
  constexpr int RPQ = kCLanes / CH;
  const int tr = (int)threadIdx.x / CH, tk = (int)threadIdx.x % CH;
#pragma unroll
  for (int q = 0; q < CH; ++q) lds[(q * RPQ + tr) * (CH + 1) + tk] = r[q];
}

template <int CH>
__device__ __forceinline__ void chunk_flush(const int32_t* lds, int32_t* __restrict__ dst,
                                            int64_t row0, int L, int nrows, int i0) {
  constexpr int RPQ = kCLanes / CH;
  const int tr = (int)threadIdx.x / CH, tk = (int)threadIdx.x % CH;
  char* blk = (char*)(dst + row0 * L + i0);
  const bool colok = i0 + tk < L;
  const uint32_t off0 = chunk_lane_off(L, CH), step = (uint32_t)(RPQ * L) * 4u;
#pragma unroll
  for (int q = 0; q < CH; ++q)
    if (colok && q * RPQ + tr < nrows)
      *(int32_t*)(blk + off0 + (uint32_t)q * step) = lds[(q * RPQ + tr) * (CH + 1) + tk];
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

// Envelope stacks of the row passes, blocked per row: entry k of row t sits
// at slot ((k >> 3) * nl + t) * 8 + (k & 7), so eight consecutive entries of
// one row share 32 bytes of E (p | f0 << 16) and 16 bytes of F1 (f1, two
// feature planes only), and the query fetches them eight at a time.  E holds
// rows_stack_slots(L, nl) slots from a 256-byte boundary, F1 as many from the
// next 256-byte boundary after E.  Slot numbers are 32-bit (rows_stack_fits).
__host__ __device__ inline int64_t rows_stack_slots(int64_t L, int64_t nl) {
  return (L + 7) / 8 * 8 * nl;
}
__host__ __device__ inline int64_t rows_stack_bytes(int64_t L, int64_t nl) {
  return 256 + (rows_stack_slots(L, nl) * 4 + 255) / 256 * 256 + rows_stack_slots(L, nl) * 2;
}
__host__ __device__ inline bool rows_stack_fits(int64_t L, int64_t nl) {
  return rows_stack_slots(L, nl) < ((int64_t)1 << 31);
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
// LDS bytes of edt_rows_final<NP, CH>: the chunk staging, NP planes in, then
// the outputs in the same space.
template <int NP, int CH>
constexpr size_t edt_rows_lds() {
  return (size_t)NP * kCLanes * (CH + 1) * sizeof(int32_t);
}

// One block of eight stack entries of a row in registers: e* from E, f*
// from F1 (two entries per word).  Named scalars rather than arrays, so that
// picking entry j stays a select tree on registers.
struct StkBlk {
  uint32_t e0, e1, e2, e3, e4, e5, e6, e7, f0, f1, f2, f3;
};

template <int NP>
__device__ __forceinline__ void stk_fetch(StkBlk& b, const uint32_t* __restrict__ E,
                                          const uint16_t* __restrict__ F1, uint32_t s) {
  const uint4 x = *(const uint4*)(E + s);
  const uint4 y = *(const uint4*)(E + s + 4);
  b.e0 = x.x; b.e1 = x.y; b.e2 = x.z; b.e3 = x.w;
  b.e4 = y.x; b.e5 = y.y; b.e6 = y.z; b.e7 = y.w;
  if (NP == 2) {
    const uint4 z = *(const uint4*)(F1 + s);
    b.f0 = z.x; b.f1 = z.y; b.f2 = z.z; b.f3 = z.w;
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

// EDT, last pass (d = ndim - 1, NP = d feature planes, 1 or 2), distances
// only.  fin: the item's feature planes (plane stride `plane`); rows
// row0 .. row0 + nrows - 1 of the nl rows whose stacks E / F1 hold (blocked
// layout above).  lds: edt_rows_lds<NP, CH>() bytes.
//
// The envelope is the one of edt_line_stack_c, entry for entry: the top two
// entries live in registers, every deeper entry is written once when it is
// buried, and the R most recently buried ones also stay in registers, so a
// pop (rarely more than two deep) reads memory only past those.  The query
// walks the stack with the block of eight holding its next entry in
// registers and the block after it already in flight.
template <int NP, int CH>
__device__ void edt_rows_final(const int32_t* __restrict__ fin, int64_t plane, const Shape& sh,
                               int64_t row0, int nrows, const Samp& sp, uint32_t* __restrict__ E,
                               uint16_t* __restrict__ F1, int64_t nl,
                               float* __restrict__ dist, int32_t* lds) {
  constexpr int RS = CH + 1;
  constexpr int R = 4;
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
  const uint32_t nl8 = (uint32_t)nl * 8u, t8 = (uint32_t)t * 8u;
  const int nch = (L + CH - 1) / CH;
  auto slot = [&](int k) -> uint32_t {
    return (uint32_t)(k >> 3) * nl8 + t8 + (uint32_t)(k & 7);
  };
  auto site_of = [&](uint32_t w, uint32_t f) -> EdtSite {
    EdtSite e;
    e.p = (int32_t)(w & 0xffffu);
    e.f0 = (int32_t)(w >> 16);
    e.f1 = NP == 2 ? (int32_t)f : 0;
    e.h = edt_partial<NP>(e.f0, e.f1, c0, c1, s0, s1);
    return e;
  };
  auto put = [&](int k, const EdtSite& e) {
    const uint32_t s = slot(k);
    E[s] = (uint32_t)(uint16_t)e.p | ((uint32_t)(uint16_t)e.f0 << 16);
    if (NP == 2) F1[s] = (uint16_t)e.f1;
  };

  // build: q1 = entry top, q0 = entry top - 1; memory holds 0 .. top - 2,
  // g0 .. g{nv - 1} = entries top - 2, top - 3, ... (raw, f1 in h*)
  int top = -1, nv = 0;
  EdtSite q1{0, 0, 0, 0.0}, q0{0, 0, 0, 0.0};
  uint32_t g0 = 0, g1 = 0, g2 = 0, g3 = 0, h0 = 0, h1 = 0, h2 = 0, h3 = 0;
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
          if (top >= 1) {
            uint32_t w, wf = 0;
            if (nv > 0) {
              w = g0;
              wf = h0;
              g0 = g1; g1 = g2; g2 = g3;
              h0 = h1; h1 = h2; h2 = h3;
              --nv;
            } else {
              const uint32_t s = slot(top - 1);
              w = E[s];
              if (NP == 2) wf = F1[s];
              // wait for these loads here, not wherever w is next written
              asm volatile("" : "+v"(w), "+v"(wf));
            }
            q0 = site_of(w, wf);
          }
        }
        if (top >= 1) {
          put(top - 1, q0);
          g3 = g2; g2 = g1; g1 = g0;
          h3 = h2; h2 = h1; h1 = h0;
          g0 = (uint32_t)(uint16_t)q0.p | ((uint32_t)(uint16_t)q0.f0 << 16);
          h0 = (uint32_t)(uint16_t)q0.f1;
          nv = nv < R ? nv + 1 : R;
        }
        ++top;
        q0 = q1;
        q1 = EdtSite{i, f0, f1, wR};
      }
    }
  }
  if (active && top >= 1) put(top - 1, q0);
  if (active && top >= 0) put(top, q1);

  // query: cur = entry l, nx = entry l + 1, n = l + 2 the next entry to
  // decode; wa = block n >> 3, wb = the block after it (loaded when it holds
  // entries)
  __syncthreads();  // the staging area holds outputs from here on
  int32_t* ob = lds;
  int l = 0, n = 2;
  EdtSite cur{0, 0, 0, 0.0}, nx{0, 0, 0, 0.0};
  StkBlk wa{}, wb{};
  if (active && top >= 0) {
    stk_fetch<NP>(wa, E, F1, t8);
    if (top >= 8) stk_fetch<NP>(wb, E, F1, nl8 + t8);
    cur = site_of(wa.e0, wa.f0 & 0xffffu);
    if (top >= 1) nx = site_of(wa.e1, wa.f0 >> 16);
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
              const int j = n & 7;
              nx = site_of(stk_pick_e(wa, j), NP == 2 ? stk_pick_f(wa, j) : 0u);
              ++n;
              if ((n & 7) == 0) {
                wa = wb;
                if ((n >> 3) + 1 <= (top >> 3))
                  stk_fetch<NP>(wb, E, F1, (uint32_t)((n >> 3) + 1) * nl8 + t8);
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
// values; the cost per voxel is its distance.  The window grows four steps
// per round: the eight loads of a round are issued together (indices
// clamped into the line, the out-of-line ones then ignored), so a voxel at
// distance r waits for ceil(r / 4) rounds of loads rather than r.
__device__ __forceinline__ int32_t cheb_window(const int32_t* __restrict__ g, int64_t base,
                                               int64_t step, int L, int i) {
  const int32_t* p = g + base;
  int32_t m = p[(int64_t)i * step];
  if (m <= 0) return m;
  const int rfull = i > L - 1 - i ? i : L - 1 - i;
  for (int r0 = 1; r0 <= rfull; r0 += 4) {
    int32_t lo[4], hi[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int a = i - r0 - q, b = i + r0 + q;
      lo[q] = p[(int64_t)(a >= 0 ? a : 0) * step];
      hi[q] = p[(int64_t)(b < L ? b : L - 1) * step];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int r = r0 + q;
      if (r > rfull) break;
      if (i - r >= 0 && lo[q] < m) m = lo[q];
      if (i + r < L && hi[q] < m) m = hi[q];
      if (m <= r) return r;
    }
  }
  return m;
}

}  // namespace tmh
