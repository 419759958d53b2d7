// Exact Euclidean distance transform for gfx950.
//
// Semantics are SciPy's distance_transform_edt, which the reference tests pin
// (reference test/test_distance_transform.py:32-69, :289-320, :397-408): the
// feature transform is the separable lower-envelope algorithm of Maurer, Qi
// and Raghavan (one pass per axis, axis 0 first), with double arithmetic,
// removal of the middle site when c*vR - b*uR - a*wR - a*b*c > 0, and a query
// that moves to the next envelope site only when it is strictly closer (the
// lower index wins ties).  Foreground voxels with no background anywhere get
// the feature (-1, 0, ..., 0), the reference's virtual background
// (edt_kernel.cu:415-437).  Distances are computed in double from the integer
// feature deltas in axis order and rounded once to float32.
//
// Kernels:
//  * edt2d_cols_kernel / edt2d_rows_kernel -- 2-D, unit sampling, H <= 1024,
//    W <= 4096.  Axis 0: nearest background row per pixel from bit-packed
//    column segments (one load round), int16 into the workspace.  Axis 1:
//    band of rows staged in LDS, exact expanding-window argmin per pixel with
//    the lower-index tie rule (exact integers, so equal to the envelope); a
//    row where some window passes kWinCap steps (sparse background) is
//    resolved through its exact integer lower envelope, built in LDS.
//  * edt_pass0_kernel   -- N-D, axis 0 from the mask: two sweeps per line.
//  * edt_stack_kernel   -- N-D, axis d >= 1: one lane per line runs the
//    double-precision envelope (stack of sites with their partial squared
//    distances) exactly in SciPy's operation order, with loads issued ahead
//    and the stack tops in registers (tm_dt_fast.h); the stack is in LDS when
//    the line fits the LDS depth, otherwise in an interleaved global scratch.
//
// Build with -ffp-contract=off: every double multiply and add must round
// separately, as SciPy's C code does.
#include <algorithm>
#include <mutex>

#include "tm_dt_chunk.h"
#include "tm_dt_fast.h"
#include "tm_dt_lines.h"
#include "tm_tune.h"

namespace tmh {

// ------------------------------------------------------------- 2-D path --
// Axis-0 pass.  A workgroup covers 64 columns of one image with 16 column
// groups (4 adjacent columns each) x 16 row segments.  Every thread issues all
// of its mask loads up front (one 4-column word per row, NCH chunks of 16
// rows), keeps them as background bit masks, and the segments exchange their
// first/last background rows through LDS: the kernel has one dependent
// memory round trip.  Output: nearest background row per pixel (int16,
// -1 = none), the lower row on ties.
constexpr int kColSegs = 16;
constexpr int kChunkRows = 16;

template <typename TIn, int CPT>
__device__ __forceinline__ uint32_t bgbits(const TIn* __restrict__ p, int x, int W, bool vec) {
  // background bits of columns x..x+CPT-1 of one row
  if (CPT == 4 && vec) {
    if constexpr (sizeof(TIn) == 1) {
      const uint32_t v = *(const uint32_t*)(p + x);
      return (uint32_t)((v & 0xffu) == 0) | ((uint32_t)((v & 0xff00u) == 0) << 1) |
             ((uint32_t)((v & 0xff0000u) == 0) << 2) | ((uint32_t)((v & 0xff000000u) == 0) << 3);
    } else {
      const float4 v = *(const float4*)(p + x);
      return (uint32_t)(v.x == 0.0f) | ((uint32_t)(v.y == 0.0f) << 1) |
             ((uint32_t)(v.z == 0.0f) << 2) | ((uint32_t)(v.w == 0.0f) << 3);
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int c = 0; c < CPT; ++c)
    if (x + c < W && p[x + c] == TIn(0)) r |= 1u << c;
  return r;
}

// Background bits of a loaded word: RAWW words are the raw data (one
// element for CPT = 1, four bytes for CPT = 4), turned into bits only after
// every load of the thread has been issued; other words are bits already.
template <typename TIn, int CPT, bool RAWW>
__device__ __forceinline__ uint32_t word_bits(uint32_t w) {
  if (!RAWW) return w;
  if (CPT == 1) {
    if constexpr (sizeof(TIn) == 4) return __uint_as_float(w) == 0.0f ? 1u : 0u;
    else return w == 0 ? 1u : 0u;
  }
  return (uint32_t)((w & 0xffu) == 0) | ((uint32_t)((w & 0xff00u) == 0) << 1) |
         ((uint32_t)((w & 0xff0000u) == 0) << 2) | ((uint32_t)((w & 0xff000000u) == 0) << 3);
}

// Column summary of one (column, segment of at most 32 rows): background
// bits of the segment's rows, the last background row above the segment and
// the first below it (-1: none).
__device__ __forceinline__ uint64_t colsum_pack(uint32_t mask, int before, int after) {
  return (uint64_t)mask | ((uint64_t)(uint16_t)before << 32) | ((uint64_t)(uint16_t)after << 48);
}

// Nearest background row of row y in a column from its summary (segment
// starting at row ys): the row above wins ties, as in the column pass.
__device__ __forceinline__ int colsum_nearest(uint64_t q, int y, int ys) {
  const uint32_t m = (uint32_t)q;
  const int b = y - ys;
  const uint32_t up = m & (b >= 31 ? 0xffffffffu : ((2u << b) - 1u)), dn = m >> b;
  const int pv = up ? ys + 31 - __builtin_clz(up) : (int16_t)(uint16_t)(q >> 32);
  const int nx = dn ? y + __builtin_ctz(dn) : (int16_t)(uint16_t)(q >> 48);
  return (pv >= 0 && (nx < 0 || (y - pv) <= (nx - y))) ? pv : nx;
}

// CPT adjacent columns per thread (1 or 4), 16 column groups x 16 row
// segments per workgroup, NCH chunks of 16 rows per segment.  VEC: 4-byte
// aligned rows of a byte mask (W % 4 == 0).
// CG column groups x (256 / CG) row segments per workgroup: fewer columns
// per workgroup give more workgroups (CG = 16: 16 columns, 16 segments).
template <typename TIn, int NCH, int CPT, bool VEC, int CG = 16>
__global__ __launch_bounds__(256) void edt2d_cols_kernel(const TIn* __restrict__ in, int H, int W,
                                                         int16_t* __restrict__ f0, int aligned) {
  constexpr int SEGS = 256 / CG;
  // raw words: one element per row (CPT = 1) or four mask bytes (CPT = 4,
  // VEC), loaded unconditionally so that all of them are in flight at once
  constexpr bool RAWW = CPT == 1 || (VEC && sizeof(TIn) == 1);
  constexpr int WGC = CG * CPT;  // columns per workgroup
  __shared__ int16_t sfirst[SEGS][WGC];
  __shared__ int16_t slast[SEGS][WGC];
  const int cg = threadIdx.x & (CG - 1);
  const int seg = threadIdx.x / CG;
  // 1-D grid of (image, column block); the column blocks of one image stay on
  // one XCD (they share the image's cache lines)
  const int ncb = (W + WGC - 1) / WGC;
  const int lb = xcd_swizzle(blockIdx.x, gridDim.x);
  const int img = lb / ncb;
  const int xb = (lb - img * ncb) * WGC + cg * CPT;
  const int seglen = (H + SEGS - 1) / SEGS;
  const int ys = seg * seglen;
  const int ye = min(H, ys + seglen);
  const TIn* im = in + (int64_t)img * H * W;
  const bool vec = aligned && (W % 4 == 0) && (xb + CPT <= W);
  const bool live = xb < W && ys < ye;

  uint32_t raw[NCH][kChunkRows];
  if (live) {
#pragma unroll
    for (int k = 0; k < NCH; ++k)
#pragma unroll
      for (int b = 0; b < kChunkRows; ++b) {
        const int y = min(ys + k * kChunkRows + b, ye - 1);
        const TIn* row = im + (int64_t)y * W;
        if constexpr (RAWW && CPT == 1) {
          if constexpr (sizeof(TIn) == 4) raw[k][b] = __float_as_uint((float)row[xb]);
          else raw[k][b] = (uint32_t)row[xb];
        } else if constexpr (RAWW) {
          raw[k][b] = *(const uint32_t*)(row + xb);
        } else {
          raw[k][b] = bgbits<TIn, CPT>(row, xb, W, vec);
        }
      }
  }
  uint32_t m[CPT][NCH];
#pragma unroll
  for (int k = 0; k < NCH; ++k) {
    uint32_t mk[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) mk[c] = 0;
#pragma unroll
    for (int b = 0; b < kChunkRows; ++b) {
      const bool rowok = live && (ys + k * kChunkRows + b < ye);
      const uint32_t v = rowok ? word_bits<TIn, CPT, RAWW>(raw[k][b]) : 0u;
#pragma unroll
      for (int c = 0; c < CPT; ++c) mk[c] |= ((v >> c) & 1u) << b;
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) m[c][k] = mk[c];
  }
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    int first = -1, last = -1;
#pragma unroll
    for (int k = NCH - 1; k >= 0; --k)
      if (m[c][k]) first = ys + k * kChunkRows + __builtin_ctz(m[c][k]);
#pragma unroll
    for (int k = 0; k < NCH; ++k)
      if (m[c][k]) last = ys + k * kChunkRows + 31 - __builtin_clz(m[c][k]);
    sfirst[seg][cg * CPT + c] = (int16_t)first;
    slast[seg][cg * CPT + c] = (int16_t)last;
  }
  __syncthreads();
  if (live) {
  int before[CPT], after[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    before[c] = -1;
    after[c] = -1;
    for (int s = seg - 1; s >= 0 && before[c] < 0; --s) before[c] = slast[s][cg * CPT + c];
    for (int s = seg + 1; s < SEGS && after[c] < 0; ++s) after[c] = sfirst[s][cg * CPT + c];
  }
#pragma unroll
  for (int k = 0; k < NCH; ++k) {
#pragma unroll
    for (int b = 0; b < kChunkRows; ++b) {
      const int y = ys + k * kChunkRows + b;
      if (y >= ye) break;
      int16_t out[CPT];
#pragma unroll
      for (int c = 0; c < CPT; ++c) {
        // nearest background at or above y
        int pv = -1;
        uint32_t mm = m[c][k] & (0xffffffffu >> (31 - b));
        if (mm) {
          pv = ys + k * kChunkRows + 31 - __builtin_clz(mm);
        } else {
#pragma unroll
          for (int kk = k - 1; kk >= 0; --kk)
            if (pv < 0 && m[c][kk]) pv = ys + kk * kChunkRows + 31 - __builtin_clz(m[c][kk]);
          if (pv < 0) pv = before[c];
        }
        // nearest background at or below y
        int nx = -1;
        mm = m[c][k] >> b;
        if (mm) {
          nx = y + __builtin_ctz(mm);
        } else {
#pragma unroll
          for (int kk = k + 1; kk < NCH; ++kk)
            if (nx < 0 && m[c][kk]) nx = ys + kk * kChunkRows + __builtin_ctz(m[c][kk]);
          if (nx < 0) nx = after[c];
        }
        int f = -1;
        if (pv >= 0 && (nx < 0 || (y - pv) <= (nx - y))) f = pv;
        else if (nx >= 0) f = nx;
        out[c] = (int16_t)f;
      }
      int16_t* dst = f0 + ((int64_t)img * H + y) * W + xb;
      if (CPT == 4 && vec) {
        uint2 pk;
        pk.x = (uint32_t)(uint16_t)out[0] | ((uint32_t)(uint16_t)out[1 % CPT] << 16);
        pk.y = (uint32_t)(uint16_t)out[2 % CPT] | ((uint32_t)(uint16_t)out[3 % CPT] << 16);
        *(uint2*)dst = pk;
      } else {
#pragma unroll
        for (int c = 0; c < CPT; ++c)
          if (xb + c < W) dst[c] = out[c];
      }
    }
  }
  }  // live
}

// Window steps a pixel may take before its row is resolved through the row's
// lower envelope instead.  The expanding window costs O(distance to the
// nearest site) per pixel, which makes a row with sparse background
// quadratic in its length; the envelope costs O(W) per row plus a binary
// search per pixel.
constexpr int kWinCap = 64;

__device__ __forceinline__ int floor_div(int n, int d) {  // d > 0
  return n >= 0 ? n / d : -((-n + d - 1) / d);
}

// Site heights of one row: h(x) = (y - fr[x])^2 for the feature row fr[x]
// of column x (fr[x] < 0: the column has no background, no site).
__device__ __forceinline__ int site_h(const int16_t* frr, int y, int x) {
  const int d = y - (int)frr[x];
  return d * d;
}

// Lower envelope of the parabolas h(x) + (q - x)^2 over the integer points
// q = 0 .. W-1, in exact integer arithmetic, built in place over the row's
// sites listed in column order as E[i] = (h << 12) | x (n entries; h <
// 2^20, x < 4096).  Entry k becomes (b << 12) | site: the site and the first
// q at which it wins (b = 0 for entry 0, strictly increasing).  A site wins
// from the first q where it is strictly below the entry before it, so the
// lower column keeps every tie: the result at q is the lowest column
// attaining the minimum, the window's rule.  A site that wins at no q of the
// row is dropped.  Entry k is written only after list entry k was read.  Run
// by one thread; returns the index of the last entry (-1: no site).
__device__ int row_env_build(uint32_t* E, int n, int W, const int16_t* frr, int y) {
  int top = -1;
  int tp = 0, tc = 0, tb = 0;  // top entry: site, h(site) + site^2, first q
  uint32_t nxt = n > 0 ? E[0] : 0u;
  for (int i = 0; i < n; ++i) {
    const uint32_t v = nxt;
    if (i + 1 < n) nxt = E[i + 1];  // entries above i are never written here
    const int x = (int)(v & 0xfffu);
    const int cq = (int)(v >> 12) + x * x;
    int bq = 0;
    while (top >= 0) {
      // first q with h + (q - x)^2 < h(tp) + (q - tp)^2, i.e. 2q(x - tp) > cq - tc
      const int B = floor_div(cq - tc, 2 * (x - tp)) + 1;
      if (B > tb) {
        bq = B;
        break;
      }
      --top;  // the top site wins at no q
      if (top >= 0) {
        const uint32_t t = E[top];
        tp = (int)(t & 0xfffu);
        tb = (int)(t >> 12);
        tc = site_h(frr, y, tp) + tp * tp;
      }
    }
    if (bq > W - 1) continue;
    ++top;
    E[top] = (uint32_t)x | ((uint32_t)bq << 12);
    tp = x;
    tc = cq;
    tb = bq;
  }
  return top;
}

// The envelope's site at q: the last entry that starts at or before q.
__device__ __forceinline__ int row_env_site(const uint32_t* E, int top, int q) {
  int lo = 0, hi = top;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if ((int)(E[mid] >> 12) <= q) lo = mid;
    else hi = mid - 1;
  }
  return (int)(E[lo] & 0xfffu);
}

// LDS of the envelope fallback: per row of the band a flag and one mark bit
// per pixel group left unresolved by the window, the site count and four
// wave totals of the compaction.  The envelope itself is built in the row's
// squared-distance area, which the window no longer reads by then.
struct RowEnvLds {
  uint32_t* mark;  // [R][mw]
  int* flag;       // [R]
  int* nsite;      // [1]
  int* wsum;       // [4]
  int mw;          // mark words per row
};

__host__ __device__ constexpr size_t row_env_lds_bytes(int R, int W, int G) {
  return (size_t)R * (((W + G - 1) / G + 31) / 32) * 4 + (size_t)R * 4 + 32;
}

__device__ __forceinline__ RowEnvLds row_env_lds(char* base, int R, int W, int G) {
  RowEnvLds e;
  e.mw = ((W + G - 1) / G + 31) / 32;
  e.mark = (uint32_t*)base;
  e.flag = (int*)(e.mark + R * e.mw);
  e.nsite = e.flag + R;
  e.wsum = e.nsite + 1;
  return e;
}

// Clears the marks and flags (before the workgroup's first barrier).
__device__ __forceinline__ void row_env_clear(const RowEnvLds& e, int R) {
  for (int i = threadIdx.x; i < R * e.mw; i += blockDim.x) e.mark[i] = 0;
  for (int i = threadIdx.x; i < R; i += blockDim.x) e.flag[i] = 0;
}

__device__ __forceinline__ void row_env_mark(const RowEnvLds& e, int r, int bit) {
  atomicOr(&e.mark[r * e.mw + (bit >> 5)], 1u << (bit & 31));
  e.flag[r] = 1;
}

// Writes G adjacent pixels (x0 .. x0+G-1 of row y) from their minima best
// (-1: no background in the whole image) and winning columns bp.
template <int G>
__device__ __forceinline__ void edt2d_store(const EdtOut& o, const int16_t* frr, int img, int H,
                                            int W, int y, int x0, const int* best,
                                            const int* bp) {
  const int64_t g = ((int64_t)img * H + y) * W + x0;
  float dist[G];
  int32_t fy[G], fx[G];
#pragma unroll
  for (int i = 0; i < G; ++i) {
    double acc;
    if (best[i] >= 0) {
      fy[i] = frr[bp[i]];
      fx[i] = bp[i];
      acc = (double)best[i];  // exact integer: same as the double delta/square/sum
    } else {
      fy[i] = -1;  // no background in the whole image
      fx[i] = 0;
      const double a = (double)(-1 - y), b = (double)(-(x0 + i));
      acc = a * a + b * b;
    }
    dist[i] = (float)sqrt(acc);
  }
  if (G == 4) {
    if (o.dist) *(float4*)(o.dist + g) = make_float4(dist[0], dist[1 % G], dist[2 % G], dist[3 % G]);
    if (o.idx) {
      if (o.idx64) {
        int64_t* ip = (int64_t*)o.idx;
        longlong2* a = (longlong2*)(ip + g);
        longlong2* b = (longlong2*)(ip + o.plane + g);
        a[0] = make_longlong2(fy[0], fy[1 % G]);
        a[1] = make_longlong2(fy[2 % G], fy[3 % G]);
        b[0] = make_longlong2(fx[0], fx[1 % G]);
        b[1] = make_longlong2(fx[2 % G], fx[3 % G]);
      } else {
        int32_t* ip = (int32_t*)o.idx;
        *(int4*)(ip + g) = make_int4(fy[0], fy[1 % G], fy[2 % G], fy[3 % G]);
        *(int4*)(ip + o.plane + g) = make_int4(fx[0], fx[1 % G], fx[2 % G], fx[3 % G]);
      }
    }
  } else {
#pragma unroll
    for (int i = 0; i < G; ++i) {
      if (o.dist) o.dist[g + i] = dist[i];
      if (o.idx) {
        if (o.idx64) {
          int64_t* ip = (int64_t*)o.idx;
          ip[g + i] = fy[i];
          ip[o.plane + g + i] = fx[i];
        } else {
          int32_t* ip = (int32_t*)o.idx;
          ip[g + i] = fy[i];
          ip[o.plane + g + i] = fx[i];
        }
      }
    }
  }
}

// Resolves the marked pixel groups of every flagged row of the band through
// the row's envelope.  Called by every thread of the workgroup, whose size is
// a multiple of 64 (barriers inside, at the top level of loops whose trip
// counts come from LDS values read after a barrier).  hs:
// squared column distances, row pitch P, entry x of row r at r * P + off (W
// entries, overwritten by the envelope); fr: feature rows, pitch W.
// One flagged row r.
template <int G>
__device__ __forceinline__ void edt2d_resolve_row(const RowEnvLds& e, int32_t* hs, int P, int off,
                                               const int16_t* fr, int r, int img, int H, int W,
                                               int y0, const EdtOut& o) {
  const int T = blockDim.x;
  const int C = (W + T - 1) / T;  // columns per thread in the compaction
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  {
    const int y = y0 + r;
    const int16_t* frr = fr + r * W;
    uint32_t* E = (uint32_t*)(hs + r * P + off);
    // the row's sites in column order, as (h << 12) | x: per-thread counts,
    // wave prefix sums, then every thread writes its own at its offset
    const int c0 = min(W, (int)threadIdx.x * C), c1 = min(W, c0 + C);
    int cnt = 0;
    for (int x = c0; x < c1; ++x) cnt += frr[x] >= 0 ? 1 : 0;
    int incl = cnt;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(incl, d);
      if (lane >= d) incl += t;
    }
    if (lane == 63) e.wsum[wave] = incl;
    __syncthreads();
    int base = 0;
    for (int w = 0; w < wave; ++w) base += e.wsum[w];
    int pos = base + incl - cnt;
    for (int x = c0; x < c1; ++x) {
      const int f = frr[x];
      if (f >= 0) E[pos++] = ((uint32_t)((y - f) * (y - f)) << 12) | (uint32_t)x;
    }
    if ((int)threadIdx.x == T - 1) *e.nsite = base + incl;
    __syncthreads();
    if (threadIdx.x == 0) *e.nsite = row_env_build(E, *e.nsite, W, frr, y);
    __syncthreads();
    const int top = *e.nsite;
    const int nb = (W + G - 1) / G;
    for (int q = threadIdx.x; q < nb; q += T) {
      if (!((e.mark[r * e.mw + (q >> 5)] >> (q & 31)) & 1u)) continue;
      int best[G], bp[G];
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const int x = q * G + i;
        if (top < 0) {
          best[i] = -1;
          bp[i] = 0;
        } else {
          const int s = row_env_site(E, top, x);
          best[i] = site_h(frr, y, s) + (x - s) * (x - s);
          bp[i] = s;
        }
      }
      edt2d_store<G>(o, frr, img, H, W, y, q * G, best, bp);
    }
    __syncthreads();  // wsum, nsite and the next row's area are reused
  }
}

template <int G>
__device__ __forceinline__ void edt2d_rows_resolve(const RowEnvLds& e, int32_t* hs, int P,
                                                   int off, const int16_t* fr, int rows, int img,
                                                   int H, int W, int y0, const EdtOut& o) {
  __syncthreads();  // marks and flags are complete, the window reads are done
  for (int r = 0; r < rows; ++r)
    if (e.flag[r]) edt2d_resolve_row<G>(e, hs, P, off, fr, r, img, H, W, y0, o);
}

// Row pass of one band of `rows` rows starting at y0: squared column
// distances hs and feature rows fr in LDS, exact expanding-window argmin per
// pixel (the lower column wins ties), outputs written to HBM; pixels whose
// window reaches kWinCap steps are resolved by their row's envelope.
template <bool ENV>
__device__ __forceinline__ void edt2d_band_rows(int32_t* hs, const int16_t* fr, int rows,
                                                int img, int H, int W, int y0, const EdtOut& o,
                                                const RowEnvLds& e) {
  const int npx = rows * W;
  for (int p = threadIdx.x; p < npx; p += blockDim.x) {
    const int r = p / W;
    const int x = p - r * W;
    const int32_t* hr = hs + r * W;
    int best = hr[x];
    int bp = x;
    const int reach = max(x, W - 1 - x);
    bool capped = false;
    // Candidates in the order x, x-1, x+1, x-2, x+2, ...: a left candidate
    // replaces an equal best and a right one does not, so the lowest column
    // attaining the minimum wins.  Steps are taken 8 at a time: the 16 LDS
    // reads of a block are issued together, then applied in that order.  A
    // step k with k^2 > best can neither beat nor tie best (hv >= 0), so the
    // steps of a block past that point change nothing, and the next block is
    // read only while (k0 + 1)^2 <= best.
#pragma unroll 1
    for (int k0 = 0; k0 < reach; k0 += 8) {
      if (best >= 0 && (k0 + 1) * (k0 + 1) > best) break;
      if (ENV && k0 >= kWinCap) {
        capped = true;
        break;
      }
      int hl[8], hv[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + 1 + u;
        hl[u] = hr[max(x - k, 0)];
        hv[u] = hr[min(x + k, W - 1)];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int k = k0 + 1 + u;
        const int kk = k * k;
        if (x - k >= 0 && hl[u] >= 0 && (best < 0 || hl[u] + kk <= best)) {
          best = hl[u] + kk;
          bp = x - k;
        }
        if (x + k < W && hv[u] >= 0 && (best < 0 || hv[u] + kk < best)) {
          best = hv[u] + kk;
          bp = x + k;
        }
      }
    }
    if (ENV && capped) {
      row_env_mark(e, r, x);
      continue;
    }
    edt2d_store<1>(o, fr + r * W, img, H, W, y0 + r, x, &best, &bp);
  }
  if (ENV) edt2d_rows_resolve<1>(e, hs, W, 0, fr, rows, img, H, W, y0, o);
}

// Rows longer than this take the envelope fallback (ENV): shorter rows bound
// every window by the row length.
constexpr int kEnvMinW = 257;

// LDS bytes of edt2d_rows_kernel<R, ENV> and edt2d_rows4_kernel<R, ENV>.
inline size_t edt2d_rows_lds(int R, int W) {
  return align256((size_t)R * W * 6) + (W >= kEnvMinW ? row_env_lds_bytes(R, W, 1) : 0);
}
inline size_t edt2d_rows4_lds(int R, int W) {
  return align256((size_t)R * (W + 16) * 4 + (size_t)R * W * 2) +
         (W >= kEnvMinW ? row_env_lds_bytes(R, W, 4) : 0);
}

// Axis-1 pass.  One workgroup per band of R rows: the band's column features
// are loaded in one round (8 int16 per thread per round), turned into squared
// column distances in LDS, then every pixel runs an exact expanding-window
// argmin over its row (the lower column wins ties).  With unit sampling all
// quantities are exact integers, so this equals SciPy's double-precision
// envelope result bit for bit.  ENV: rows of at least kEnvMinW pixels, with
// the envelope fallback.
template <int R, bool ENV>
__global__ __launch_bounds__(256) void edt2d_rows_kernel(const int16_t* __restrict__ f0, int H,
                                                         int W, EdtOut o) {
  extern __shared__ int32_t lds[];
  int32_t* hs = lds;                       // [R][W] squared column distance, -1 = none
  int16_t* fr = (int16_t*)(lds + R * W);   // [R][W] feature row
  const RowEnvLds e = row_env_lds((char*)lds + ((R * W * 6 + 255) & ~255), R, W, 1);
  const int img = blockIdx.y;
  const int y0 = blockIdx.x * R;
  const int rows = min(R, H - y0);
  const int npx = rows * W;
  const int16_t* src = f0 + ((int64_t)img * H + y0) * W;
  const bool vec = (W % 8 == 0);
  if (ENV) row_env_clear(e, R);
  for (int base = 0; base < npx; base += 256 * 8) {
    const int p0 = base + threadIdx.x * 8;
    int16_t v[8];
    if (vec && p0 + 8 <= npx) {
      const uint4 q = *(const uint4*)(src + p0);
      v[0] = (int16_t)(q.x & 0xffff); v[1] = (int16_t)(q.x >> 16);
      v[2] = (int16_t)(q.y & 0xffff); v[3] = (int16_t)(q.y >> 16);
      v[4] = (int16_t)(q.z & 0xffff); v[5] = (int16_t)(q.z >> 16);
      v[6] = (int16_t)(q.w & 0xffff); v[7] = (int16_t)(q.w >> 16);
    } else {
#pragma unroll
      for (int i = 0; i < 8; ++i) v[i] = p0 + i < npx ? src[p0 + i] : (int16_t)-1;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int p = p0 + i;
      if (p < npx) {
        const int r = p / W;
        const int f = v[i];
        const int dy = y0 + r - f;
        fr[p] = (int16_t)f;
        hs[p] = f >= 0 ? dy * dy : -1;
      }
    }
  }
  __syncthreads();
  edt2d_band_rows<ENV>(hs, fr, rows, img, H, W, y0, o, e);
}

// Window keys of the 4-pixel row pass: column x of a band row holds
// key = (h << 12) | x, h its squared column distance, or kKeyNone for a
// column without background (and the 8 pad entries on both sides).  With H
// <= 1024 and k <= 8, every h + k^2 of a real column is below kKeyNone, and
// (kKeyNone + 64) << 12 still fits 32 bits, so the candidate of step k is
// key + (k^2 << 12) and the minimum key is the lowest column attaining the
// minimum: one add and one unsigned min per candidate.
constexpr uint32_t kKeyNone = (1u << 20) - 65u;

__device__ __forceinline__ uint32_t window_key(int f, int y, int x) {
  const int dy = y - f;
  return ((f >= 0 ? (uint32_t)(dy * dy) : kKeyNone) << 12) | (uint32_t)x;
}

// The window argmin of the four pixels x0 .. x0 + 3 of one row (W <= 4096):
// kr[-8 .. W + 7] readable, 16-byte aligned at kr - 8.  The 20 keys around
// the four pixels come as five 16-byte LDS reads and the first 8 window steps
// are key minima on registers; a pixel whose best still admits a step beyond
// 8 (none found, or best >= 81) continues with the one-pixel blocked loop of
// edt2d_band_rows (same candidate order and tie rule).  best = -1: no
// candidate.  With CAP, returns whether a pixel reached kWinCap steps (its
// best is then not final).
template <bool CAP>
__device__ __forceinline__ bool rows4_quad(const uint32_t* kr, int x0, int W, int* best,
                                           int* bp) {
  uint32_t h[20];  // kr[x0 - 8 .. x0 + 11]
#pragma unroll
  for (int v = 0; v < 5; ++v) {
    const uint4 t = *(const uint4*)(kr + x0 - 8 + 4 * v);
    h[4 * v] = t.x;
    h[4 * v + 1] = t.y;
    h[4 * v + 2] = t.z;
    h[4 * v + 3] = t.w;
  }
  bool capped = false;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    uint32_t m = h[8 + i];
#pragma unroll
    for (int k = 1; k <= 8; ++k) {
      const uint32_t add = (uint32_t)(k * k) << 12;
      m = min(m, h[8 + i - k] + add);
      m = min(m, h[8 + i + k] + add);
    }
    const uint32_t mv = m >> 12;
    best[i] = mv >= kKeyNone ? -1 : (int)mv;
    bp[i] = (int)(m & 4095u);
    const int x = x0 + i;
    const int reach = max(x, W - 1 - x);
    if (best[i] < 0 || best[i] >= 81) {
      // steps 9, 10, ...: the blocked loop of edt2d_band_rows (rolled: the
      // cap bounds its trip count, and unrolled it would quadruple the code)
#pragma unroll 1
      for (int k0 = 8; k0 < reach; k0 += 8) {
        if (best[i] >= 0 && (k0 + 1) * (k0 + 1) > best[i]) break;
        if (CAP && k0 >= kWinCap) {
          capped = true;
          break;
        }
        int hl[8], hv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int k = k0 + 1 + u;
          const uint32_t a = kr[max(x - k, 0)] >> 12, c = kr[min(x + k, W - 1)] >> 12;
          hl[u] = a >= kKeyNone ? -1 : (int)a;
          hv[u] = c >= kKeyNone ? -1 : (int)c;
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const int k = k0 + 1 + u;
          const int kk = k * k;
          if (x - k >= 0 && hl[u] >= 0 && (best[i] < 0 || hl[u] + kk <= best[i])) {
            best[i] = hl[u] + kk;
            bp[i] = x - k;
          }
          if (x + k < W && hv[u] >= 0 && (best[i] < 0 || hv[u] + kk < best[i])) {
            best[i] = hv[u] + kk;
            bp[i] = x + k;
          }
        }
      }
    }
  }
  return capped;
}

// Axis-1 pass with four adjacent pixels per lane (W % 4 == 0, W <= 4096).
// One workgroup per band of R rows.  The band's window keys sit in LDS with
// 8 kKeyNone entries on both sides of every row; every lane runs rows4_quad
// on its four pixels, so the results are edt2d_band_rows's bit for bit.
// Outputs leave as 16-byte stores: four float32 distances, and per index
// plane four int32 or two pairs of int64.  (The envelope fallback rebuilds a
// flagged row from the feature rows fr, overwriting its keys.)
template <int R, bool ENV>
__global__ __launch_bounds__(256) void edt2d_rows4_kernel(const int16_t* __restrict__ f0, int H,
                                                          int W, EdtOut o) {
  extern __shared__ int32_t lds[];
  const int P = W + 16;                 // padded row pitch of hs
  int32_t* hs = lds;                    // [R][W + 16], entry x of row r at r * P + 8 + x
  int16_t* fr = (int16_t*)(lds + R * P);  // [R][W]
  const RowEnvLds e =
      row_env_lds((char*)lds + ((R * P * 4 + R * W * 2 + 255) & ~255), R, W, 4);
  const int img = blockIdx.y;
  const int y0 = blockIdx.x * R;
  const int rows = min(R, H - y0);
  const int npx = rows * W;
  const int16_t* src = f0 + ((int64_t)img * H + y0) * W;
  if (ENV) row_env_clear(e, R);
  for (int i = threadIdx.x; i < rows * 16; i += blockDim.x) {
    const int r = i >> 4, k = i & 15;
    hs[r * P + (k < 8 ? k : W + k)] = (int32_t)(kKeyNone << 12);
  }
  for (int p0 = threadIdx.x * 8; p0 < npx; p0 += blockDim.x * 8) {
    const uint4 q = *(const uint4*)(src + p0);  // 8 features, W % 8 == 0 keeps rows whole
    const int r = R == 1 ? 0 : p0 / W;
    const int x = p0 - r * W;
    const uint32_t w4[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int f = (int16_t)((w4[i >> 1] >> (16 * (i & 1))) & 0xffffu);
      fr[r * W + x + i] = (int16_t)f;
      hs[r * P + 8 + x + i] = (int32_t)window_key(f, y0 + r, x + i);
    }
  }
  __syncthreads();
  for (int qd = threadIdx.x; qd * 4 < npx; qd += blockDim.x) {
    const int p0 = qd * 4;
    const int r = R == 1 ? 0 : p0 / W;
    const int x0 = p0 - r * W;
    const uint32_t* kr = (const uint32_t*)(hs + r * P + 8);  // kr[-8 .. W + 7] readable
    int best[4], bp[4];
    const bool capped = rows4_quad<ENV>(kr, x0, W, best, bp);
    if (ENV && capped) {
      row_env_mark(e, r, x0 >> 2);
      continue;
    }
    edt2d_store<4>(o, fr + r * W, img, H, W, y0 + r, x0, best, bp);
  }
  if (ENV) edt2d_rows_resolve<4>(e, hs, P, 8, fr, rows, img, H, W, y0, o);
}

// Summary path (unit sampling, H <= 1024, W % 8 == 0).  Column pass: a
// workgroup covers CG columns x (256 / CG) segments of seglen <= 32 rows;
// each thread loads its segment's rows of one column, keeps them as a
// background bit mask, the segments exchange first/last background rows
// through LDS, and each thread writes one summary (colsum_pack) to
// sm[image][segment][column] -- 8 bytes per 16 or 32 pixels instead of one
// int16 feature row per pixel.
template <typename TIn, int NCH, int CG>
__global__ __launch_bounds__(256) void edt2d_colsum_kernel(const TIn* __restrict__ in, int H,
                                                           int W, int seglen,
                                                           uint64_t* __restrict__ sm) {
  constexpr int SEGS = 256 / CG;
  __shared__ int16_t sfirst[SEGS][CG];
  __shared__ int16_t slast[SEGS][CG];
  const int cg = threadIdx.x & (CG - 1);
  const int seg = threadIdx.x / CG;
  const int ncb = (W + CG - 1) / CG;
  const int lb = xcd_swizzle(blockIdx.x, gridDim.x);  // an image's column blocks share an XCD
  const int img = lb / ncb;
  const int xb = (lb - img * ncb) * CG + cg;
  const int ys = seg * seglen;
  const int nr = min(seglen, H - ys);
  const bool live = xb < W && nr > 0;
  TIn raw[NCH * 16];
  if (live) {
    const TIn* col = in + ((int64_t)img * H + ys) * W + xb;
#pragma unroll
    for (int b = 0; b < NCH * 16; ++b) raw[b] = col[(int64_t)min(b, nr - 1) * W];
  }
  uint32_t m = 0;
  if (live) {
#pragma unroll
    for (int b = 0; b < NCH * 16; ++b) m |= (uint32_t)(b < nr && raw[b] == TIn(0)) << b;
  }
  sfirst[seg][cg] = m ? (int16_t)(ys + __builtin_ctz(m)) : (int16_t)-1;
  slast[seg][cg] = m ? (int16_t)(ys + 31 - __builtin_clz(m)) : (int16_t)-1;
  __syncthreads();
  if (!live) return;
  int before = -1, after = -1;
  for (int s = seg - 1; s >= 0 && before < 0; --s) before = slast[s][cg];
  for (int s = seg + 1; s < SEGS && after < 0; ++s) after = sfirst[s][cg];
  sm[((int64_t)img * SEGS + seg) * W + xb] = colsum_pack(m, before, after);
}

// Summary path, row pass: one workgroup per row (min(256, W / 4) lanes,
// rounded up to whole waves), a 1-D grid of batch * H rows whose launch
// order is remapped so that every XCD takes a contiguous range of rows: the
// rows of one segment read the same summaries, which one L2 then fetches
// once instead of once per XCD.  Every lane resolves four columns of the row
// from their summaries (two 16-byte loads), writes their window keys and
// feature rows to LDS, then runs rows4_quad on four pixels -- the results
// of edt2d_rows4_kernel bit for bit.  ENV: rows of at least kEnvMinW pixels,
// with the envelope fallback.
template <bool ENV>
__global__ __launch_bounds__(256) void edt2d_rowsum_kernel(const uint64_t* __restrict__ sm,
                                                           int H, int W, int segs, int seglen,
                                                           EdtOut o) {
  extern __shared__ int32_t lds[];
  const int P = W + 16;
  uint32_t* kb = (uint32_t*)lds;           // [W + 16] window keys, entry x at 8 + x
  int16_t* fr = (int16_t*)(lds + P);       // [W] feature rows
  const RowEnvLds e = row_env_lds((char*)lds + ((P * 4 + W * 2 + 255) & ~255), 1, W, 4);
  const int lb = xcd_swizzle(blockIdx.x, gridDim.x);
  const int img = lb / H;
  const int y = lb - img * H;
  const int sg = y / seglen;
  const uint64_t* srow = sm + ((int64_t)img * segs + sg) * W;
  if (ENV) row_env_clear(e, 1);
  if (threadIdx.x < 16) kb[threadIdx.x < 8 ? threadIdx.x : W + threadIdx.x] = kKeyNone << 12;
  for (int x = threadIdx.x * 4; x < W; x += blockDim.x * 4) {
    const ulonglong2 q0 = *(const ulonglong2*)(srow + x);
    const ulonglong2 q1 = *(const ulonglong2*)(srow + x + 2);
    const uint64_t qq[4] = {q0.x, q0.y, q1.x, q1.y};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int f = colsum_nearest(qq[i], y, sg * seglen);
      fr[x + i] = (int16_t)f;
      kb[8 + x + i] = window_key(f, y, x + i);
    }
  }
  __syncthreads();
  for (int x0 = threadIdx.x * 4; x0 < W; x0 += blockDim.x * 4) {
    int best[4], bp[4];
    const bool capped = rows4_quad<ENV>(kb + 8, x0, W, best, bp);
    if (ENV && capped) {
      row_env_mark(e, 0, x0 >> 2);
      continue;
    }
    edt2d_store<4>(o, fr, img, H, W, y, x0, best, bp);
  }
  if (ENV) edt2d_rows_resolve<4>(e, (int32_t*)kb, P, 8, fr, 1, img, H, W, y, o);
}

// --------------------------------------------------------------- N-D --
// Axis 0 from the mask, one lane per line (tm_dt_lines.h).
template <typename TIn, bool FINAL>
__global__ __launch_bounds__(256) void edt_pass0_kernel(const TIn* __restrict__ in,
                                                        int32_t* __restrict__ f0, Shape sh,
                                                        int64_t nlines, Samp sp, EdtOut o) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nlines) return;
  edt_line_axis0<TIn, FINAL>(in, f0, sh, t, sp, o);
}

// Axis d >= 1, one lane per line (tm_dt_fast.h, loads PF voxels ahead).
// COMPACT (items of up to 3 axes): 4/6-byte stack entries
// (edt_line_stack_c), in LDS (LDS_DEPTH > 0, requires L <= LDS_DEPTH) or in
// the blocked global stacks from gpos on; otherwise site + double partial
// sum, in LDS or in the interleaved global scratch gpos/gh (entry k of line
// t at k * nlines + t, coalesced across lanes).
template <bool FINAL, int PF, int LDS_DEPTH, int BLOCK, bool COMPACT>
__global__ __launch_bounds__(BLOCK) void edt_stack_kernel(
    const int32_t* __restrict__ fin, int32_t* __restrict__ fout, int64_t plane, Shape sh, int d,
    int64_t nlines, Samp sp, int32_t* __restrict__ gpos, double* __restrict__ gh, EdtOut o) {
  constexpr int NE = LDS_DEPTH > 0 ? LDS_DEPTH * BLOCK : 1;
  __shared__ uint32_t lA[COMPACT ? NE : 1];
  __shared__ uint16_t lB[NE];
  __shared__ double lh[COMPACT ? 1 : NE];
  const int64_t t = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  if (t >= nlines) return;
  if (COMPACT) {
    constexpr bool BLK = LDS_DEPTH == 0;
    StkAddr<BLK> A;
    if (BLK) {
      const int64_t L = sh.size[d];
      A.E = rows_stack_e(gpos);
      A.F1 = rows_stack_f1(gpos, L, nlines);
      A.ss = 0;
      A.nl8 = nlines * 8;
      A.t8 = t * 8;
    } else {
      A.E = lA + threadIdx.x;
      A.F1 = lB + threadIdx.x;
      A.ss = BLOCK;
      A.nl8 = A.t8 = 0;
    }
    if (d == 1) edt_line_stack_c<FINAL, 1, PF, BLK>(fin, fout, plane, sh, t, sp, A, o);
    else edt_line_stack_c<FINAL, 2, PF, BLK>(fin, fout, plane, sh, t, sp, A, o);
  } else if (LDS_DEPTH > 0) {
    edt_line_stack_fast<FINAL, PF>(fin, fout, plane, sh, d, t, sp, lB + threadIdx.x,
                                   lh + threadIdx.x, BLOCK, o);
  } else {
    edt_line_stack_fast<FINAL, PF>(fin, fout, plane, sh, d, t, sp, gpos + t, gh + t, nlines, o);
  }
}

// Last pass (axis ndim - 1, the contiguous one) of 2-D / 3-D items without
// indices, row-staged through LDS (tm_dt_chunk.h): 256 rows per block.
template <int NP, int CH>
__global__ __launch_bounds__(kCLanes) void edt_rows_kernel(const int32_t* __restrict__ fin,
                                                           int64_t plane, Shape sh,
                                                           int64_t nlines, Samp sp,
                                                           uint32_t* __restrict__ E,
                                                           uint16_t* __restrict__ F1,
                                                           float* __restrict__ dist) {
  extern __shared__ int32_t lds_rows[];
  const int64_t row0 = (int64_t)xcd_swizzle(blockIdx.x, gridDim.x) * kCLanes;
  const int nrows = (int)(nlines - row0 < kCLanes ? nlines - row0 : kCLanes);
  edt_rows_final<NP, CH>(fin, plane, sh, row0, nrows, sp, E, F1, nlines, dist, lds_rows);
}

// Bytes of the blocked envelope stacks (tm_dt_fast.h) of the passes d >= 1
// of items of 2 or 3 axes, the largest over the passes; 0 otherwise.
static int64_t compact_stack_bytes(int64_t batch, int ndim, const int64_t* shape) {
  if (ndim != 2 && ndim != 3) return 0;
  int64_t n = batch, m = 0;
  for (int j = 0; j < ndim; ++j) n *= shape[j];
  for (int d = 1; d < ndim; ++d) {
    const int64_t b = rows_stack_bytes(shape[d], n / shape[d]);
    m = b > m ? b : m;
  }
  return m;
}

static bool unit_sampling(int ndim, const double* s) {
  if (!s) return true;
  for (int j = 0; j < ndim; ++j)
    if (s[j] != 1.0) return false;
  return true;
}

static bool use_2d_path(int ndim, const int64_t* shape, const double* s) {
  return ndim == 2 && unit_sampling(ndim, s) && shape[0] <= kColSegs * 4 * kChunkRows &&
         shape[1] <= 4096;
}

template <typename TIn>
static int edt_dispatch(const TIn* in, int64_t batch, int ndim, const int64_t* shape,
                        const double* sampling, EdtOut o, void* ws, size_t ws_bytes,
                        hipStream_t st) {
  Shape sh = make_shape(ndim, shape);
  Samp sp;
  for (int j = 0; j < kMaxNdim; ++j) sp.s[j] = (sampling && j < ndim) ? sampling[j] : 1.0;
  o.plane = batch * sh.numel;

  if (use_2d_path(ndim, shape, sampling)) {
    const int H = (int)shape[0], W = (int)shape[1];
    const size_t need = (size_t)o.plane * sizeof(int16_t);
    if (!ws || ws_bytes < need) {
      set_error("tm_edt: workspace of %zu bytes is smaller than the %zu required", ws_bytes, need);
      return TM_ERR_WORKSPACE;
    }
    if (((uintptr_t)ws % 16) != 0) {
      set_error("tm_edt: workspace must be 16-byte aligned");
      return TM_ERR_ARG;
    }
    int16_t* f0 = (int16_t*)ws;
    const int seglen = (H + kColSegs - 1) / kColSegs;
    const int aligned = ((uintptr_t)in % 16) == 0;
    const bool vecw = sizeof(TIn) == 1 && aligned && W % 4 == 0;
    // variant v < 8: columns per thread = (v & 1) ? 4 : 1, rows per band =
    // 1 << (v >> 1); v = 8, 9, 10: one column per thread and the four-pixel
    // row pass (edt2d_rows4_kernel) with bands of 1, 2, 4 rows; v = 11, 12:
    // bands of 1 row, 8 or 4 columns per column-pass workgroup (more
    // workgroups for narrow images); v = 13, 14, 15: bands of 1 row, four
    // adjacent columns per column-pass thread (one 4-byte load per row, one
    // 8-byte store per row), 16, 8 or 4 column groups per workgroup; v = 16,
    // 17: bands of 1 row, the column pass writes one summary per (column,
    // segment of at most 32 rows) instead of one feature row per pixel, 16 or
    // 8 column groups per workgroup
    auto run = [&](int v) -> int {
      if (v >= 8) {
        const int R = v <= 10 ? 1 << (v - 8) : 1;  // 1, 2, 4 rows per band
        const int CPT = (v >= 13 && v <= 15) ? 4 : 1;
        const int CG = (v <= 10 || v == 13 || v == 16) ? 16 : ((v == 11 || v == 14 || v == 17) ? 8 : 4);
        const bool sum = v >= 16;
        const int segs = 256 / CG, seglen = (H + segs - 1) / segs;
        if (sum && (seglen > 32 || (size_t)segs * 8 > (size_t)H * 2)) {
          set_error("tm_edt: column summaries do not apply to this height");
          return TM_ERR_ARG;
        }
        const size_t lds = edt2d_rows4_lds(R, W);
        auto a16 = [](const void* p) { return ((uintptr_t)p % 16) == 0; };
        if (W % 8 != 0 || lds > 96 * 1024 || batch > 65535 || !a16(o.dist) || !a16(o.idx)) {
          set_error("tm_edt: four-pixel row variant does not apply");
          return TM_ERR_ARG;
        }
        const int64_t ncb = (W + CG * CPT - 1) / (CG * CPT);
        const dim3 cgrid((unsigned)(ncb * batch));
        const int seg = (H + 256 / CG - 1) / (256 / CG);
        if (seg > 4 * kChunkRows) {
          set_error("tm_edt: column segments too long for this variant");
          return TM_ERR_ARG;
        }
#define TM_COLS1(NCH, G)                                                                    \
  if (vecw)                                                                                 \
    hipLaunchKernelGGL((edt2d_cols_kernel<TIn, NCH, 1, true, G>), cgrid, dim3(256), 0, st, in, \
                       H, W, f0, aligned);                                                  \
  else                                                                                      \
    hipLaunchKernelGGL((edt2d_cols_kernel<TIn, NCH, 1, false, G>), cgrid, dim3(256), 0, st, \
                       in, H, W, f0, aligned)
#define TM_COLS4(NCH, G)                                                                    \
  if (vecw)                                                                                 \
    hipLaunchKernelGGL((edt2d_cols_kernel<TIn, NCH, 4, true, G>), cgrid, dim3(256), 0, st, in, \
                       H, W, f0, aligned);                                                  \
  else                                                                                      \
    hipLaunchKernelGGL((edt2d_cols_kernel<TIn, NCH, 4, false, G>), cgrid, dim3(256), 0, st, \
                       in, H, W, f0, aligned)
#define TM_COLS_CG(M, G)                     \
  if (seg <= 16) M(1, G);                    \
  else if (seg <= 32) M(2, G);               \
  else M(4, G)
#define TM_COLSS(NCH, G)                                                                \
  hipLaunchKernelGGL((edt2d_colsum_kernel<TIn, NCH, G>), cgrid, dim3(256), 0, st, in, H, W, \
                     seglen, sm)
        if (sum) {
          uint64_t* sm = (uint64_t*)f0;
          if (CG == 16) {
            if (seg <= 16) TM_COLSS(1, 16);
            else TM_COLSS(2, 16);
          } else {
            if (seg <= 16) TM_COLSS(1, 8);
            else TM_COLSS(2, 8);
          }
        } else if (CPT == 1) {
          if (CG == 16) { TM_COLS_CG(TM_COLS1, 16); }
          else if (CG == 8) { TM_COLS_CG(TM_COLS1, 8); }
          else { TM_COLS_CG(TM_COLS1, 4); }
        } else {
          if (CG == 16) { TM_COLS_CG(TM_COLS4, 16); }
          else if (CG == 8) { TM_COLS_CG(TM_COLS4, 8); }
          else { TM_COLS_CG(TM_COLS4, 4); }
        }
#undef TM_COLSS
#undef TM_COLS_CG
#undef TM_COLS4
#undef TM_COLS1
        TM_LAUNCH_CHECK();
        const int threads = (int)std::min<int64_t>(256, ((int64_t)R * W / 4 + 63) / 64 * 64);
        const dim3 rgrid((H + R - 1) / R, (unsigned)batch);
#define TM_ROWS4(RR, E) \
  hipLaunchKernelGGL((edt2d_rows4_kernel<RR, E>), rgrid, dim3(threads), lds, st, f0, H, W, o)
        const bool env = W >= kEnvMinW;
        if (sum) {
          const uint64_t* sm = (const uint64_t*)f0;
          if ((int64_t)H * batch >= ((int64_t)1 << 31)) {
            set_error("tm_edt: too many rows for one launch of this variant");
            return TM_ERR_ARG;
          }
          const dim3 sgrid((unsigned)(H * batch));
          if (env)
            hipLaunchKernelGGL((edt2d_rowsum_kernel<true>), sgrid, dim3(threads), lds, st, sm, H,
                               W, segs, seglen, o);
          else
            hipLaunchKernelGGL((edt2d_rowsum_kernel<false>), sgrid, dim3(threads), lds, st, sm, H,
                               W, segs, seglen, o);
          TM_LAUNCH_CHECK();
          return TM_OK;
        }
        switch (R) {
          case 1: if (env) TM_ROWS4(1, true); else TM_ROWS4(1, false); break;
          case 2: if (env) TM_ROWS4(2, true); else TM_ROWS4(2, false); break;
          default: if (env) TM_ROWS4(4, true); else TM_ROWS4(4, false); break;
        }
#undef TM_ROWS4
        TM_LAUNCH_CHECK();
        return TM_OK;
      }
      const int cpt = (v & 1) ? 4 : 1;
      const int R = 1 << (v >> 1);
      if (edt2d_rows_lds(R, W) > 112 * 1024) {
        set_error("tm_edt: band does not fit LDS");
        return TM_ERR_ARG;
      }
      const int64_t ncb = (W + 16 * cpt - 1) / (16 * cpt);
      if (ncb * batch >= ((int64_t)1 << 31)) {
        set_error("tm_edt: batch too large for one launch");
        return TM_ERR_ARG;
      }
      const dim3 cgrid((unsigned)(ncb * batch));
#define TM_COLS(NCH, CPT)                                                                   \
  if (vecw)                                                                                 \
    hipLaunchKernelGGL((edt2d_cols_kernel<TIn, NCH, CPT, true>), cgrid, dim3(256), 0, st, in, \
                       H, W, f0, aligned);                                                  \
  else                                                                                      \
    hipLaunchKernelGGL((edt2d_cols_kernel<TIn, NCH, CPT, false>), cgrid, dim3(256), 0, st, in, \
                       H, W, f0, aligned)
      if (cpt == 4) {
        if (seglen <= 16) TM_COLS(1, 4);
        else if (seglen <= 32) TM_COLS(2, 4);
        else TM_COLS(4, 4);
      } else {
        if (seglen <= 16) TM_COLS(1, 1);
        else if (seglen <= 32) TM_COLS(2, 1);
        else TM_COLS(4, 1);
      }
#undef TM_COLS
      TM_LAUNCH_CHECK();
      const dim3 rgrid((H + R - 1) / R, (unsigned)batch);
      const size_t lds = edt2d_rows_lds(R, W);
#define TM_ROWS(RR, E) \
  hipLaunchKernelGGL((edt2d_rows_kernel<RR, E>), rgrid, dim3(256), lds, st, f0, H, W, o)
      const bool env = W >= kEnvMinW;
      switch (R) {
        case 1: if (env) TM_ROWS(1, true); else TM_ROWS(1, false); break;
        case 2: if (env) TM_ROWS(2, true); else TM_ROWS(2, false); break;
        case 4: if (env) TM_ROWS(4, true); else TM_ROWS(4, false); break;
        default: if (env) TM_ROWS(8, true); else TM_ROWS(8, false); break;
      }
#undef TM_ROWS
      TM_LAUNCH_CHECK();
      return TM_OK;
    };
    auto p2 = [](int64_t v) { int64_t p = 1; while (p < v) p <<= 1; return p; };
    const int64_t key[4] = {p2(H), p2(W), o.idx ? (o.idx64 ? 64 : 32) : 0, o.dist ? 1 : 0};
    const int dflt = W >= 512 ? 1 : 3;  // 4 columns/thread; 1 or 2 rows per band
    bool ran = false;
    const int v = tune_pick("edt2d", key, 4, 18, dflt, run, st, &ran);
    if (ran) return TM_OK;
    const int rc = run(v);
    // a cached choice that does not apply here (imported from another
    // device, or an LDS limit) falls back to the default variant
    if (rc == TM_ERR_ARG && v != dflt) return run(dflt);
    return rc;
  }

  const int64_t N = o.plane;
  const size_t need = tm_edt_workspace_bytes(batch, ndim, shape, sampling);
  if (ws_bytes < need || (need && !ws)) {
    set_error("tm_edt: workspace of %zu bytes is smaller than the %zu required", ws_bytes, need);
    return TM_ERR_WORKSPACE;
  }
  const int nbuf = ndim > 1 ? ndim - 1 : 1;
  int32_t* bufA = (int32_t*)ws;
  int32_t* bufB = bufA + (size_t)nbuf * N;
  int32_t* gpos = bufB + (size_t)nbuf * N;
  double* gh = (double*)(((uintptr_t)(gpos + N) + 255) & ~(uintptr_t)255);

  const bool rows_ok = !o.idx && (ndim == 2 || ndim == 3);
  // launch variant v (one choice for the whole call):
  //   v & 3   where each envelope stack lives and the workgroup size:
  //           0: LDS (128 deep, 80 KB per 64 lanes) when the line fits, else
  //              global with 64 lanes; 1: global, 64 lanes; 2: global, 256
  //              lanes; 3: LDS 64 deep when the line fits, else global, 256
  //   v >> 2 & 1  loads issued 8 (0) or 16 (1) voxels ahead
  //   v >> 3  1: the last pass of a 2-D / 3-D item without indices runs
  //           row-staged (edt_rows_kernel)
  auto run = [&](int v) -> int {
    const int w = v & 3;
    const bool pf16 = ((v >> 2) & 1) != 0;
    const bool rows = (v >> 3) != 0 && rows_ok;
    {
      const int64_t nlines = batch * sh.stride[0];
      const dim3 g(grid_for(nlines, 256)), blk(256);
      if (ndim == 1)
        hipLaunchKernelGGL((edt_pass0_kernel<TIn, true>), g, blk, 0, st, in, bufA, sh, nlines, sp,
                           o);
      else
        hipLaunchKernelGGL((edt_pass0_kernel<TIn, false>), g, blk, 0, st, in, bufA, sh, nlines,
                           sp, o);
      TM_LAUNCH_CHECK();
    }
    int32_t* fin = bufA;
    int32_t* fout = bufB;
    for (int d = 1; d < ndim; ++d) {
      const int64_t nlines = batch * (sh.numel / sh.size[d]);
      const int64_t L = sh.size[d];
      const bool fin_pass = (d == ndim - 1);
      int depth = 0, block = 256;
      if (w == 0) {
        depth = L <= 128 ? 128 : 0;
        block = 64;
      } else if (w == 1) {
        block = 64;
      } else if (w == 3 && L <= 64) {
        depth = 64;
        block = 64;
      }
      if (fin_pass && rows) {
        const dim3 g(grid_for(nlines, kCLanes)), blk(kCLanes);
        uint32_t* E = rows_stack_e(gpos);
        uint16_t* F1 = rows_stack_f1(gpos, L, nlines);
        if (ndim == 2)
          hipLaunchKernelGGL((edt_rows_kernel<1, 32>), g, blk, (edt_rows_lds<1, 32>()), st, fin,
                             N, sh, nlines, sp, E, F1, o.dist);
        else
          hipLaunchKernelGGL((edt_rows_kernel<2, 16>), g, blk, (edt_rows_lds<2, 16>()), st, fin,
                             N, sh, nlines, sp, E, F1, o.dist);
        TM_LAUNCH_CHECK();
        break;
      }
      const dim3 g(grid_for(nlines, block)), blk(block);
#define TM_EDT_STACK(FINAL, PF, DEPTH, B)                                                      \
  if (ndim <= 3)                                                                              \
    hipLaunchKernelGGL((edt_stack_kernel<FINAL, PF, DEPTH, B, true>), g, blk, 0, st, fin, fout, \
                       N, sh, d, nlines, sp, gpos, gh, o);                                    \
  else                                                                                        \
    hipLaunchKernelGGL((edt_stack_kernel<FINAL, PF, DEPTH, B, false>), g, blk, 0, st, fin,      \
                       fout, N, sh, d, nlines, sp, gpos, gh, o)
#define TM_EDT_STACK_ANY(FINAL, PF)                 \
  if (depth == 128) {                              \
    TM_EDT_STACK(FINAL, PF, 128, 64);              \
  } else if (depth == 64) {                        \
    TM_EDT_STACK(FINAL, PF, 64, 64);               \
  } else if (block == 64) {                        \
    TM_EDT_STACK(FINAL, PF, 0, 64);                \
  } else {                                         \
    TM_EDT_STACK(FINAL, PF, 0, 256);               \
  }
      if (fin_pass) {
        if (pf16) {
          TM_EDT_STACK_ANY(true, 16)
        } else {
          TM_EDT_STACK_ANY(true, 8)
        }
      } else {
        if (pf16) {
          TM_EDT_STACK_ANY(false, 16)
        } else {
          TM_EDT_STACK_ANY(false, 8)
        }
      }
#undef TM_EDT_STACK_ANY
#undef TM_EDT_STACK
      TM_LAUNCH_CHECK();
      int32_t* tmp = fin;
      fin = fout;
      fout = tmp;
    }
    return TM_OK;
  };
  auto p2 = [](int64_t v) { int64_t q = 1; while (q < v) q <<= 1; return q; };
  int64_t key[kMaxNdim + 3];
  key[0] = ndim;
  key[1] = p2(batch);
  key[2] = o.idx ? 1 : 0;
  for (int j = 0; j < ndim; ++j) key[3 + j] = p2(sh.size[j]);
  bool ran = false;
  const int v = tune_pick("edtnd", key, 3 + ndim, 16, 8, run, st, &ran);
  if (ran) return TM_OK;
  return run(v);
}

}  // namespace tmh

extern "C" size_t tm_edt_workspace_bytes(int64_t batch, int ndim, const int64_t* shape,
                                         const double* sampling) {
  if (ndim < 1 || ndim > tmh::kMaxNdim || batch < 1) return 0;
  int64_t n = batch;
  for (int j = 0; j < ndim; ++j) n *= shape[j];
  if (tmh::use_2d_path(ndim, shape, sampling)) return (size_t)n * sizeof(int16_t);
  const int nbuf = ndim > 1 ? ndim - 1 : 1;
  // feature planes A and B, then the envelope stacks: interleaved positions
  // and partial sums (4 or more axes) or the blocked compact stacks
  const size_t b = (size_t)n * 4 * (2 * nbuf + 1) + 256 + (size_t)n * 8;
  return std::max(b, (size_t)n * 8 * nbuf + (size_t)tmh::compact_stack_bytes(batch, ndim, shape));
}

extern "C" int tm_edt(const void* input, int input_dtype, int64_t batch, int ndim,
                      const int64_t* shape, const double* sampling, float* distances,
                      void* indices, int index_dtype, void* workspace, size_t workspace_bytes,
                      void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(input != nullptr, "tm_edt: input is NULL");
  TM_CHECK_ARG(ndim >= 1 && ndim <= kMaxNdim, "tm_edt: ndim must be in 1..8, got %d", ndim);
  TM_CHECK_ARG(batch >= 1, "tm_edt: batch must be >= 1");
  for (int j = 0; j < ndim; ++j) {
    TM_CHECK_ARG(shape[j] >= 1 && shape[j] <= 65535, "tm_edt: axis %d size %lld out of range 1..65535",
                 j, (long long)shape[j]);
    if (sampling)
      TM_CHECK_ARG(sampling[j] > 0.0 && sampling[j] < 1e300, "tm_edt: sampling must be finite > 0");
  }
  TM_CHECK_ARG(distances || indices, "tm_edt: at least one output must be requested");
  TM_CHECK_ARG(!indices || index_dtype == TM_I32 || index_dtype == TM_I64,
               "tm_edt: index dtype must be int32 or int64");
  EdtOut o{distances, indices, index_dtype == TM_I64 ? 1 : 0, 0};
  hipStream_t st = (hipStream_t)stream;
  if (input_dtype == TM_U8)
    return edt_dispatch((const uint8_t*)input, batch, ndim, shape, sampling, o, workspace,
                        workspace_bytes, st);
  if (input_dtype == TM_F32)
    return edt_dispatch((const float*)input, batch, ndim, shape, sampling, o, workspace,
                        workspace_bytes, st);
  set_error("tm_edt: input dtype must be uint8/bool or float32");
  return TM_ERR_ARG;
}

// ------------------------------------------------------------- no-alloc ---
// tm_edt_noalloc runs tm_edt with its workspace in a file-scope device array
// instead of caller memory, for callers that cannot allocate (code under
// stream capture, tight serving loops).  Each GPU the library runs on holds
// its own copy of the array until the process ends; every call on a GPU
// shares it, so calls that overlap on one GPU race.  No kernel of tm_edt
// reads workspace it has not written in the same call, so nothing is
// cleared between calls.

namespace tmh {

constexpr size_t kEdtScratchBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kEdtScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

// The device address of `scratch` on the current device, looked up once per
// device (host code must never pass the array's name to a kernel: on the
// host it is only a stand-in for the symbol).
static int edt_scratch(void** out) {
  int dev = 0;
  TM_CHECK_HIP(hipGetDevice(&dev));
  TM_CHECK_ARG(dev >= 0 && dev < kMaxDevices, "tm_edt_noalloc: device %d beyond %d", dev,
               kMaxDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    TM_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TM_CHECK_ARG(p != nullptr, "tm_edt_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return TM_OK;
}

}  // namespace tmh

extern "C" int tm_edt_noalloc(const void* input, int input_dtype, int64_t batch, int ndim,
                              const int64_t* shape, const double* sampling, float* distances,
                              void* indices, int index_dtype, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(ndim >= 1 && ndim <= kMaxNdim, "tm_edt_noalloc: ndim must be in 1..8, got %d",
               ndim);
  TM_CHECK_ARG(batch >= 1 && shape, "tm_edt_noalloc: batch must be >= 1 and shape given");
  const size_t need = tm_edt_workspace_bytes(batch, ndim, shape, sampling);
  TM_CHECK_ARG(need <= kEdtScratchBytes,
               "tm_edt_noalloc: this problem needs %zu bytes of scratch, above the %zu-byte "
               "maximum; use tm_edt with a workspace",
               need, kEdtScratchBytes);
  void* ws = nullptr;
  if (need) {
    const int s = edt_scratch(&ws);
    if (s != TM_OK) return s;
  }
  TuneSuspend no_timing;  // cached or default launch variant: no events, no extra launches
  return tm_edt(input, input_dtype, batch, ndim, shape, sampling, distances, indices,
                index_dtype, ws, need, stream);
}
