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

// LDS bytes of edt_rows_final<NP, CH>: the chunk staging, NP planes in, then
// the outputs in the same space.
template <int NP, int CH>
constexpr size_t edt_rows_lds() {
  return (size_t)NP * kCLanes * (CH + 1) * sizeof(int32_t);
}

// EDT, last pass (d = ndim - 1, NP = d feature planes, 1 or 2), distances
// only.  fin: the item's feature planes (plane stride `plane`); rows
// row0 .. row0 + nrows - 1 of the nl rows whose stacks E / F1 hold (blocked
// layout, rows_stack_bytes).  lds: edt_rows_lds<NP, CH>() bytes.  The
// envelope and its walk are those of edt_line_stack_c (tm_dt_fast.h).
template <int NP, int CH>
__device__ void edt_rows_final(const int32_t* __restrict__ fin, int64_t plane, const Shape& sh,
                               int64_t row0, int nrows, const Samp& sp, uint32_t* __restrict__ E,
                               uint16_t* __restrict__ F1, int64_t nl,
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
  Envelope<NP, true> env;
  env.A.E = E;
  env.A.F1 = F1;
  env.A.ss = 0;
  env.A.nl8 = nl * 8;
  env.A.t8 = t * 8;
  env.sd = sp.s[d];
  env.s0 = sp.s[0];
  env.s1 = sp.s[1];
  env.c0 = (int32_t)c[0];
  env.c1 = (int32_t)c[1];
  int32_t* in0 = lds;
  int32_t* in1 = lds + kCLanes * RS;
  const int nch = (L + CH - 1) / CH;

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
        env.add(i, f0, f1, edt_partial<NP>(f0, f1, env.c0, env.c1, env.s0, env.s1));
      }
    }
  }
  if (active) env.close();

  __syncthreads();  // the staging area holds outputs from here on
  int32_t* ob = lds;
  EnvWalk<NP, true> walk;
  if (active && env.top >= 0) walk.start(env);
  int32_t fnone[kMaxNdim];  // no site on the line: feature (-1, 0, ..., 0)
  for (int j = 0; j < kMaxNdim; ++j) fnone[j] = 0;
  fnone[0] = -1;
  for (int q = 0; q < nch; ++q) {
    if (active) {
      for (int k = 0; k < CH; ++k) {
        const int i = q * CH + k;
        if (i >= L) break;
        float out;
        if (env.top < 0) {
          c[d] = i;
          out = edt_dist_of(sh.ndim, c, fnone, sp);
        } else {
          out = (float)sqrt(walk.next(env, i));
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
