// Lane-per-line passes along strided axes with batched loads.  First pass
// (axis 0) of the EDT and of the chamfer transforms: for every voxel the
// nearest background at or before it and at or after it; and the taxicab
// pass along later strided axes (cdt_taxicab_chunked, at the end).
//
// Two sweeps over chunks of 32 voxels.  The forward sweep reads the mask once
// (a chunk's 32 loads in flight together) and keeps, per chunk, its background
// bitmask and the last background before it as two words of a chunk-major
// scratch (chunk k of line t at k * nl + t, so the lanes of a wave touch
// consecutive words).  The backward sweep resolves every voxel from those two
// words with bit scans and writes it once.  Per voxel that is one mask read
// and one result write, against a read, a write, a second mask read and a
// read-modify-write of the result for a forward/backward pair of sweeps.
//
// The results are those of edt_line_axis0 / cdt_line_axis0 (tm_dt_lines.h):
// the same nearest sites, the same tie rule (the earlier site when both are
// as near) and the same stores.  The lanes of a workgroup walk lines of one
// item (one length), so their loads and stores stay coalesced.
#pragma once
#include "tm_dt_lines.h"

namespace tmh {

constexpr int kA0Chunk = 32;

// Bytes of the chunk words of nl lines of length L.
__host__ __device__ inline int64_t axis0_words_bytes(int64_t L, int64_t nl) {
  return (L + kA0Chunk - 1) / kA0Chunk * nl * 8;
}

// Calls emit(i, offset, prev, next) for i = L - 1 down to 0, prev / next the
// nearest background at or before / at or after i (-1: none).
template <typename TIn, typename Emit>
__device__ __forceinline__ void axis0_nearest(const TIn* __restrict__ in, int64_t base,
                                              int64_t step, int L, int64_t t, int64_t nl,
                                              uint32_t* __restrict__ words, Emit&& emit) {
  const int nch = (L + kA0Chunk - 1) / kA0Chunk;
  uint32_t* bmw = words + t;                               // background bitmask of a chunk
  int32_t* pvw = (int32_t*)words + (int64_t)nch * nl + t;  // last background before it
  int32_t last = -1;
  for (int c = 0; c < nch; ++c) {
    const int i0 = c * kA0Chunk;
    const int n = L - i0 < kA0Chunk ? L - i0 : kA0Chunk;
    const TIn* p = in + base + (int64_t)i0 * step;
    // every load in range (past the line's end: its last voxel), issued
    // together, the bits formed afterwards
    TIn m[kA0Chunk];
#pragma unroll
    for (int k = 0; k < kA0Chunk; ++k) m[k] = p[(int64_t)(k < n ? k : n - 1) * step];
    uint32_t bm = 0;
#pragma unroll
    for (int k = 0; k < kA0Chunk; ++k) bm |= (k < n && m[k] == TIn(0) ? 1u : 0u) << k;
    bmw[(int64_t)c * nl] = bm;
    pvw[(int64_t)c * nl] = last;
    if (bm) last = i0 + 31 - __clz(bm);
  }
  int32_t nx = -1;
  uint32_t bm = bmw[(int64_t)(nch - 1) * nl];
  int32_t pv = pvw[(int64_t)(nch - 1) * nl];
  for (int c = nch - 1; c >= 0; --c) {
    const int i0 = c * kA0Chunk;
    const int hi = L - 1 - i0 < kA0Chunk - 1 ? L - 1 - i0 : kA0Chunk - 1;
    uint32_t bmn = 0;
    int32_t pvn = 0;
    if (c > 0) {  // the next chunk's words, in flight meanwhile
      bmn = bmw[(int64_t)(c - 1) * nl];
      pvn = pvw[(int64_t)(c - 1) * nl];
    }
#pragma unroll
    for (int k = kA0Chunk - 1; k >= 0; --k) {
      if (k > hi) continue;
      const uint32_t le = bm & (0xffffffffu >> (31 - k));
      const int32_t prev = le ? i0 + 31 - __clz(le) : pv;
      const uint32_t ge = bm >> k;
      const int32_t next = ge ? i0 + k + __ffs(ge) - 1 : nx;
      emit(i0 + k, base + (int64_t)(i0 + k) * step, prev, next);
    }
    if (bm) nx = i0 + __ffs(bm) - 1;
    bm = bmn;
    pv = pvn;
  }
}

// EDT axis 0 of line t (edt_line_axis0 with the chunk words).
template <typename TIn, bool FINAL>
__device__ __forceinline__ void edt_axis0_words(const TIn* __restrict__ in,
                                                int32_t* __restrict__ f0, const Shape& sh,
                                                int64_t t, const Samp& sp, const EdtOut& o,
                                                uint32_t* __restrict__ words) {
  int64_t c[kMaxNdim], base, b;
  decode_line(sh, 0, t, c, &base, &b);
  const int64_t nl = sh.numel / sh.size[0];
  axis0_nearest(in, base, sh.stride[0], (int)sh.size[0], t, nl, words,
                [&](int i, int64_t v, int32_t pv, int32_t nx) {
                  int32_t r;
                  if (pv < 0) r = nx;
                  else if (nx < 0) r = pv;
                  else r = (i - pv <= nx - i) ? pv : nx;
                  if (FINAL) {
                    c[0] = i;
                    int32_t f[kMaxNdim];
                    f[0] = r;
                    for (int j = 1; j < sh.ndim; ++j) f[j] = r >= 0 ? (int32_t)c[j] : 0;
                    edt_store(o, sh.ndim, v, c, f, sp);
                  } else {
                    f0[v] = r;
                  }
                });
}

// Chamfer axis 0 of the line starting at base (cdt_line_axis0 with the
// chunk words): distance to the nearest background along the line.
template <typename TIn, bool FINAL>
__device__ __forceinline__ void cdt_axis0_words(const TIn* __restrict__ in,
                                                int32_t* __restrict__ w, int64_t base,
                                                int64_t step, int L, int64_t t, int64_t nl,
                                                uint32_t* __restrict__ words) {
  axis0_nearest(in, base, step, L, t, nl, words,
                [&](int i, int64_t v, int32_t pv, int32_t nx) {
                  int32_t a = pv >= 0 ? i - pv : kCdtInf;
                  if (nx >= 0 && nx - i < a) a = nx - i;
                  w[v] = FINAL ? cdt_final_bits(a) : a;
                });
}

// Taxicab pass along a strided axis d >= 1, in place (cdt_line_taxicab's
// forward and backward sweeps, same values), in chunks of 16 voxels whose
// loads are issued together: a lane waits once per chunk instead of once per
// voxel.
template <bool FINAL>
__device__ __forceinline__ void cdt_taxicab_chunked(int32_t* __restrict__ w, int64_t base,
                                                    int64_t step, int L) {
  constexpr int K = 16;
  int32_t* p = w + base;
  int32_t prev = kCdtInf;
  for (int i0 = 0; i0 < L; i0 += K) {
    const int n = L - i0 < K ? L - i0 : K;
    int32_t a[K];
#pragma unroll
    for (int k = 0; k < K; ++k) a[k] = p[(int64_t)(i0 + (k < n ? k : n - 1)) * step];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if (k >= n) break;
      if (prev + 1 < a[k]) a[k] = prev + 1;
      prev = a[k];
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (k < n) p[(int64_t)(i0 + k) * step] = a[k];
  }
  int32_t nxt = kCdtInf;
  for (int i1 = L; i1 > 0; i1 -= K) {
    const int lo = i1 > K ? i1 - K : 0, n = i1 - lo;
    int32_t a[K];
#pragma unroll
    for (int k = 0; k < K; ++k) a[k] = p[(int64_t)(lo + (k < n ? k : n - 1)) * step];
#pragma unroll
    for (int k = K - 1; k >= 0; --k) {
      if (k >= n) continue;
      if (nxt + 1 < a[k]) a[k] = nxt + 1;
      nxt = a[k];
    }
#pragma unroll
    for (int k = 0; k < K; ++k)
      if (k < n) p[(int64_t)(lo + k) * step] = FINAL ? cdt_final_bits(a[k]) : a[k];
  }
}

}  // namespace tmh
