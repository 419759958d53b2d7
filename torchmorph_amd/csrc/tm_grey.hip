// Grey-scale erosion / dilation for gfx950.
//
// Semantics follow the reference gather formulation (grey_kernel.cu:101-170,
// offsets grey_kernel.cu:247-292, border modes grey_kernel.cu:63-97):
//   erosion   out[p] = min_i ( E(p + d_i) - s_i )
//   dilation  out[p] = max_i ( E(p - d_i) + s_i ),   d_i = coord_i - (size/2 + origin)
// over the active elements i of the structuring element, where E extends the
// input by the border mode (constant -> cval).  min/max are exact, so every
// evaluation order below returns the same bits as the reference (and SciPy).
//
// Kernels (float32 in/out, items of up to 3 spatial dims padded to 3-D):
//  * grey_box3_kernel   flat box SE.  2.5-D streaming: a workgroup owns a
//    64 x 16 (x, y) output tile and a chunk of z planes; each input plane's
//    halo tile is staged in LDS once (coalesced along x), reduced along x then
//    y in LDS, and pushed into a ring of kz reduced planes; every output plane
//    is the min/max over the ring.  HBM traffic is one read + one write per
//    voxel plus halo re-reads that mostly hit L2.
//  * grey_table3_kernel non-flat or non-box SE (<= 256 active elements).
//    Same streaming, but the ring holds raw input halo planes and each output
//    folds the element table (offset, value) held in kernel arguments.
//  * grey_axis_kernel   flat box fallback for any ndim: one 1-D window pass
//    per axis (ping-pong through the workspace).
//  * grey_gather_kernel general fallback for any ndim and any SE: one thread
//    per output folds the element table from device memory.
#include <algorithm>
#include <vector>

#include "tm_common.h"
#include "tm_tune.h"

namespace tmh {

// Border remap of coordinate c of an axis of `size` voxels (reference
// grey_kernel.cu:63-97): constant -> cval, nearest -> clamp, reflect
// (half-sample symmetric, period 2 size), mirror (whole-sample symmetric,
// period 2 (size - 1)), wrap (period size).  Outside the axis, reflections /
// shifts by one period repeat until c lands inside: the same index as the
// modulo form, in a few instructions per call site (a 64-bit modulo inlined
// at every load of an unrolled tile load made each kernel larger than the
// 64 KB instruction cache).  Each pass moves c at least one voxel closer to
// the axis, so the loop ends; callers pass c within a halo of the axis, where
// it runs once or twice unless the axis is shorter than the halo.
__device__ __forceinline__ int64_t gmap(int64_t c, int64_t size, int mode, bool* use_cval) {
  *use_cval = false;
  if (c >= 0 && c < size) return c;
  if (mode == 2) return c < 0 ? 0 : size - 1;  // nearest
  if (mode == 3 && size == 1) return 0;        // mirror of a single voxel
  if (mode != 1 && mode != 3 && mode != 4) {   // constant
    *use_cval = true;
    return 0;
  }
  while (c < 0 || c >= size) {
    if (mode == 1) c = c < 0 ? -c - 1 : 2 * size - 1 - c;  // reflect
    else if (mode == 3) c = c < 0 ? -c : 2 * (size - 1) - c;  // mirror
    else c = c < 0 ? c + size : c - size;  // wrap
  }
  return c;
}

template <bool MAX>
__device__ __forceinline__ float gop(float a, float b) {
  return MAX ? fmaxf(a, b) : fminf(a, b);
}

template <bool MAX>
__device__ __forceinline__ float gid() {
  return MAX ? -INFINITY : INFINITY;
}

constexpr int kTX = 64, kTY = 16;
// Halo tile elements a lane keeps in registers per plane: NE = 8 covers
// tiles of up to 2048 values (structuring elements up to about 9 x 9 in y,
// x: more waves per SIMD), NE = 16 up to 4096; elements beyond are staged
// directly (large elements).
constexpr int kPlaneElemsSmall = 8, kPlaneElemsLarge = 16;

// One mode-extended input plane halo tile [SY][SX] starting at (y0, x0),
// flattened over the workgroup: lane t holds elements e = t + 256 i, so the
// loads of a row are coalesced and the LDS stores are consecutive words.
// Two planes are in flight: the registers of plane n + 2 are filled while
// planes n and n + 1 are reduced.
// `bias` is added to every value: +0.0f for flat dilation (the reference
// adds its zero structure, which turns -0.0 into +0.0), -0.0f otherwise (an
// exact identity).
struct PlaneGeom {
  int D, H, W, SY, SX, mode;
  float cval, bias;
};

__device__ __forceinline__ float plane_value(const float* __restrict__ vol, const PlaneGeom& g,
                                             int64_t zm, bool zc, int y, int x) {
  if (zc) return g.cval + g.bias;
  bool yc, xc;
  const int64_t ym = gmap(y, g.H, g.mode, &yc);
  const int64_t xm = gmap(x, g.W, g.mode, &xc);
  return ((yc || xc) ? g.cval : vol[(zm * g.H + ym) * g.W + xm]) + g.bias;
}

// The in-plane offsets a lane loads are the same for every plane of its
// tile: mode-mapped once per workgroup (-1 = constant border, or past the
// tile's last element).  Plane offsets fit 32 bits (H W < 2^31, checked by
// the host).
template <int NE>
struct PlaneMap {
  int32_t off[NE];
  int ne;  // register elements in use, ceil(SY SX / 256) capped at NE (uniform)
};

template <int NE>
__device__ __forceinline__ PlaneMap<NE> plane_map(const PlaneGeom& g, int y0, int x0) {
  PlaneMap<NE> m;
  const int n = g.SY * g.SX;
  m.ne = min(NE, (n + 255) >> 8);
#pragma unroll
  for (int i = 0; i < NE; ++i) {
    const int e = (int)threadIdx.x + 256 * i;
    int32_t o = -1;
    if (i < m.ne && e < n) {
      const int r = e / g.SX, c = e - r * g.SX;
      bool yc, xc;
      const int64_t ym = gmap(y0 + r, g.H, g.mode, &yc);
      const int64_t xm = gmap(x0 + c, g.W, g.mode, &xc);
      o = (yc || xc) ? -1 : (int32_t)(ym * g.W + xm);
    }
    m.off[i] = o;
  }
  return m;
}

// Raw loads of one plane's halo tile, in flight until plane_store: every
// load has an in-range address (border positions read offset 0), and the
// border substitution and bias are applied at the store, so nothing
// consumes the loaded registers early.
template <int NE>
struct PlaneRegs {
  float v[NE];
  bool zc;  // the plane lies in the constant border
};

template <int NE>
__device__ __forceinline__ void plane_fetch(PlaneRegs<NE>& R, const float* __restrict__ vol,
                                            const PlaneGeom& g, const PlaneMap<NE>& m, int zi) {
  bool zc;
  const int64_t zm = gmap(zi, g.D, g.mode, &zc);
  R.zc = zc;
  const float* pl = vol + zm * g.H * g.W;
  // all NE loads, unconditionally (unused ones read offset 0): a fixed
  // number of loads per plane lets the compiler wait for exactly the older
  // plane's loads while the newer plane's stay in flight
#pragma unroll
  for (int i = 0; i < NE; ++i) R.v[i] = pl[m.off[i] < 0 ? 0 : m.off[i]];
}

template <int NE>
__device__ __forceinline__ void plane_store(const PlaneRegs<NE>& R, const PlaneMap<NE>& m,
                                            float* stage, const float* __restrict__ vol,
                                            const PlaneGeom& g, int zi, int y0, int x0) {
  const int n = g.SY * g.SX;
#pragma unroll
  for (int i = 0; i < NE; ++i) {
    const int e = (int)threadIdx.x + 256 * i;
    if (i < m.ne && e < n) stage[e] = (R.zc || m.off[i] < 0 ? g.cval : R.v[i]) + g.bias;
  }
  if (n > 256 * NE) {
    bool zc;
    const int64_t zm = gmap(zi, g.D, g.mode, &zc);
    for (int e = 256 * NE + (int)threadIdx.x; e < n; e += 256) {
      const int r = e / g.SX, c = e - r * g.SX;
      stage[e] = plane_value(vol, g, zm, zc, y0 + r, x0 + c);
    }
  }
}

// Logical block (tile fastest, then z chunk, then item) of a 1-D launch,
// XCD-swizzled so neighbouring tiles (shared halos) share an L2.
__device__ __forceinline__ void grey_block(int ntx, int ntiles, int nz, int* tx, int* ty, int* zc,
                                           int64_t* item) {
  const int lb = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = lb % ntiles;
  const int rest = lb / ntiles;
  *tx = t % ntx;
  *ty = t / ntx;
  *zc = rest % nz;
  *item = rest / nz;
}

// Epilogue of the composite operators (reference grey.py:200-289), applied
// to the min/max `acc` of output voxel p; a and b are item-relative arrays of
// the output's layout (a may be the output itself: each voxel reads its own
// value before writing it).  Same float32 operations, in the same order, as
// the reference's tensor expressions:
//   1: acc - a[p]            gradient (a = erosion), black tophat (a = input)
//   2: a[p] - acc            white tophat (a = input)
//   3: (acc + a[p]) - 2 b[p] laplace (a = erosion, b = input)
struct GreyEpi {
  const float* a;
  const float* b;
};

template <int EPI>
__device__ __forceinline__ float grey_epi(float acc, const GreyEpi& e, int64_t p) {
  if (EPI == 1) return acc - e.a[p];
  if (EPI == 2) return e.a[p] - acc;
  if (EPI == 3) return (acc + e.a[p]) - 2.0f * e.b[p];
  return acc;
}

// Windows of at least kVhgw planes along z are reduced by van Herk /
// Gil-Werman: the planes are cut into blocks of k, each block gets its
// suffix min/max in place once complete and a running prefix per output
// voxel, and the window [z, z + k - 1] is suffix[z] combined with
// prefix[z + k - 1] (just the prefix when z starts a block).  min/max are
// exact, so any grouping gives the same values.
constexpr int kVhgw = 10;

// x and y: each lane reduces a run of T consecutive outputs of one line from
// LDS, out[i] = op(v[(i + j) es], j < k).  For k >= T the windows share the
// middle M = v[T-1 .. k-1]; window i is S_i op M op P_i with the suffix
// S_i = v[i .. T-2] and the prefix P_i = v[k .. k+i-1], so T outputs take
// k + 3 T - 4 operations and k + T - 1 loads, all independent of the other
// lanes (no scans, no extra barriers).  For k < T the k + T - 1 values are
// held in registers and each window is folded directly.
template <bool MAX, int T>
__device__ __forceinline__ void win_run(const float* v, int es, int k, float (&out)[T]) {
  if (k >= T) {
    float lo[T - 1], hi[T - 1];
#pragma unroll
    for (int i = 0; i < T - 1; ++i) {
      lo[i] = v[i * es];
      hi[i] = v[(k + i) * es];
    }
    float m = v[(T - 1) * es];
    for (int j = T; j < k; ++j) m = gop<MAX>(m, v[j * es]);
    // out[i] = S_i op M op P_i: S from the right, P from the left
    float sfx = lo[T - 2];
    out[T - 2] = gop<MAX>(sfx, m);
#pragma unroll
    for (int i = T - 3; i >= 0; --i) {
      sfx = gop<MAX>(lo[i], sfx);
      out[i] = gop<MAX>(sfx, m);
    }
    float pre = hi[0];
    out[T - 1] = m;
#pragma unroll
    for (int i = 1; i < T; ++i) {
      if (i > 1) pre = gop<MAX>(pre, hi[i - 1]);
      out[i] = gop<MAX>(out[i], pre);
    }
  } else {
    float w[2 * T - 1];
#pragma unroll
    for (int j = 0; j < 2 * T - 1; ++j) w[j] = j < T + k - 1 ? v[j * es] : 0.0f;
#pragma unroll
    for (int i = 0; i < T; ++i) {
      float acc = w[i];
#pragma unroll
      for (int j = 1; j < T; ++j)
        if (j < k) acc = gop<MAX>(acc, w[i + j]);
      out[i] = acc;
    }
  }
}

// Runs per line of the x and y passes: 8 outputs along x (8 runs per
// 64-wide row, rows in sweeps of 32 lanes), 4 along y (4 runs per 16-high
// column, 64 columns).  Windows shorter than a run are folded per output,
// one output per lane at a time (consecutive lanes on consecutive LDS words,
// which measured faster than runs for such windows).
constexpr int kRunX = 8, kRunY = 4;
// Pitch of the x-pass results: one word more than a row, so the 32 lanes of
// a run position that store rows r .. r + 31 hit 32 different LDS banks.
constexpr int kXP = kTX + 1;

// x pass of one staged plane: rows of SX values (pitch SX) to rows of kTX
// window results (pitch kXP).
template <bool MAX>
__device__ __forceinline__ void xpass(const float* stage, int SX, int SY, int kx, float* xr) {
  if (kx < kRunX) {
    for (int e = threadIdx.x; e < SY * kTX; e += blockDim.x) {
      const int r = e >> 6, c = e & 63;
      const float* s = stage + r * SX + c;
      float acc = s[0];
      for (int l = 1; l < kx; ++l) acc = gop<MAX>(acc, s[l]);
      xr[r * kXP + c] = acc;
    }
    return;
  }
  const int g = threadIdx.x >> 5, r0 = threadIdx.x & 31;  // run, first row
  for (int r = r0; r < SY; r += 32) {
    float o[kRunX];
    win_run<MAX, kRunX>(stage + r * SX + g * kRunX, 1, kx, o);
#pragma unroll
    for (int i = 0; i < kRunX; ++i) xr[r * kXP + g * kRunX + i] = o[i];
  }
}

// y pass: SY rows of kTX x-results to kTY rows of window results.  Lane
// (g, c) writes rows kRunY g .. kRunY g + kRunY - 1 of column c: the outputs
// it later folds along z, so a z ring slot is only ever touched by its lane.
template <bool MAX>
__device__ __forceinline__ void ypass(const float* xr, int ky, float* slot) {
  const int c = threadIdx.x & 63, g = threadIdx.x >> 6;
  const float* col = xr + g * kRunY * kXP + c;
  if (ky < kRunY) {
#pragma unroll
    for (int i = 0; i < kRunY; ++i) {
      float acc = col[i * kXP];
      for (int j = 1; j < ky; ++j) acc = gop<MAX>(acc, col[(i + j) * kXP]);
      slot[(g * kRunY + i) * kTX + c] = acc;
    }
    return;
  }
  float o[kRunY];
  win_run<MAX, kRunY>(col, kXP, ky, o);
#pragma unroll
  for (int i = 0; i < kRunY; ++i) slot[(g * kRunY + i) * kTX + c] = o[i];
}

// One (x, y) tile of one item over output planes [zb, ze).
template <bool MAX, int NR, int EPI = 0>
__device__ __forceinline__ void grey_box3_tile(const float* __restrict__ vol,
                                               float* __restrict__ ovol, int D, int H, int W,
                                               int kz, int ky, int kx, int sz, int sy, int sx,
                                               int mode, float cval, int tx, int ty, int zb,
                                               int ze, float* sm, GreyEpi epi = GreyEpi{}) {
  const int SX = kTX + kx - 1, SY = kTY + ky - 1;
  float* stage = sm;
  float* xr = stage + SY * SX;
  float* ring = xr + SY * kXP;
  const int x0 = tx * kTX, y0 = ty * kTY;
  const float ID = gid<MAX>();
  const PlaneGeom g{D, H, W, SY, SX, mode, cval, MAX ? 0.0f : -0.0f};
  const int nplanes = (ze - zb) + kz - 1;
  const PlaneMap<NR> pm = plane_map<NR>(g, y0 + sy, x0 + sx);
  const bool vz = kz >= kVhgw;
  // z by blocks of kz planes: a running prefix per output voxel (this
  // thread's four), and the ring holding the previous block's suffixes
  float zpre[kTY * kTX / 256];
  auto step = [&](PlaneRegs<NR>& R, int n) {
    const int zi = zb + sz + n;
    plane_store(R, pm, stage, vol, g, zi, y0 + sy, x0 + sx);
    __syncthreads();
    // plane n + 2's loads are in flight while planes n and n + 1 are reduced
    plane_fetch(R, vol, g, pm, zi + 2);  // past the last plane: mapped, unused
    xpass<MAX>(stage, SX, SY, kx, xr);
    __syncthreads();
    float* slot = ring + (n % kz) * (kTY * kTX);
    ypass<MAX>(xr, ky, slot);
    // no barrier: this lane's z fold reads only the ring entries its own
    // y pass wrote (and the next plane's stage store waits for the barrier
    // after its x pass, which every lane reaches after this y pass)
    const int z = zb + n - (kz - 1);
    const int pos = n % kz;  // place of plane n in its block of kz
#pragma unroll
    for (int q = 0; q < kRunY; ++q) {
      const int e = ((threadIdx.x >> 6) * kRunY + q) * kTX + (threadIdx.x & 63);
      float acc;
      if (vz) {
        const float v = ring[pos * (kTY * kTX) + e];
        zpre[q] = pos == 0 ? v : gop<MAX>(zpre[q], v);
        // window planes n - kz + 1 .. n: the previous block's suffix from
        // n - kz + 1 (ring slot pos + 1) and this block's prefix
        acc = pos == kz - 1 ? zpre[q] : gop<MAX>(ring[(pos + 1) * (kTY * kTX) + e], zpre[q]);
        if (pos == kz - 1) {  // block complete: its suffixes, in place
          for (int i = kz - 2; i >= 0; --i)
            ring[i * (kTY * kTX) + e] =
                gop<MAX>(ring[i * (kTY * kTX) + e], ring[(i + 1) * (kTY * kTX) + e]);
        }
      } else {
        acc = ID;
        if (z >= zb)
          for (int i = 0; i < kz; ++i) acc = gop<MAX>(acc, ring[i * (kTY * kTX) + e]);
      }
      const int r = e >> 6, c = e & 63;
      const int y = y0 + r, x = x0 + c;
      if (z >= zb && y < H && x < W) {
        const int64_t p = ((int64_t)z * H + y) * W + x;
        ovol[p] = grey_epi<EPI>(acc, epi, p);
      }
    }
  };
  PlaneRegs<NR> RA, RB;
  plane_fetch(RA, vol, g, pm, zb + sz);
  if (nplanes > 1) plane_fetch(RB, vol, g, pm, zb + sz + 1);
  for (int n = 0; n < nplanes; n += 2) {
    step(RA, n);
    if (n + 1 < nplanes) step(RB, n + 1);
  }
}

template <bool MAX, int NR, int EPI = 0>
__global__ __launch_bounds__(256) void grey_box3_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, int D, int H,
                                                        int W, int kz, int ky, int kx, int sz,
                                                        int sy, int sx, int mode, float cval,
                                                        int zchunk, GreyEpi epi) {
  extern __shared__ float sm[];
  const int ntx = (W + kTX - 1) / kTX, nty = (H + kTY - 1) / kTY;
  const int nz = (D + zchunk - 1) / zchunk;
  int tx, ty, zci;
  int64_t item;
  grey_block(ntx, ntx * nty, nz, &tx, &ty, &zci, &item);
  const int zb = zci * zchunk;
  const int64_t n = (int64_t)D * H * W;
  const GreyEpi ei{epi.a ? epi.a + item * n : nullptr, epi.b ? epi.b + item * n : nullptr};
  grey_box3_tile<MAX, NR, EPI>(in + item * n, out + item * n, D, H, W, kz, ky, kx, sz, sy, sx,
                               mode, cval, tx, ty, zb, min(D, zb + zchunk), sm, ei);
}

// Flat boxes of at most 3 voxels per axis, register streaming.  A wave owns a
// 64-wide x strip of TY output rows and a chunk of z planes; the four waves
// of a workgroup take four consecutive y tiles.  Lane l holds input column
// x0 + sx + l of every row it streams, and lanes 0 .. KX - 2 also the voxels
// past the strip.  The x window reads the right-hand neighbours through DPP
// lane shifts, the wave's last lanes taking the voxels past the strip.  The
// y window folds KY consecutive x-results, and the z window keeps the last
// KZ - 1 y-results of each output row in registers.  No LDS and no barriers.
// The next plane's rows are loaded while the current plane is reduced: two
// row buffers alternate, and a buffer is reloaded right after it is
// consumed, so the loads land in the registers they are consumed from.
// Every load address is in range (border-mapped; constant borders read
// offset 0 and are substituted when consumed), and a load is a uniform plane
// base plus a 32-bit byte offset (planes of at most 2^30 voxels).
// EPI 0..3: one erosion / dilation with grey_epi's epilogue.  EPI 4 / 5: the
// one-pass gradient (max - min) / laplace ((max + min) - 2 input) of
// grey_box3_dual_tile, min and max folded side by side (MAX unused).
constexpr int kRwWaves = 4;

template <int KX, int KY, int TY>
struct RwRows {
  float v[TY + KY - 1];  // this lane's column of each input row
  float e[KX > 1 ? TY + KY - 1 : 1];  // lanes 0 .. KX - 2: the voxels past the strip
  bool zc;               // the plane lies in the constant border
};

// Value of lane l + 1 (lane 63: `last`), by a DPP wave shift.
__device__ __forceinline__ float rw_next_lane(float v, float last) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(last), __float_as_int(v),
                                                    0x130 /* wave_shl:1 */, 0xf, 0xf, false));
}
// Value of lane l - 1 (lane 0: `first`); the integer forms of both shifts.
__device__ __forceinline__ float lane_prev(float v, float first) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(first), __float_as_int(v),
                                                    0x138 /* wave_shr:1 */, 0xf, 0xf, false));
}
__device__ __forceinline__ int lane_prev_i(int v, int first) {
  return __builtin_amdgcn_update_dpp(first, v, 0x138, 0xf, 0xf, false);
}
__device__ __forceinline__ int lane_next_i(int v, int last) {
  return __builtin_amdgcn_update_dpp(last, v, 0x130, 0xf, 0xf, false);
}

template <bool MAX, int KX, int KY, int KZ, int TY, int EPI>
__global__ __launch_bounds__(256) void grey_boxr_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, int D, int H,
                                                        int W, int sz, int sy, int sx, int mode,
                                                        float cval, int zchunk, GreyEpi epi) {
  constexpr int NR = TY + KY - 1;
  const int ntx = (W + 63) >> 6;
  const int nty = (H + TY - 1) / TY;
  const int nyb = (nty + kRwWaves - 1) / kRwWaves;
  const int nz = (D + zchunk - 1) / zchunk;
  int tx, tyb, zci;
  int64_t item;
  grey_block(ntx, ntx * nyb, nz, &tx, &tyb, &zci, &item);
  const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int ty = tyb * kRwWaves + wave;
  if (ty >= nty) return;  // whole wave; no barriers in this kernel
  const int zb = zci * zchunk, ze = min(D, zb + zchunk);
  const int x0 = tx * 64, y0 = ty * TY;
  const int64_t n = (int64_t)D * H * W;
  const float* vol = in + item * n;
  float* ovol = out + item * n;
  const GreyEpi ei{epi.a ? epi.a + item * n : nullptr, epi.b ? epi.b + item * n : nullptr};
  const float bias = MAX ? 0.0f : -0.0f;  // the flat dilation's +0.0f; an identity otherwise
  constexpr bool DU = EPI >= 4;

  // border maps (byte offsets) of this lane's column, of the voxel past the
  // strip this lane loads (lanes 0 .. KX - 2) and of the rows
  bool xc, ec;
  const uint32_t xo = (uint32_t)gmap(x0 + sx + lane, W, mode, &xc) * 4u;
  const uint32_t eo =
      (uint32_t)gmap(x0 + sx + 64 + min(lane, KX > 1 ? KX - 2 : 0), W, mode, &ec) * 4u;
  uint32_t hcm = 0;  // voxels past the strip in the constant border
#pragma unroll
  for (int j = 0; j < KX - 1; ++j) hcm |= (__builtin_amdgcn_readlane((int)ec, j) ? 1u : 0u) << j;
  uint32_t ro[NR];
  uint32_t rcm = 0;  // rows in the constant border
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    bool c;
    ro[r] = (uint32_t)gmap(y0 + sy + r, H, mode, &c) * (uint32_t)W * 4u;
    rcm |= (c ? 1u : 0u) << r;
  }

  auto fetch = [&](RwRows<KX, KY, TY>& R, int zi) {
    bool zc;
    const int64_t zm = gmap(zi, D, mode, &zc);
    R.zc = zc;
    const char* pl = (const char*)(vol + zm * H * W);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      R.v[r] = *(const float*)(pl + (ro[r] + xo));
      if (KX > 1) R.e[r] = *(const float*)(pl + (ro[r] + eo));
    }
  };

  const int nplanes = (ze - zb) + KZ - 1;
  constexpr int NZ = KZ > 1 ? KZ - 1 : 1;
  float zf[NZ][TY];               // y-results of the previous KZ - 1 planes
  float zn[DU ? NZ : 1][TY];      // DU: their minima (zf: maxima)
#pragma unroll
  for (int k = 0; k < NZ; ++k)
#pragma unroll
    for (int t = 0; t < TY; ++t) {
      zf[k][t] = DU ? -INFINITY : gid<MAX>();
      if (DU) zn[k][t] = INFINITY;
    }
  // epilogue operands of an output plane: a (EPI 1..3) or the input (EPI 5),
  // and b (EPI 3); loaded when the plane's step starts, consumed at its
  // stores (clamped coordinates: every address in range, unused values)
  constexpr bool EA = (EPI >= 1 && EPI <= 3) || EPI == 5, EB = EPI == 3;
  const float* pa = EPI == 5 ? vol : ei.a;
  const int xl = min(x0 + lane, W - 1);
  auto step = [&](RwRows<KX, KY, TY>& R, int p) {
    const int z = zb + p - (KZ - 1);
    float ea[EA ? TY : 1], eb[EB ? TY : 1];
    if (EA) {
      const int64_t zq = (int64_t)max(z, zb) * H;
#pragma unroll
      for (int t = 0; t < TY; ++t) {
        const int64_t q = (zq + min(y0 + t, H - 1)) * W + xl;
        ea[t] = pa[q];
        if (EB) eb[EB ? t : 0] = ei.b[q];
      }
    }
    float xw[NR], xn[DU ? NR : 1];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const bool cr = R.zc || ((rcm >> r) & 1u);
      const float raw = (cr || xc) ? cval : R.v[r];
      if (!DU) {
        float acc = raw + bias;
        float s = acc;
#pragma unroll
        for (int j = 0; j < KX - 1; ++j) {
          const float h = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(R.e[r]), j));
          s = rw_next_lane(s, ((cr || ((hcm >> j) & 1u)) ? cval : h) + bias);
          acc = gop<MAX>(acc, s);
        }
        xw[r] = acc;
      } else {  // max of the values + 0.0f, min of the raw values
        float hi = raw + 0.0f, lo = raw, s = raw;
#pragma unroll
        for (int j = 0; j < KX - 1; ++j) {
          const float h = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(R.e[r]), j));
          s = rw_next_lane(s, (cr || ((hcm >> j) & 1u)) ? cval : h);
          hi = fmaxf(hi, s + 0.0f);
          lo = fminf(lo, s);
        }
        xw[r] = hi;
        xn[DU ? r : 0] = lo;
      }
    }
    // this buffer's registers are free: plane p + 2 (past the last plane:
    // mapped, unused)
    fetch(R, zb + sz + p + 2);
#pragma unroll
    for (int t = 0; t < TY; ++t) {
      const int y = y0 + t, x = x0 + lane;
      const bool st = z >= zb && y < H && x < W;
      const int64_t q = ((int64_t)z * H + y) * W + x;
      if (!DU) {
        float yw = xw[t];
#pragma unroll
        for (int j = 1; j < KY; ++j) yw = gop<MAX>(yw, xw[t + j]);
        float acc = yw;
#pragma unroll
        for (int k = 0; k < KZ - 1; ++k) acc = gop<MAX>(acc, zf[k][t]);
#pragma unroll
        for (int k = 0; k + 1 < KZ - 1; ++k) zf[k][t] = zf[k + 1][t];
        if (KZ > 1) zf[KZ - 2][t] = yw;
        if (st) {
          float o = acc;  // grey_epi's operations on the preloaded operands
          if (EPI == 1) o = acc - ea[EA ? t : 0];
          if (EPI == 2) o = ea[EA ? t : 0] - acc;
          if (EPI == 3) o = (acc + ea[EA ? t : 0]) - 2.0f * eb[EB ? t : 0];
          ovol[q] = o;
        }
      } else {
        float yh = xw[t], yl = xn[DU ? t : 0];
#pragma unroll
        for (int j = 1; j < KY; ++j) {
          yh = fmaxf(yh, xw[t + j]);
          yl = fminf(yl, xn[DU ? t + j : 0]);
        }
        float hi = yh, lo = yl;
#pragma unroll
        for (int k = 0; k < KZ - 1; ++k) {
          hi = fmaxf(hi, zf[k][t]);
          lo = fminf(lo, zn[DU ? k : 0][t]);
        }
#pragma unroll
        for (int k = 0; k + 1 < KZ - 1; ++k) {
          zf[k][t] = zf[k + 1][t];
          zn[DU ? k : 0][t] = zn[DU ? k + 1 : 0][t];
        }
        if (KZ > 1) {
          zf[KZ - 2][t] = yh;
          zn[DU ? KZ - 2 : 0][t] = yl;
        }
        if (st) ovol[q] = EPI == 4 ? hi - lo : (hi + lo) - 2.0f * ea[EA ? t : 0];
      }
    }
  };
  RwRows<KX, KY, TY> A, B;
  fetch(A, zb + sz);
  fetch(B, zb + sz + 1);
  for (int p = 0; p < nplanes; p += 2) {
    step(A, p);
    if (p + 1 < nplanes) step(B, p + 1);
  }
}

// The ball of radius 3 with heights floor(3 - r) (r <= 3, zero origin),
// register streaming.  Grouping the elements by superlevel sets of their
// height h (balls of radius 3, 2, 1 and 0 for h = 0..3), a dilation is
//   max_h (h + max of x over the ball B_h),
// exactly: x -> x + h rounds monotonically, so the terms the larger balls
// add are never above the element's own, and max over a set of x + h is
// (max over the set of x) + h bit for bit (no ties between -0 and +0 arise:
// a sum x + h that is zero is +0).  Erosion likewise with min and -h (h = 0
// adds -0.0f, the identity).  A ball is a union of z planes of 2-D disks, and
// a disk a union of rows of 1-D x runs of half-width 0..3, so per input
// plane a wave forms, for each of its rows, the x-run extrema S_0..S_3 (DPP
// lane shifts; lanes 0..5 load the six columns past the strip), then per
// output row the seven disks (radius^2 9, 8, 5, 4, 3, 1, 0) that the four
// balls cut from the planes at |dz| = 0..3, adds the heights, and folds the
// plane into the seven output planes z = p - 3 .. p + 3 still open.  A wave
// owns a 64-wide x strip of kBallRows output rows and a z chunk; the four
// waves of a workgroup take consecutive y tiles; no LDS, no barriers.
constexpr int kBallRows = 4;

template <bool MAX>
__global__ __launch_bounds__(256) void grey_ballr_kernel(const float* __restrict__ in,
                                                         float* __restrict__ out, int D, int H,
                                                         int W, int mode, float cval,
                                                         int zchunk, GreyEpi epi, int epil) {
  constexpr int TY = kBallRows, NR = TY + 6;
  const int ntx = (W + 63) >> 6;
  const int nty = (H + TY - 1) / TY;
  const int nyb = (nty + kRwWaves - 1) / kRwWaves;
  const int nz = (D + zchunk - 1) / zchunk;
  int tx, tyb, zci;
  int64_t item;
  grey_block(ntx, ntx * nyb, nz, &tx, &tyb, &zci, &item);
  const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int ty = tyb * kRwWaves + wave;
  if (ty >= nty) return;  // whole wave; no barriers in this kernel
  const int zb = zci * zchunk, ze = min(D, zb + zchunk);
  const int x0 = tx * 64, y0 = ty * TY;
  const int64_t n = (int64_t)D * H * W;
  const float* vol = in + item * n;
  float* ovol = out + item * n;
  const GreyEpi ei{epi.a ? epi.a + item * n : nullptr, epi.b ? epi.b + item * n : nullptr};
  const float ID = gid<MAX>();
  const float h0 = MAX ? 0.0f : -0.0f;
  const float hs = MAX ? 1.0f : -1.0f;  // heights 1..3 enter as h * hs

  // border maps (byte offsets): this lane's column, the column past the strip
  // it loads (lanes 0..2: x0 - 3 + l; lanes 3..5: x0 + 61 + l), the rows
  bool xc, ec;
  const uint32_t xo = (uint32_t)gmap(x0 + lane, W, mode, &xc) * 4u;
  const int el = lane < 3 ? x0 - 3 + lane : x0 + 61 + min(lane, 5);
  const uint32_t eo = (uint32_t)gmap(el, W, mode, &ec) * 4u;
  uint32_t hcm = 0;
#pragma unroll
  for (int j = 0; j < 6; ++j) hcm |= (__builtin_amdgcn_readlane((int)ec, j) ? 1u : 0u) << j;
  uint32_t ro[NR];
  uint32_t rcm = 0;
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    bool c;
    ro[r] = (uint32_t)gmap(y0 - 3 + r, H, mode, &c) * (uint32_t)W * 4u;
    rcm |= (c ? 1u : 0u) << r;
  }
  auto fetch = [&](RwRows<2, NR - TY + 1, TY>& R, int zi) {
    bool zc;
    const int64_t zm = gmap(zi, D, mode, &zc);
    R.zc = zc;
    const char* pl = (const char*)(vol + zm * H * W);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      R.v[r] = *(const float*)(pl + (ro[r] + xo));
      R.e[r] = *(const float*)(pl + (ro[r] + eo));
    }
  };
  auto ex = [&](const RwRows<2, NR - TY + 1, TY>& R, int r, int j, bool cr) {
    const float h = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(R.e[r]), j));
    return (cr || ((hcm >> j) & 1u)) ? cval : h;
  };

  // acc[k][t]: output plane z = p - 3 + k of output row t, open while the
  // planes p - 3 + k - 3 .. p - 3 + k + 3 arrive
  float acc[7][TY];
#pragma unroll
  for (int k = 0; k < 7; ++k)
#pragma unroll
    for (int t = 0; t < TY; ++t) acc[k][t] = ID;
  const int nplanes = (ze - zb) + 6;
  auto step = [&](RwRows<2, NR - TY + 1, TY>& R, int p) {
    // x-run extrema of every row: S[w][r] over columns x - w .. x + w
    float S0[NR], S1[NR], S2[NR], S3[NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
      const bool cr = R.zc || ((rcm >> r) & 1u);
      const float v = (cr || xc) ? cval : R.v[r];
      const float l1 = lane_prev(v, ex(R, r, 2, cr));
      const float l2 = lane_prev(l1, ex(R, r, 1, cr));
      const float l3 = lane_prev(l2, ex(R, r, 0, cr));
      const float r1 = rw_next_lane(v, ex(R, r, 3, cr));
      const float r2 = rw_next_lane(r1, ex(R, r, 4, cr));
      const float r3 = rw_next_lane(r2, ex(R, r, 5, cr));
      S0[r] = v;
      S1[r] = gop<MAX>(gop<MAX>(l1, v), r1);
      S2[r] = gop<MAX>(gop<MAX>(l2, S1[r]), r2);
      S3[r] = gop<MAX>(gop<MAX>(l3, S2[r]), r3);
    }
    // this buffer's registers are free: plane p + 2 (mapped, unused past the end)
    fetch(R, zb - 3 + p + 2);
    const int z = zb - 3 + p - 3;  // the output plane this plane completes
#pragma unroll
    for (int t = 0; t < TY; ++t) {
      const int c = t + 3;  // row of output row t among the plane's rows
      auto m3 = [&](const float* A, int i) { return gop<MAX>(gop<MAX>(A[c - i], A[c]), A[c + i]); };
      auto m2 = [&](const float* A, int i) { return gop<MAX>(A[c - i], A[c + i]); };
      const float d9 = gop<MAX>(gop<MAX>(gop<MAX>(S3[c], m2(S2, 1)), m2(S2, 2)), m2(S0, 3));
      const float d8 = gop<MAX>(gop<MAX>(S2[c], m2(S2, 1)), m2(S2, 2));
      const float d5 = gop<MAX>(gop<MAX>(S2[c], m2(S2, 1)), m2(S1, 2));
      const float d4 = gop<MAX>(gop<MAX>(S2[c], m2(S1, 1)), m2(S0, 2));
      const float d3 = m3(S1, 1);
      const float d1 = gop<MAX>(S1[c], m2(S0, 1));
      const float d0 = S0[c];
      // the plane's part of each ball level at |dz| = 0..3, heights added
      const float c0 = gop<MAX>(gop<MAX>(d9 + h0, d4 + hs), gop<MAX>(d1 + 2.0f * hs, d0 + 3.0f * hs));
      const float c1 = gop<MAX>(gop<MAX>(d8 + h0, d3 + hs), d0 + 2.0f * hs);
      const float c2 = gop<MAX>(d5 + h0, d0 + hs);
      const float c3 = d0 + h0;
      // open outputs z + k (k = 0..6) at dz = 3 - k
      acc[0][t] = gop<MAX>(acc[0][t], c3);
      acc[1][t] = gop<MAX>(acc[1][t], c2);
      acc[2][t] = gop<MAX>(acc[2][t], c1);
      acc[3][t] = gop<MAX>(acc[3][t], c0);
      acc[4][t] = gop<MAX>(acc[4][t], c1);
      acc[5][t] = gop<MAX>(acc[5][t], c2);
      acc[6][t] = gop<MAX>(acc[6][t], c3);
      const int y = y0 + t, x = x0 + lane;
      if (z >= zb && y < H && x < W) {
        const int64_t q = ((int64_t)z * H + y) * W + x;
        ovol[q] = epil == 1 ? acc[0][t] - ei.a[q] : epil == 2 ? ei.a[q] - acc[0][t] : acc[0][t];
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) acc[k][t] = acc[k + 1][t];
      acc[6][t] = ID;
    }
  };
  RwRows<2, NR - TY + 1, TY> A, B;
  fetch(A, zb - 3);
  fetch(B, zb - 2);
  for (int p = 0; p < nplanes; p += 2) {
    step(A, p);
    if (p + 1 < nplanes) step(B, p + 1);
  }
}

// Gradient (dilation - erosion, COMB 0) or laplace ((dilation + erosion) -
// 2 input, COMB 1) of a flat box whose erosion and dilation windows coincide
// (odd sizes, zero origin), in one pass: each staged plane is reduced to its
// min and its max together.  The stage holds the raw mode-extended values;
// the max path adds the flat dilation's +0.0f (as grey_box3_tile's bias
// does), the min path adds nothing (the erosion's -0.0f is an identity).
template <int NR, int COMB>
__device__ __forceinline__ void grey_box3_dual_tile(const float* __restrict__ vol,
                                                    float* __restrict__ ovol, int D, int H,
                                                    int W, int kz, int ky, int kx, int sz, int sy,
                                                    int sx, int mode, float cval, int tx, int ty,
                                                    int zb, int ze, float* sm) {
  const int SX = kTX + kx - 1, SY = kTY + ky - 1;
  float* stage = sm;
  float* xmn = stage + SY * SX;
  float* xmx = xmn + SY * kTX;
  float* rmn = xmx + SY * kTX;
  float* rmx = rmn + kz * kTY * kTX;
  const int x0 = tx * kTX, y0 = ty * kTY;
  const PlaneGeom g{D, H, W, SY, SX, mode, cval, -0.0f};
  const int nplanes = (ze - zb) + kz - 1;
  const PlaneMap<NR> pm = plane_map<NR>(g, y0 + sy, x0 + sx);
  auto step = [&](PlaneRegs<NR>& R, int n) {
    const int zi = zb + sz + n;
    plane_store(R, pm, stage, vol, g, zi, y0 + sy, x0 + sx);
    __syncthreads();
    plane_fetch(R, vol, g, pm, zi + 2);  // past the last plane: mapped, unused
    for (int e = threadIdx.x; e < SY * kTX; e += blockDim.x) {
      const int r = e >> 6, c = e & 63;
      const float* s = stage + r * SX + c;
      float lo = INFINITY, hi = -INFINITY;
      for (int l = 0; l < kx; ++l) {
        lo = fminf(lo, s[l]);
        hi = fmaxf(hi, s[l] + 0.0f);
      }
      xmn[e] = lo;
      xmx[e] = hi;
    }
    __syncthreads();
    const int slot = (n % kz) * (kTY * kTX);
    for (int e = threadIdx.x; e < kTY * kTX; e += blockDim.x) {
      const int r = e >> 6, c = e & 63;
      float lo = INFINITY, hi = -INFINITY;
      for (int j = 0; j < ky; ++j) {
        lo = fminf(lo, xmn[(r + j) * kTX + c]);
        hi = fmaxf(hi, xmx[(r + j) * kTX + c]);
      }
      rmn[slot + e] = lo;
      rmx[slot + e] = hi;
    }
    __syncthreads();
    const int z = zb + n - (kz - 1);
    if (z >= zb) {
      for (int e = threadIdx.x; e < kTY * kTX; e += blockDim.x) {
        const int r = e >> 6, c = e & 63;
        const int y = y0 + r, x = x0 + c;
        if (y < H && x < W) {
          float lo = INFINITY, hi = -INFINITY;
          for (int i = 0; i < kz; ++i) {
            lo = fminf(lo, rmn[i * (kTY * kTX) + e]);
            hi = fmaxf(hi, rmx[i * (kTY * kTX) + e]);
          }
          const int64_t p = ((int64_t)z * H + y) * W + x;
          ovol[p] = COMB == 0 ? hi - lo : (hi + lo) - 2.0f * vol[p];
        }
      }
    }
  };
  PlaneRegs<NR> RA, RB;
  plane_fetch(RA, vol, g, pm, zb + sz);
  if (nplanes > 1) plane_fetch(RB, vol, g, pm, zb + sz + 1);
  for (int n = 0; n < nplanes; n += 2) {
    step(RA, n);
    if (n + 1 < nplanes) step(RB, n + 1);
  }
}

template <int NR, int COMB>
__global__ __launch_bounds__(256) void grey_box3_dual_kernel(const float* __restrict__ in,
                                                             float* __restrict__ out, int D,
                                                             int H, int W, int kz, int ky, int kx,
                                                             int sz, int sy, int sx, int mode,
                                                             float cval, int zchunk) {
  extern __shared__ float sm[];
  const int ntx = (W + kTX - 1) / kTX, nty = (H + kTY - 1) / kTY;
  const int nz = (D + zchunk - 1) / zchunk;
  int tx, ty, zci;
  int64_t item;
  grey_block(ntx, ntx * nty, nz, &tx, &ty, &zci, &item);
  const int zb = zci * zchunk;
  const int64_t n = (int64_t)D * H * W;
  grey_box3_dual_tile<NR, COMB>(in + item * n, out + item * n, D, H, W, kz, ky, kx, sz, sy, sx,
                                mode, cval, tx, ty, zb, min(D, zb + zchunk), sm);
}

constexpr int kMaxTable = 256;
struct GreyTable {
  int n;
  int8_t dz[kMaxTable], dy[kMaxTable], dx[kMaxTable];
  float v[kMaxTable];
};

// Ring of raw input halo planes; zlo..zhi, ylo..yhi, xlo..xhi bound the table
// offsets.  Per plane step the LDS offset of every table element (ring slot
// of its dz, then dy, dx) is computed once, so the inner loop is one LDS
// read, one add and one min/max per element.
template <bool MAX, int NR, int EPI = 0>
__device__ __forceinline__ void grey_table3_tile(const float* __restrict__ vol,
                                                 float* __restrict__ ovol, int D, int H, int W,
                                                 const GreyTable& tab, int zlo, int zhi, int ylo,
                                                 int yhi, int xlo, int xhi, int mode, float cval,
                                                 int tx, int ty, int zb, int ze, float* sm,
                                                 int* soff, float* sval,
                                                 GreyEpi epi = GreyEpi{}) {
  const int kz = zhi - zlo + 1;
  const int SX = kTX + xhi - xlo, SY = kTY + yhi - ylo;
  const int plane = SY * SX;
  const int x0 = tx * kTX, y0 = ty * kTY;
  const float ID = gid<MAX>();
  const PlaneGeom g{D, H, W, SY, SX, mode, cval, -0.0f};
  const int nplanes = (ze - zb) + kz - 1;
  const PlaneMap<NR> pm = plane_map<NR>(g, y0 + ylo, x0 + xlo);
  auto step = [&](PlaneRegs<NR>& R, int n) {
    const int zi = zb + zlo + n;
    plane_store(R, pm, sm + (n % kz) * plane, vol, g, zi, y0 + ylo, x0 + xlo);
    // plane of offset dz sits at ring index (n - (zhi - dz)) mod kz; the
    // table is padded to a multiple of 4 with elements that read offset 0
    // and add the identity (they change no min / max)
    const int npad = (tab.n + 3) & ~3;
    for (int t = threadIdx.x; t < npad; t += blockDim.x) {
      if (t < tab.n) {
        const int ridx = ((n - (zhi - tab.dz[t])) % kz + kz) % kz;
        soff[t] = ridx * plane + (tab.dy[t] - ylo) * SX + (tab.dx[t] - xlo);
        sval[t] = tab.v[t];
      } else {
        soff[t] = 0;
        sval[t] = ID;
      }
    }
    __syncthreads();
    plane_fetch(R, vol, g, pm, zi + 2);  // past the last plane: mapped, unused
    const int z = zb + n - (kz - 1);
    if (z >= zb) {
      // four adjacent outputs of one row per lane: each element's offset and
      // value are read once for the four, four independent chains
      const int r = threadIdx.x >> 4, c0 = (threadIdx.x & 15) * 4;
      const float* base = sm + r * SX + c0;
      float a0 = ID, a1 = ID, a2 = ID, a3 = ID;
      for (int t = 0; t < npad; t += 4) {
        const int4 o = *(const int4*)(soff + t);
        const float4 v = *(const float4*)(sval + t);
        const int oo[4] = {o.x, o.y, o.z, o.w};
        const float vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float* b = base + oo[u];
          a0 = gop<MAX>(a0, b[0] + vv[u]);
          a1 = gop<MAX>(a1, b[1] + vv[u]);
          a2 = gop<MAX>(a2, b[2] + vv[u]);
          a3 = gop<MAX>(a3, b[3] + vv[u]);
        }
      }
      const int y = y0 + r;
      const float acc[4] = {a0, a1, a2, a3};
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int x = x0 + c0 + u;
        if (y < H && x < W) {
          const int64_t p = ((int64_t)z * H + y) * W + x;
          ovol[p] = grey_epi<EPI>(acc[u], epi, p);
        }
      }
    }
    __syncthreads();
  };
  PlaneRegs<NR> RA, RB;
  plane_fetch(RA, vol, g, pm, zb + zlo);
  if (nplanes > 1) plane_fetch(RB, vol, g, pm, zb + zlo + 1);
  for (int n = 0; n < nplanes; n += 2) {
    step(RA, n);
    if (n + 1 < nplanes) step(RB, n + 1);
  }
}

template <bool MAX, int NR, int EPI = 0>
__global__ __launch_bounds__(256) void grey_table3_kernel(const float* __restrict__ in,
                                                          float* __restrict__ out, int D, int H,
                                                          int W, GreyTable tab, int zlo, int zhi,
                                                          int ylo, int yhi, int xlo, int xhi,
                                                          int mode, float cval, int zchunk,
                                                          GreyEpi epi) {
  extern __shared__ float sm[];
  __shared__ __align__(16) int soff[kMaxTable];
  __shared__ __align__(16) float sval[kMaxTable];
  const int ntx = (W + kTX - 1) / kTX, nty = (H + kTY - 1) / kTY;
  const int nz = (D + zchunk - 1) / zchunk;
  int tx, ty, zci;
  int64_t item;
  grey_block(ntx, ntx * nty, nz, &tx, &ty, &zci, &item);
  const int zb = zci * zchunk;
  const int64_t n = (int64_t)D * H * W;
  const GreyEpi ei{epi.a ? epi.a + item * n : nullptr, epi.b ? epi.b + item * n : nullptr};
  grey_table3_tile<MAX, NR, EPI>(in + item * n, out + item * n, D, H, W, tab, zlo, zhi, ylo, yhi,
                                 xlo, xhi, mode, cval, tx, ty, zb, min(D, zb + zchunk), sm, soff,
                                 sval, ei);
}

// Ragged lists (no reference counterpart): items of different (D, H, W) at
// their own flat offsets share one launch.  Item k owns the logical blocks
// [block0, block0 + ntiles * nz); a per-block table names the item, so a
// workgroup finds its item with one load instead of a search.  The logical
// block is XCD-swizzled as in the dense kernels.
struct GreyRgItem {
  int64_t in_off, out_off;  // float32 elements
  int D, H, W;
  int ntx, ntiles;  // tiles per row of tiles, per plane
  int block0;
};
constexpr int kRgZChunk = 16;

__device__ __forceinline__ GreyRgItem grey_rg_block(const GreyRgItem* __restrict__ items,
                                                    const int32_t* __restrict__ blk_item, int* tx,
                                                    int* ty, int* zb) {
  const int lb = xcd_swizzle(blockIdx.x, gridDim.x);
  const GreyRgItem it = items[blk_item[lb]];
  const int local = lb - it.block0;
  const int t = local % it.ntiles;
  *tx = t % it.ntx;
  *ty = t / it.ntx;
  *zb = (local / it.ntiles) * kRgZChunk;
  return it;
}

template <bool MAX, int NR>
__global__ __launch_bounds__(256) void grey_box3_ragged_kernel(
    const float* __restrict__ in, float* __restrict__ out, const GreyRgItem* __restrict__ items,
    const int32_t* __restrict__ blk_item, int kz, int ky, int kx, int sz, int sy, int sx, int mode,
    float cval) {
  extern __shared__ float sm[];
  int tx, ty, zb;
  const GreyRgItem it = grey_rg_block(items, blk_item, &tx, &ty, &zb);
  grey_box3_tile<MAX, NR>(in + it.in_off, out + it.out_off, it.D, it.H, it.W, kz, ky, kx, sz, sy,
                          sx, mode, cval, tx, ty, zb, min(it.D, zb + kRgZChunk), sm);
}

template <bool MAX, int NR>
__global__ __launch_bounds__(256) void grey_table3_ragged_kernel(
    const float* __restrict__ in, float* __restrict__ out, const GreyRgItem* __restrict__ items,
    const int32_t* __restrict__ blk_item, GreyTable tab, int zlo, int zhi, int ylo, int yhi,
    int xlo, int xhi, int mode, float cval) {
  extern __shared__ float sm[];
  __shared__ __align__(16) int soff[kMaxTable];
  __shared__ __align__(16) float sval[kMaxTable];
  int tx, ty, zb;
  const GreyRgItem it = grey_rg_block(items, blk_item, &tx, &ty, &zb);
  grey_table3_tile<MAX, NR>(in + it.in_off, out + it.out_off, it.D, it.H, it.W, tab, zlo, zhi,
                            ylo, yhi, xlo, xhi, mode, cval, tx, ty, zb,
                            min(it.D, zb + kRgZChunk), sm, soff, sval);
}

// One 1-D window pass along `axis` (flat SE): out = op over in[p + (start+l)e_axis].
template <bool MAX>
__global__ __launch_bounds__(256) void grey_axis_kernel(const float* __restrict__ in,
                                                        float* __restrict__ out, Shape sh,
                                                        int64_t total, int axis, int k, int start,
                                                        int mode, float cval) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int64_t stride = sh.stride[axis], size = sh.size[axis];
  const int64_t c = (g / stride) % size;
  const int64_t base = g - c * stride;
  float acc = gid<MAX>();
  for (int l = 0; l < k; ++l) {
    bool uc;
    const int64_t m = gmap(c + start + l, size, mode, &uc);
    acc = gop<MAX>(acc, (uc ? cval : in[base + m * stride]) + (MAX ? 0.0f : -0.0f));
  }
  out[g] = acc;
}

// General fallback: table rows are (ndim offsets, value) in device memory.
template <bool MAX>
__global__ __launch_bounds__(256) void grey_gather_kernel(const float* __restrict__ in,
                                                          float* __restrict__ out, Shape sh,
                                                          int64_t total,
                                                          const int32_t* __restrict__ toff,
                                                          const float* __restrict__ tval, int n,
                                                          int mode, float cval) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int nd = sh.ndim;
  const int64_t item = g / sh.numel;
  int64_t rem = g - item * sh.numel;
  int64_t c[kMaxNdim];
  for (int j = 0; j < nd; ++j) {
    c[j] = rem / sh.stride[j];
    rem -= c[j] * sh.stride[j];
  }
  const float* vol = in + item * sh.numel;
  float acc = gid<MAX>();
  for (int t = 0; t < n; ++t) {
    bool use_c = false;
    int64_t flat = 0;
    for (int j = 0; j < nd && !use_c; ++j) {
      bool uc;
      const int64_t m = gmap(c[j] + toff[t * nd + j], sh.size[j], mode, &uc);
      use_c |= uc;
      flat += m * sh.stride[j];
    }
    acc = gop<MAX>(acc, (use_c ? cval : vol[flat]) + tval[t]);
  }
  out[g] = acc;
}

static size_t box3_lds(int kz, int ky, int kx) {
  const size_t SX = kTX + kx - 1, SY = kTY + ky - 1;
  return (SY * SX + SY * kXP + (size_t)kz * kTY * kTX) * sizeof(float);
}

static size_t table3_lds(int kz, int ky, int kx) {
  return (size_t)kz * (kTY + ky - 1) * (kTX + kx - 1) * sizeof(float);
}

constexpr size_t kLdsBudget = 120 * 1024;

// Active elements of a structuring element, in the reference's element order
// (grey_kernel.cu:247-292): offsets o_i (erosion d_i, dilation -d_i) and
// values v_i (erosion -s_i, dilation s_i).  The tables live in per-thread
// storage that the next call reuses.
struct GreyElems {
  int n = 0;
  bool flat = true;  // every value is +0.0
  bool full = true;  // every element of the box is active
  int64_t lo[kMaxNdim], hi[kMaxNdim];  // offset bounds per axis
  const int32_t* off = nullptr;        // n x ndim
  const float* val = nullptr;          // n
};

static int grey_elements(const char* who, int ndim, const float* structure,
                         const uint8_t* footprint, const int64_t* se_shape,
                         const int64_t* origin, int dilation, GreyElems* E) {
  int64_t ne = 1;
  for (int j = 0; j < ndim; ++j) {
    TM_CHECK_ARG(se_shape[j] >= 1, "%s: structuring element axis %d is empty", who, j);
    ne *= se_shape[j];
  }
  // ne x ndim offsets and ne values (sized by both: the same element count
  // with more axes needs a larger offset table)
  static thread_local std::vector<int32_t> hoff_buf;
  static thread_local std::vector<float> hval_buf;
  if (hoff_buf.size() < (size_t)(ne * ndim)) hoff_buf.resize((size_t)(ne * ndim));
  if (hval_buf.size() < (size_t)ne) hval_buf.resize((size_t)ne);
  int32_t* hoff = hoff_buf.data();
  float* hval = hval_buf.data();
  for (int j = 0; j < ndim; ++j) {
    E->lo[j] = INT64_MAX;
    E->hi[j] = INT64_MIN;
  }
  int n = 0;
  for (int64_t e = 0; e < ne; ++e) {
    if (footprint && !footprint[e]) {
      E->full = false;
      continue;
    }
    const float s = structure ? structure[e] : 0.0f;
    if (s != 0.0f || (structure && std::signbit(s))) E->flat = false;
    int64_t rem = e;
    int64_t r[kMaxNdim];
    for (int j = ndim - 1; j >= 0; --j) {
      r[j] = rem % se_shape[j];
      rem /= se_shape[j];
    }
    for (int j = 0; j < ndim; ++j) {
      int64_t d = r[j] - (se_shape[j] / 2 + origin[j]);
      if (dilation) d = -d;
      hoff[n * ndim + j] = (int32_t)d;
      E->lo[j] = d < E->lo[j] ? d : E->lo[j];
      E->hi[j] = d > E->hi[j] ? d : E->hi[j];
    }
    hval[n] = dilation ? s : -s;
    ++n;
  }
  TM_CHECK_ARG(n > 0, "%s: footprint must contain at least one active position", who);
  E->n = n;
  E->off = hoff;
  E->val = hval;
  return TM_OK;
}

// Offset bounds of an element of <= 3 axes, padded to (z, y, x).
struct Grey3 {
  int64_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int kz, ky, kx;
};

static Grey3 grey3_of(int ndim, const GreyElems& E) {
  Grey3 g;
  for (int j = 0; j < ndim; ++j) {
    g.lo[3 - ndim + j] = E.lo[j];
    g.hi[3 - ndim + j] = E.hi[j];
  }
  g.kz = (int)(g.hi[0] - g.lo[0] + 1);
  g.ky = (int)(g.hi[1] - g.lo[1] + 1);
  g.kx = (int)(g.hi[2] - g.lo[2] + 1);
  return g;
}

// The streaming kernels take the element as a flat box (box3) or as a table
// of <= kMaxTable elements with int8 offsets (table3), within the LDS budget.
static bool grey_box3_fits(const GreyElems& E, const Grey3& g) {
  return E.flat && E.full && box3_lds(g.kz, g.ky, g.kx) <= kLdsBudget;
}

static bool grey_table3_fits(const GreyElems& E, const Grey3& g) {
  bool ok = E.n <= kMaxTable;
  for (int j = 0; j < 3; ++j) ok = ok && g.lo[j] >= -128 && g.hi[j] <= 127;
  return ok && table3_lds(g.kz, g.ky, g.kx) <= kLdsBudget;
}

// Register size class of a streaming launch: the halo tile of a plane,
// (kTY + ky - 1) x (kTX + kx - 1) values, in at most 8 per lane.
static bool plane_small(int ky, int kx) {
  return (int64_t)(kTY + ky - 1) * (kTX + kx - 1) <= 256 * kPlaneElemsSmall;
}

// The streaming kernels address a plane with 32-bit offsets.
static bool plane32(int64_t H, int64_t W) { return H * W < ((int64_t)1 << 31); }

static GreyTable grey_table3_of(int ndim, const GreyElems& E) {
  GreyTable tab{};
  tab.n = E.n;
  for (int t = 0; t < E.n; ++t) {
    int32_t o3[3] = {0, 0, 0};
    for (int j = 0; j < ndim; ++j) o3[3 - ndim + j] = E.off[t * ndim + j];
    tab.dz[t] = (int8_t)o3[0];
    tab.dy[t] = (int8_t)o3[1];
    tab.dx[t] = (int8_t)o3[2];
    tab.v[t] = E.val[t];
  }
  return tab;
}

}  // namespace tmh

extern "C" size_t tm_grey_workspace_bytes(int64_t batch, int ndim, const int64_t* shape,
                                          const int64_t* se_shape) {
  if (ndim < 1 || ndim > tmh::kMaxNdim || batch < 1) return 0;
  int64_t n = batch, ne = 1;
  for (int j = 0; j < ndim; ++j) {
    n *= shape[j];
    ne *= se_shape[j];
  }
  // ping-pong volume for the per-axis fallback + an element table
  return (size_t)n * sizeof(float) + 256 + (size_t)ne * (ndim + 1) * sizeof(int32_t);
}

namespace tmh {

template <bool MX, int NR>
static void launch_box3(int epi, unsigned G, size_t lds, hipStream_t st, const float* in,
                        float* out, int D, int H, int W, int kz, int ky, int kx, int sz, int sy,
                        int sx, int mode, float cval, int zchunk, GreyEpi e) {
#define TM_BOX(EP)                                                                               \
  hipLaunchKernelGGL((grey_box3_kernel<MX, NR, EP>), dim3(G), dim3(256), lds, st, in, out, D, H, W, \
                     kz, ky, kx, sz, sy, sx, mode, cval, zchunk, e)
  switch (epi) {
    case 1: TM_BOX(1); break;
    case 2: TM_BOX(2); break;
    case 3: TM_BOX(3); break;
    default: TM_BOX(0); break;
  }
#undef TM_BOX
}

// Register-streaming flat box (grey_boxr_kernel): K x K x K or, for 2-D
// items, K x K, for K = 3, 5, 7; rw_rows(K) output rows per wave (fewer for
// K = 7, whose row buffers and z window fill the registers); epi 4 / 5 the
// one-pass gradient / laplace.  Returns false when the box has another shape.
constexpr int rw_rows(int k) { return k >= 7 ? 4 : 8; }

template <int K>
static void launch_boxr_k(bool mx, int epi, unsigned G, int kz, const float* in, float* out,
                          int D, int H, int W, int sz, int sy, int sx, int mode, float cval,
                          int zchunk, GreyEpi e, hipStream_t st) {
#define TM_RW(MX, KZ, EP)                                                                   \
  hipLaunchKernelGGL((grey_boxr_kernel<MX, K, K, KZ, rw_rows(K), EP>), dim3(G), dim3(256), 0, \
                     st, in, out, D, H, W, sz, sy, sx, mode, cval, zchunk, e)
#define TM_RW_EPI(MX, KZ)            \
  switch (epi) {                     \
    case 1: TM_RW(MX, KZ, 1); break; \
    case 2: TM_RW(MX, KZ, 2); break; \
    case 3: TM_RW(MX, KZ, 3); break; \
    default: TM_RW(MX, KZ, 0); break; \
  }
  if (epi >= 4) {  // one-pass gradient / laplace
    if (kz == K) {
      if (epi == 4) TM_RW(true, K, 4); else TM_RW(true, K, 5);
    } else {
      if (epi == 4) TM_RW(true, 1, 4); else TM_RW(true, 1, 5);
    }
  } else if (mx) {
    if (kz == K) TM_RW_EPI(true, K) else TM_RW_EPI(true, 1)
  } else {
    if (kz == K) TM_RW_EPI(false, K) else TM_RW_EPI(false, 1)
  }
#undef TM_RW_EPI
#undef TM_RW
}

static bool launch_boxr(bool mx, int epi, int64_t batch, int D, int H, int W, int kz, int ky,
                        int kx, int sz, int sy, int sx, int mode, float cval, int zchunk,
                        GreyEpi e, const float* in, float* out, hipStream_t st) {
  if (kx != ky || (kz != kx && kz != 1) || (kx != 3 && kx != 5 && kx != 7)) return false;
  if ((int64_t)H * W > ((int64_t)1 << 30)) return false;  // 32-bit byte offsets in a plane
  const int rows = rw_rows(kx);
  const int64_t ntx = (W + 63) / 64, nty = (H + rows - 1) / rows;
  const int64_t nyb = (nty + kRwWaves - 1) / kRwWaves, nz = (D + zchunk - 1) / zchunk;
  const int64_t G = ntx * nyb * nz * batch;
  if (G >= ((int64_t)1 << 31)) return false;
  const unsigned g = (unsigned)G;
  auto go = [&](auto kk) {
    launch_boxr_k<decltype(kk)::value>(mx, epi, g, kz, in, out, D, H, W, sz, sy, sx, mode, cval,
                                       zchunk, e, st);
  };
  if (kx == 3) go(std::integral_constant<int, 3>{});
  else if (kx == 5) go(std::integral_constant<int, 5>{});
  else go(std::integral_constant<int, 7>{});
  return true;
}

// The ball of radius 3 with heights floor(3 - r) on 3-D items (grey_ballr_kernel):
// 123 elements at every offset with r^2 <= 9, value h (dilation) or -h
// (erosion, -0.0f for h = 0), h = 3, 2, 1, 0 for r^2 = 0, 1, 2..4, 5..9.
static bool grey_is_ball3(int ndim, const GreyElems& E, bool is_max) {
  if (ndim != 3 || E.n != 123) return false;
  for (int t = 0; t < E.n; ++t) {
    const int32_t* o = E.off + 3 * t;
    const int r2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
    if (r2 > 9) return false;
    const int h = r2 == 0 ? 3 : r2 == 1 ? 2 : r2 <= 4 ? 1 : 0;
    const float want = is_max ? (float)h : (h == 0 ? -0.0f : -(float)h);
    uint32_t a, b;
    memcpy(&a, &E.val[t], 4);
    memcpy(&b, &want, 4);
    if (a != b) return false;
  }
  return true;
}

static bool launch_ballr(bool mx, int epi, int64_t batch, int D, int H, int W, int mode,
                         float cval, int zchunk, GreyEpi e, const float* in, float* out,
                         hipStream_t st) {
  if (epi < 0 || epi > 2) return false;
  if ((int64_t)H * W > ((int64_t)1 << 30)) return false;  // 32-bit byte offsets in a plane
  const int64_t ntx = (W + 63) / 64, nty = (H + kBallRows - 1) / kBallRows;
  const int64_t nyb = (nty + kRwWaves - 1) / kRwWaves, nz = (D + zchunk - 1) / zchunk;
  const int64_t G = ntx * nyb * nz * batch;
  if (G >= ((int64_t)1 << 31)) return false;
  if (mx)
    hipLaunchKernelGGL(grey_ballr_kernel<true>, dim3((unsigned)G), dim3(256), 0, st, in, out, D,
                       H, W, mode, cval, zchunk, e, epi);
  else
    hipLaunchKernelGGL(grey_ballr_kernel<false>, dim3((unsigned)G), dim3(256), 0, st, in, out, D,
                       H, W, mode, cval, zchunk, e, epi);
  return true;
}

template <bool MX, int NR>
static void launch_table3(int epi, unsigned G, size_t lds, hipStream_t st, const float* in,
                          float* out, int D, int H, int W, const GreyTable& tab, const Grey3& g3,
                          int mode, float cval, int zchunk, GreyEpi e) {
#define TM_TAB(EP)                                                                                 \
  hipLaunchKernelGGL((grey_table3_kernel<MX, NR, EP>), dim3(G), dim3(256), lds, st, in, out, D, H, \
                     W, tab, (int)g3.lo[0], (int)g3.hi[0], (int)g3.lo[1], (int)g3.hi[1],          \
                     (int)g3.lo[2], (int)g3.hi[2], mode, cval, zchunk, e)
  switch (epi) {
    case 1: TM_TAB(1); break;
    case 2: TM_TAB(2); break;
    case 3: TM_TAB(3); break;
    default: TM_TAB(0); break;
  }
#undef TM_TAB
}

static int grey_run(const float* input, int64_t batch, int ndim, const int64_t* shape,
                    const float* structure, const uint8_t* footprint, const int64_t* se_shape,
                    const int64_t* origin, int mode, float cval, int dilation, int epi,
                    const float* epi_a, const float* epi_b, float* output, void* workspace,
                    size_t workspace_bytes, hipStream_t st) {
  TM_CHECK_ARG(input && output, "tm_grey: input/output is NULL");
  TM_CHECK_ARG(input != output, "tm_grey: output must not alias input");
  TM_CHECK_ARG(ndim >= 1 && ndim <= kMaxNdim, "tm_grey: ndim must be in 1..8, got %d", ndim);
  TM_CHECK_ARG(mode >= 0 && mode <= 4, "tm_grey: mode must be in 0..4");
  TM_CHECK_ARG(batch >= 1, "tm_grey: batch must be >= 1");
  const GreyEpi ge{epi_a, epi_b};
  Shape sh = make_shape(ndim, shape);
  const int64_t total = batch * sh.numel;
  for (int j = 0; j < ndim; ++j) TM_CHECK_ARG(shape[j] >= 1, "tm_grey: input axis %d is empty", j);
  GreyElems E;
  {
    const int rc = grey_elements("tm_grey", ndim, structure, footprint, se_shape, origin,
                                 dilation, &E);
    if (rc != TM_OK) return rc;
  }
  const int n = E.n;
  const bool flat = E.flat, full = E.full;
  const int32_t* hoff = E.off;
  const float* hval = E.val;
  const int64_t* lo = E.lo;
  const int64_t* hi = E.hi;

  const bool is_max = dilation != 0;
  if (ndim <= 3) {
    int64_t s3[3] = {1, 1, 1};
    for (int j = 0; j < ndim; ++j) s3[3 - ndim + j] = shape[j];
    const Grey3 g3 = grey3_of(ndim, E);
    const int64_t* lo3 = g3.lo;
    const int64_t* hi3 = g3.hi;
    const int D = (int)s3[0], H = (int)s3[1], W = (int)s3[2];
    const int kz = g3.kz, ky = g3.ky, kx = g3.kx;
    const int64_t ntiles = ((W + kTX - 1) / kTX) * (int64_t)((H + kTY - 1) / kTY);
    // launch variant v: z planes per workgroup = 8 << v (capped at D); the
    // autotuner trades workgroup count against the kz - 1 halo planes every
    // chunk re-reads
    auto grid_of = [&](int zchunk) -> int64_t {
      return ntiles * ((D + zchunk - 1) / zchunk) * batch;
    };
    auto zchunk_of = [&](int v) { return std::min(D, 8 << v); };
    auto p2 = [](int64_t v) { int64_t q = 1; while (q < v) q <<= 1; return q; };
    if (grey_box3_fits(E, g3) && plane32(H, W)) {
      const size_t lds = box3_lds(kz, ky, kx);
      // variants 4..7: the register-streaming kernel for K x K (x K) boxes,
      // K = 3, 5, 7, with z chunks of 8 << (v - 4) planes; not applicable to
      // other boxes
      auto run = [&](int v) -> int {
        if (v >= 4) {
          const int zc = std::min(D, 8 << (v - 4));
          if (!launch_boxr(is_max, epi, batch, D, H, W, kz, ky, kx, (int)lo3[0], (int)lo3[1],
                           (int)lo3[2], mode, cval, zc, ge, input, output, st))
            return TM_ERR_ARG;
          TM_LAUNCH_CHECK();
          return TM_OK;
        }
        const int zchunk = zchunk_of(v);
        const int64_t G = grid_of(zchunk);
        if (G >= ((int64_t)1 << 31)) {
          set_error("tm_grey: batch too large for one launch");
          return TM_ERR_ARG;
        }
        const bool small_se = plane_small(ky, kx);
#define TM_BOX(MX, NR)                                                                         \
  launch_box3<MX, NR>(epi, (unsigned)G, lds, st, input, output, D, H, W, kz, ky, kx, (int)lo3[0], \
                      (int)lo3[1], (int)lo3[2], mode, cval, zchunk, ge)
        if (is_max) {
          if (small_se) TM_BOX(true, kPlaneElemsSmall);
          else TM_BOX(true, kPlaneElemsLarge);
        } else {
          if (small_se) TM_BOX(false, kPlaneElemsSmall);
          else TM_BOX(false, kPlaneElemsLarge);
        }
#undef TM_BOX
        TM_LAUNCH_CHECK();
        return TM_OK;
      };
      const int64_t key[8] = {1, kz, ky, kx, p2(D), p2(H), p2(W), p2(batch)};
      bool ran = false;
      const int v = tune_pick("grey3", key, 8, 8, 1, run, st, &ran);
      if (ran) return TM_OK;
      const int rc = run(v);
      return (rc == TM_ERR_ARG && v != 1) ? run(1) : rc;  // a cached variant that does not apply
    }
    if (grey_table3_fits(E, g3) && plane32(H, W)) {
      const GreyTable tab = grey_table3_of(ndim, E);
      const size_t lds = table3_lds(kz, ky, kx);
      const bool ball3 = grey_is_ball3(ndim, E, is_max);
      // variants 4..7: the ball of radius 3 with heights floor(3 - r)
      // (grey_ballr_kernel), z chunks of 8 << (v - 4) planes; not applicable
      // to other elements
      auto run = [&](int v) -> int {
        if (v >= 4) {
          if (!ball3 || !launch_ballr(is_max, epi, batch, D, H, W, mode, cval,
                                      std::min(D, 8 << (v - 4)), ge, input, output, st))
            return TM_ERR_ARG;
          TM_LAUNCH_CHECK();
          return TM_OK;
        }
        const int zchunk = zchunk_of(v);
        const int64_t G = grid_of(zchunk);
        if (G >= ((int64_t)1 << 31)) {
          set_error("tm_grey: batch too large for one launch");
          return TM_ERR_ARG;
        }
        const bool small_se = plane_small(ky, kx);
#define TM_TAB(MX, NR)                                                                        \
  launch_table3<MX, NR>(epi, (unsigned)G, lds, st, input, output, D, H, W, tab, g3, mode, cval, \
                        zchunk, ge)
        if (is_max) {
          if (small_se) TM_TAB(true, kPlaneElemsSmall);
          else TM_TAB(true, kPlaneElemsLarge);
        } else {
          if (small_se) TM_TAB(false, kPlaneElemsSmall);
          else TM_TAB(false, kPlaneElemsLarge);
        }
#undef TM_TAB
        TM_LAUNCH_CHECK();
        return TM_OK;
      };
      const int64_t key[8] = {2, kz, ky, kx, p2(D), p2(H), p2(W), p2(batch)};
      bool ran = false;
      const int v = tune_pick("grey3", key, 8, 8, 1, run, st, &ran);
      if (ran) return TM_OK;
      const int rc = run(v);
      return (rc == TM_ERR_ARG && v != 1) ? run(1) : rc;  // a cached variant that does not apply
    }
  }
  TM_CHECK_ARG(epi == 0,
               "tm_grey_ex: a composite epilogue needs items of at most 3 axes and a "
               "structuring element of the streaming kernels (tm_grey_fused_ok), and planes of "
               "at most 2^31 - 1 voxels");
  const size_t need = tm_grey_workspace_bytes(batch, ndim, shape, se_shape);
  if (!workspace || workspace_bytes < need) {
    set_error("tm_grey: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, need);
    return TM_ERR_WORKSPACE;
  }
  if (flat && full) {
    // separable passes: axis 0 .. ndim-1, alternating workspace / output so
    // that the last pass lands in output
    float* tmp = (float*)workspace;
    const float* src = input;
    for (int j = 0; j < ndim; ++j) {
      float* dst = ((ndim - 1 - j) % 2 == 0) ? output : tmp;
      const int start = (int)lo[j];
      const int k = (int)(hi[j] - lo[j] + 1);
      if (is_max)
        hipLaunchKernelGGL((grey_axis_kernel<true>), dim3(grid_for(total, 256)), dim3(256), 0, st,
                           src, dst, sh, total, j, k, start, mode, cval);
      else
        hipLaunchKernelGGL((grey_axis_kernel<false>), dim3(grid_for(total, 256)), dim3(256), 0, st,
                           src, dst, sh, total, j, k, start, mode, cval);
      TM_LAUNCH_CHECK();
      src = dst;
    }
    return TM_OK;
  }
  int32_t* doff = (int32_t*)(((uintptr_t)workspace + (size_t)total * sizeof(float) + 255) &
                             ~(uintptr_t)255);
  float* dval = (float*)(doff + (size_t)n * ndim);
  TM_CHECK_HIP(hipMemcpyAsync(doff, hoff, sizeof(int32_t) * n * ndim, hipMemcpyHostToDevice, st));
  TM_CHECK_HIP(hipMemcpyAsync(dval, hval, sizeof(float) * n, hipMemcpyHostToDevice, st));
  if (is_max)
    hipLaunchKernelGGL((grey_gather_kernel<true>), dim3(grid_for(total, 256)), dim3(256), 0, st,
                       input, output, sh, total, doff, dval, n, mode, cval);
  else
    hipLaunchKernelGGL((grey_gather_kernel<false>), dim3(grid_for(total, 256)), dim3(256), 0, st,
                       input, output, sh, total, doff, dval, n, mode, cval);
  TM_LAUNCH_CHECK();
  // the pageable copies above are staged synchronously, so the host table
  // may be reused once this call returns
  return TM_OK;
}

}  // namespace tmh

extern "C" int tm_grey(const float* input, int64_t batch, int ndim, const int64_t* shape,
                       const float* structure, const uint8_t* footprint, const int64_t* se_shape,
                       const int64_t* origin, int mode, float cval, int dilation, float* output,
                       void* workspace, size_t workspace_bytes, void* stream) {
  return tmh::grey_run(input, batch, ndim, shape, structure, footprint, se_shape, origin, mode,
                       cval, dilation, 0, nullptr, nullptr, output, workspace, workspace_bytes,
                       (hipStream_t)stream);
}

extern "C" int tm_grey_fused_ok(int ndim, const float* structure, const uint8_t* footprint,
                                const int64_t* se_shape, const int64_t* origin, int dilation) {
  using namespace tmh;
  if (ndim < 1 || ndim > 3 || !se_shape || !origin) return 0;
  GreyElems E;
  if (grey_elements("tm_grey_fused_ok", ndim, structure, footprint, se_shape, origin, dilation,
                    &E) != TM_OK)
    return 0;
  const Grey3 g3 = grey3_of(ndim, E);
  return (grey_box3_fits(E, g3) || grey_table3_fits(E, g3)) ? 1 : 0;
}

namespace tmh {

static size_t box3_dual_lds(int kz, int ky, int kx) {
  const size_t SX = kTX + kx - 1, SY = kTY + ky - 1;
  return (SY * SX + 2 * SY * kTX + 2 * (size_t)kz * kTY * kTX) * sizeof(float);
}

// Whether gradient / laplace of this element runs in one pass
// (grey_box3_dual_kernel): a flat box of odd sizes with zero origin, within
// the LDS budget.
static bool grey_dual_fits(int ndim, const GreyElems& E, const Grey3& g3) {
  if (!(E.flat && E.full) || ndim > 3) return false;
  for (int j = 0; j < 3; ++j)
    if (g3.lo[j] != -g3.hi[j]) return false;
  return box3_dual_lds(g3.kz, g3.ky, g3.kx) <= kLdsBudget;
}

static int grey_dual_run(const float* input, int64_t batch, int ndim, const int64_t* shape,
                         const int64_t* se_shape, const int64_t* origin, int mode, float cval,
                         int comb, float* output, hipStream_t st) {
  TM_CHECK_ARG(input && output && input != output, "tm_grey_ex: input/output invalid");
  TM_CHECK_ARG(mode >= 0 && mode <= 4, "tm_grey_ex: mode must be in 0..4");
  TM_CHECK_ARG(batch >= 1, "tm_grey_ex: batch must be >= 1");
  for (int j = 0; j < ndim; ++j) TM_CHECK_ARG(shape[j] >= 1, "tm_grey_ex: input axis %d is empty", j);
  GreyElems E;
  int rc = grey_elements("tm_grey_ex", ndim, nullptr, nullptr, se_shape, origin, 1, &E);
  if (rc != TM_OK) return rc;
  const Grey3 g3 = grey3_of(ndim, E);
  TM_CHECK_ARG(grey_dual_fits(ndim, E, g3),
               "tm_grey_ex: one-pass gradient / laplace needs a flat box of odd sizes with zero "
               "origin (tm_grey_dual_ok)");
  int64_t s3[3] = {1, 1, 1};
  for (int j = 0; j < ndim; ++j) s3[3 - ndim + j] = shape[j];
  const int D = (int)s3[0], H = (int)s3[1], W = (int)s3[2];
  TM_CHECK_ARG(plane32(H, W), "tm_grey_ex: planes of at most 2^31 - 1 voxels");
  const int kz = g3.kz, ky = g3.ky, kx = g3.kx;
  const int64_t ntiles = ((W + kTX - 1) / kTX) * (int64_t)((H + kTY - 1) / kTY);
  const size_t lds = box3_dual_lds(kz, ky, kx);
  const bool small_se = plane_small(ky, kx);
  // variants 4..7: the register-streaming kernel (K x K (x K) boxes, K = 3,
  // 5, 7), z chunks as grey_run's
  auto run = [&](int v) -> int {
    if (v >= 4) {
      const int zc = std::min(D, 8 << (v - 4));
      if (!launch_boxr(true, 4 + comb, batch, D, H, W, kz, ky, kx, (int)g3.lo[0],
                       (int)g3.lo[1], (int)g3.lo[2], mode, cval, zc, GreyEpi{}, input, output,
                       st))
        return TM_ERR_ARG;
      TM_LAUNCH_CHECK();
      return TM_OK;
    }
    const int zchunk = std::min(D, 8 << v);
    const int64_t G = ntiles * ((D + zchunk - 1) / zchunk) * batch;
    if (G >= ((int64_t)1 << 31)) {
      set_error("tm_grey_ex: batch too large for one launch");
      return TM_ERR_ARG;
    }
#define TM_DUAL(NR, CB)                                                                          \
  hipLaunchKernelGGL((grey_box3_dual_kernel<NR, CB>), dim3((unsigned)G), dim3(256), lds, st, input, \
                     output, D, H, W, kz, ky, kx, (int)g3.lo[0], (int)g3.lo[1], (int)g3.lo[2],      \
                     mode, cval, zchunk)
    if (comb == 0) {
      if (small_se) TM_DUAL(kPlaneElemsSmall, 0);
      else TM_DUAL(kPlaneElemsLarge, 0);
    } else {
      if (small_se) TM_DUAL(kPlaneElemsSmall, 1);
      else TM_DUAL(kPlaneElemsLarge, 1);
    }
#undef TM_DUAL
    TM_LAUNCH_CHECK();
    return TM_OK;
  };
  auto p2 = [](int64_t v) { int64_t q = 1; while (q < v) q <<= 1; return q; };
  const int64_t key[8] = {3, kz, ky, kx, p2(D), p2(H), p2(W), p2(batch)};
  bool ran = false;
  const int v = tune_pick("grey3", key, 8, 8, 1, run, st, &ran);
  if (ran) return TM_OK;
  const int rc2 = run(v);
  return (rc2 == TM_ERR_ARG && v != 1) ? run(1) : rc2;  // a cached variant that does not apply
}

}  // namespace tmh

extern "C" int tm_grey_dual_ok(int ndim, const int64_t* se_shape, const int64_t* origin) {
  using namespace tmh;
  if (ndim < 1 || ndim > 3 || !se_shape || !origin) return 0;
  GreyElems E;
  if (grey_elements("tm_grey_dual_ok", ndim, nullptr, nullptr, se_shape, origin, 1, &E) != TM_OK)
    return 0;
  return grey_dual_fits(ndim, E, grey3_of(ndim, E)) ? 1 : 0;
}

extern "C" int tm_grey_ex(const float* input, int64_t batch, int ndim, const int64_t* shape,
                          const float* structure, const uint8_t* footprint,
                          const int64_t* se_shape, const int64_t* origin, int mode, float cval,
                          int dilation, int epilogue, const float* a, const float* b,
                          float* output, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(epilogue >= 1 && epilogue <= 5, "tm_grey_ex: epilogue must be in 1..5");
  TM_CHECK_ARG(ndim >= 1 && ndim <= 3, "tm_grey_ex: items must have 1..3 axes");
  if (epilogue >= 4) {
    TM_CHECK_ARG(!structure && !footprint,
                 "tm_grey_ex: one-pass gradient / laplace takes a flat box (no structure or "
                 "footprint)");
    return grey_dual_run(input, batch, ndim, shape, se_shape, origin, mode, cval, epilogue - 4,
                         output, (hipStream_t)stream);
  }
  TM_CHECK_ARG(a != nullptr && (epilogue != 3 || b != nullptr),
               "tm_grey_ex: epilogue operand is NULL");
  TM_CHECK_ARG(a != output && b != output,
               "tm_grey_ex: epilogue operands must not alias the output");
  return grey_run(input, batch, ndim, shape, structure, footprint, se_shape, origin, mode, cval,
                  dilation, epilogue, a, b, output, nullptr, 0, (hipStream_t)stream);
}

// ------------------------------------------------------- ragged lists --
namespace tmh {

struct GreyRgPlan {
  std::vector<GreyRgItem> items;
  int nd = 0;
  int64_t blocks = 0;
  size_t items_bytes = 0, total = 0;
};

static int grey_rg_plan(const tm_ragged_job* jobs, int64_t njobs, GreyRgPlan& p) {
  TM_CHECK_ARG(jobs != nullptr, "tm_ragged_grey: jobs is NULL");
  TM_CHECK_ARG(njobs >= 1 && njobs < ((int64_t)1 << 30), "tm_ragged_grey: njobs must be in 1..2^30");
  const int nd = jobs[0].ndim;
  TM_CHECK_ARG(nd >= 1 && nd <= 3, "tm_ragged_grey: items must have 1..3 spatial axes, got %d",
               nd);
  p.nd = nd;
  p.items.resize((size_t)njobs);
  int64_t blocks = 0;
  for (int64_t k = 0; k < njobs; ++k) {
    const tm_ragged_job& j = jobs[k];
    TM_CHECK_ARG(j.ndim == nd,
                 "tm_ragged_grey: job %lld has %d axes but job 0 has %d (one structuring element "
                 "serves every item)",
                 (long long)k, j.ndim, nd);
    TM_CHECK_ARG(j.in_offset >= 0 && j.out_offset >= 0,
                 "tm_ragged_grey: job %lld has a negative offset", (long long)k);
    int64_t s3[3] = {1, 1, 1};
    for (int a = 0; a < nd; ++a) {
      TM_CHECK_ARG(j.shape[a] >= 1 && j.shape[a] <= INT32_MAX,
                   "tm_ragged_grey: job %lld axis %d has size %lld", (long long)k, a,
                   (long long)j.shape[a]);
      s3[3 - nd + a] = j.shape[a];
    }
    GreyRgItem& it = p.items[(size_t)k];
    it.in_off = j.in_offset;
    it.out_off = j.out_offset;
    it.D = (int)s3[0];
    it.H = (int)s3[1];
    it.W = (int)s3[2];
    TM_CHECK_ARG(plane32(it.H, it.W), "tm_ragged_grey: planes of at most 2^31 - 1 voxels");
    const int64_t ntx = (s3[2] + kTX - 1) / kTX, nty = (s3[1] + kTY - 1) / kTY;
    const int64_t nz = (s3[0] + kRgZChunk - 1) / kRgZChunk;
    TM_CHECK_ARG(ntx * nty * nz < ((int64_t)1 << 31) - blocks,
                 "tm_ragged_grey: more tiles than one launch holds");
    it.ntx = (int)ntx;
    it.ntiles = (int)(ntx * nty);
    it.block0 = (int)blocks;
    blocks += ntx * nty * nz;
  }
  p.blocks = blocks;
  p.items_bytes = align256(p.items.size() * sizeof(GreyRgItem));
  p.total = p.items_bytes + align256((size_t)blocks * sizeof(int32_t));
  return TM_OK;
}

static thread_local PinnedStage g_grey_stage;

}  // namespace tmh

extern "C" size_t tm_ragged_grey_workspace_bytes(const tm_ragged_job* jobs, int64_t njobs) {
  tmh::GreyRgPlan p;
  return tmh::grey_rg_plan(jobs, njobs, p) == TM_OK ? p.total : 0;
}

extern "C" int tm_ragged_grey(const tm_ragged_job* jobs, int64_t njobs, const float* input,
                              const float* structure, const uint8_t* footprint,
                              const int64_t* se_shape, const int64_t* origin, int mode,
                              float cval, int dilation, float* output, int upload,
                              void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(input && output, "tm_ragged_grey: input/output is NULL");
  TM_CHECK_ARG(input != output, "tm_ragged_grey: output must not alias input");
  TM_CHECK_ARG(mode >= 0 && mode <= 4, "tm_ragged_grey: mode must be in 0..4");
  GreyRgPlan p;
  int rc = grey_rg_plan(jobs, njobs, p);
  if (rc != TM_OK) return rc;
  if (!workspace || workspace_bytes < p.total) {
    set_error("tm_ragged_grey: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, p.total);
    return TM_ERR_WORKSPACE;
  }
  TM_CHECK_ARG(((uintptr_t)workspace % 256) == 0,
               "tm_ragged_grey: workspace must be 256-byte aligned");
  const int nd = p.nd;
  GreyElems E;
  rc = grey_elements("tm_ragged_grey", nd, structure, footprint, se_shape, origin, dilation, &E);
  if (rc != TM_OK) return rc;
  const Grey3 g3 = grey3_of(nd, E);
  const bool box = grey_box3_fits(E, g3);
  TM_CHECK_ARG(box || grey_table3_fits(E, g3),
               "tm_ragged_grey: the structuring element fits neither streaming kernel (a flat "
               "box within the LDS budget, or at most %d active elements with offsets in "
               "-128..127)",
               kMaxTable);
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;
  const GreyRgItem* items = (const GreyRgItem*)ws;
  const int32_t* blk_item = (const int32_t*)(ws + p.items_bytes);
  if (upload) {
    const size_t need = p.items_bytes + (size_t)p.blocks * sizeof(int32_t);
    PinnedStage& s = g_grey_stage;
    rc = stage_acquire(s, need);
    if (rc != TM_OK) return rc;
    char* h = (char*)s.host;
    memcpy(h, p.items.data(), p.items.size() * sizeof(GreyRgItem));
    int32_t* hb = (int32_t*)(h + p.items_bytes);
    for (size_t k = 0; k < p.items.size(); ++k) {
      const int64_t b1 = k + 1 < p.items.size() ? p.items[k + 1].block0 : p.blocks;
      for (int64_t b = p.items[k].block0; b < b1; ++b) hb[b] = (int32_t)k;
    }
    rc = stage_upload(s, ws, need, st);
    if (rc != TM_OK) return rc;
  }
  const unsigned G = (unsigned)p.blocks;
  const bool is_max = dilation != 0;
  const int kz = g3.kz, ky = g3.ky, kx = g3.kx;
  if (box) {
    const size_t lds = box3_lds(kz, ky, kx);
    const bool small_se = plane_small(ky, kx);
#define TM_BOX(MX, NR)                                                                            \
  hipLaunchKernelGGL((grey_box3_ragged_kernel<MX, NR>), dim3(G), dim3(256), lds, st, input, output, \
                     items, blk_item, kz, ky, kx, (int)g3.lo[0], (int)g3.lo[1], (int)g3.lo[2],      \
                     mode, cval)
    if (is_max) {
      if (small_se) TM_BOX(true, kPlaneElemsSmall);
      else TM_BOX(true, kPlaneElemsLarge);
    } else {
      if (small_se) TM_BOX(false, kPlaneElemsSmall);
      else TM_BOX(false, kPlaneElemsLarge);
    }
#undef TM_BOX
  } else {
    const GreyTable tab = grey_table3_of(nd, E);
    const size_t lds = table3_lds(kz, ky, kx);
    const bool small_se = plane_small(ky, kx);
#define TM_TAB(MX, NR)                                                                             \
  hipLaunchKernelGGL((grey_table3_ragged_kernel<MX, NR>), dim3(G), dim3(256), lds, st, input,       \
                     output, items, blk_item, tab, (int)g3.lo[0], (int)g3.hi[0], (int)g3.lo[1],     \
                     (int)g3.hi[1], (int)g3.lo[2], (int)g3.hi[2], mode, cval)
    if (is_max) {
      if (small_se) TM_TAB(true, kPlaneElemsSmall);
      else TM_TAB(true, kPlaneElemsLarge);
    } else {
      if (small_se) TM_TAB(false, kPlaneElemsSmall);
      else TM_TAB(false, kPlaneElemsLarge);
    }
#undef TM_TAB
  }
  TM_LAUNCH_CHECK();
  return TM_OK;
}

namespace tmh {

// ------------------------------------------------------------- backward --
// Hand-written backward of grey erosion / dilation (no reference
// counterpart: the reference ops are not differentiable, test_grey.py:218-222).
// Unified form y[p] = OP_i E(p + o_i) + v_i (erosion: OP = min, o = d,
// v = -s; dilation: OP = max, o = -d, v = s).  The gradient of y[p] goes to
// the input voxel read by the FIRST element (table order = the reference's
// element order) attaining the extremum; elements reading cval route nothing.
// Two kernels, no atomics: arg[p] = winning element, then every input voxel
// gathers dy[p] over the outputs whose winning element reads it, in a fixed
// order (deterministic).

struct GreyTabN {
  int n;
  int32_t ext[kMaxNdim];  // max |offset| per axis
  int8_t o[kMaxTable][kMaxNdim];
  float v[kMaxTable];
  int32_t flat[kMaxTable];  // sum_j o[t][j] * stride[j]: the element's offset inside an item
};

// Coordinates of element g of a batch; true when every read at an offset of
// the table (|o_j| <= ext_j + pad) stays inside the item.
__device__ __forceinline__ bool grey_decode(const Shape& sh, int64_t g, const int32_t* ext, int pad,
                                            int64_t* item, int64_t* local, int64_t* c) {
  *item = g / sh.numel;
  int64_t rem = g - *item * sh.numel;
  *local = rem;
  bool interior = true;
  for (int j = 0; j < sh.ndim; ++j) {
    c[j] = rem / sh.stride[j];
    rem -= c[j] * sh.stride[j];
    if (c[j] < ext[j] + pad || c[j] > sh.size[j] - 1 - ext[j] - pad) interior = false;
  }
  return interior;
}

template <bool MAX>
__global__ __launch_bounds__(256) void grey_arg_kernel(const float* __restrict__ in,
                                                       int16_t* __restrict__ arg, Shape sh,
                                                       int64_t total, GreyTabN tab, int mode,
                                                       float cval) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int nd = sh.ndim;
  int64_t item, local, c[kMaxNdim];
  const bool interior = grey_decode(sh, g, tab.ext, 0, &item, &local, c);
  const float* vol = in + item * sh.numel;
  float best = 0.f;
  int bi = 0;
  if (interior) {
    // every read in bounds: one load per element at a precomputed offset
    const float* vp = vol + local;
    for (int t = 0; t < tab.n; ++t) {
      const float v = vp[tab.flat[t]] + tab.v[t];
      if (t == 0 || (MAX ? v > best : v < best)) {
        best = v;
        bi = t;
      }
    }
    arg[g] = (int16_t)bi;
    return;
  }
  for (int t = 0; t < tab.n; ++t) {
    bool use_c = false;
    int64_t flat = 0;
    for (int j = 0; j < nd && !use_c; ++j) {
      bool uc;
      const int64_t m = gmap(c[j] + tab.o[t][j], sh.size[j], mode, &uc);
      use_c |= uc;
      flat += m * sh.stride[j];
    }
    const float v = (use_c ? cval : vol[flat]) + tab.v[t];
    if (t == 0 || (MAX ? v > best : v < best)) {
      best = v;
      bi = t;
    }
  }
  arg[g] = (int16_t)bi;
}

__global__ __launch_bounds__(256) void grey_grad_gather_kernel(
    const int16_t* __restrict__ arg, const float* __restrict__ dy, float* __restrict__ dx,
    Shape sh, int64_t total, GreyTabN tab, int mode) {
  const int32_t* ext = tab.ext;
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int nd = sh.ndim;
  int64_t item, local, q[kMaxNdim];
  // mirror maps a read at -k onto k (not k-1): its border zone is one wider
  const bool interior = grey_decode(sh, g, ext, 1, &item, &local, q);
  const int64_t base = item * sh.numel;
  float acc = 0.f;
  if (interior) {
    // every output reading q through element t is p = q - o_t, in bounds
    const int16_t* ap = arg + g;
    const float* dp = dy + g;
    for (int t = 0; t < tab.n; ++t) {
      const int32_t off = tab.flat[t];
      if (ap[-off] == t) acc += dp[-off];
    }
  } else {
    // border voxel: scan the outputs within the element extent of q (per
    // axis q-ext..q+ext, wrapped around for mode wrap, clipped otherwise), in
    // raster order; each output routes to exactly one input, checked here
    int64_t cnt[kMaxNdim], k[kMaxNdim];
    bool wrapd[kMaxNdim];
    for (int j = 0; j < nd; ++j) {
      const int64_t n = sh.size[j], e = ext[j] + 1;
      wrapd[j] = (mode == 4) && (2 * e + 1 < n);
      cnt[j] = (mode == 4) ? (wrapd[j] ? 2 * e + 1 : n)
                           : (q[j] + e > n - 1 ? n - 1 : q[j] + e) - (q[j] - e < 0 ? 0 : q[j] - e) + 1;
      k[j] = 0;
    }
    for (;;) {
      int64_t flat = 0;
      int64_t p[kMaxNdim];
      for (int j = 0; j < nd; ++j) {
        const int64_t n = sh.size[j], e = ext[j] + 1;
        if (mode == 4) p[j] = wrapd[j] ? (((q[j] + k[j] - e) % n) + n) % n : k[j];
        else p[j] = (q[j] - e < 0 ? 0 : q[j] - e) + k[j];
        flat += p[j] * sh.stride[j];
      }
      const int t = arg[base + flat];
      bool hit = true;
      for (int j = 0; j < nd && hit; ++j) {
        bool uc;
        const int64_t m = gmap(p[j] + tab.o[t][j], sh.size[j], mode, &uc);
        hit = !uc && m == q[j];
      }
      if (hit) acc += dy[base + flat];
      int j = nd - 1;
      for (; j >= 0; --j) {
        if (++k[j] < cnt[j]) break;
        k[j] = 0;
      }
      if (j < 0) break;
    }
  }
  dx[g] = acc;
}

// Fused backward for 2-D items (modes other than wrap): a workgroup owns a
// 64 x 16 tile of inputs q.  The outputs p that can route into the tile lie
// within e' = ext + 1 of it (an element reads p + o; a mode maps a read
// outside the item back within ext + 1 of the border); their winning
// elements are recomputed from x staged with a further ext halo (the
// forward's arithmetic and first-element tie rule), and their dy staged
// beside them.  Input q then sums, in element order t, dy[p] over the
// outputs p = q' - o_t whose winning element is t, for every q' that the
// mode maps to q (q itself, plus reflections of border positions); an
// interior q has only q' = q, so its sum is grey_grad_gather_kernel's, term
// for term.
constexpr int kBwTX = 64, kBwTY = 16;

// Candidate i of the positions the mode may map to c: i = 0 is c itself,
// i = 1 .. e are -e .. -1, i = e + 1 .. 2e are n .. n - 1 + e.
__device__ __forceinline__ int grey_pre_cand(int i, int c, int n, int e) {
  return i == 0 ? c : (i <= e ? i - 1 - e : n + i - e - 1);
}

// Bit i set when candidate i maps to c (c itself always; constant mode maps
// nothing from outside the item).
__device__ __forceinline__ uint64_t grey_preimages(int c, int n, int e, int mode) {
  uint64_t m = 1;
  if (mode == 0 || (c >= e && c <= n - 1 - e)) return m;
  for (int i = 1; i <= 2 * e; ++i) {
    bool uc;
    if (gmap(grey_pre_cand(i, c, n, e), n, mode, &uc) == c && !uc) m |= (uint64_t)1 << i;
  }
  return m;
}

// gmap for 32-bit coordinates.
__device__ __forceinline__ int gmap32(int c, int size, int mode, bool* use_cval) {
  *use_cval = false;
  if ((unsigned)c < (unsigned)size) return c;
  bool uc;
  const int64_t m = gmap(c, size, mode, &uc);
  *use_cval = uc;
  return (int)m;
}

// E > 0: both element extents equal E, known at compile time (tile extents
// are constants and the element table lives in registers); E == 0: any
// extents up to 16, the table in LDS.
template <bool MAX, int E>
__global__ __launch_bounds__(256) void grey_bw2d_kernel(const float* __restrict__ in,
                                                        const float* __restrict__ dy,
                                                        float* __restrict__ dx, int H, int W,
                                                        GreyTabN tab, int mode, float cval) {
  extern __shared__ float smem[];
  constexpr bool CT = E > 0;
  constexpr int NK = CT ? (2 * E + 1) * (2 * E + 1) : 1;
  const int ey = CT ? E : tab.ext[0], ex = CT ? E : tab.ext[1];
  const int py = ey + 1, px = ex + 1;  // output halo
  const int OR = kBwTY + 2 * py, OC = kBwTX + 2 * px;
  const int XR = OR + 2 * ey, XC = OC + 2 * ex;
  float* xs = smem;                     // [XR][XC] mode-extended input
  float* ds = xs + XR * XC;             // [OR][OC] dy of in-item outputs
  int32_t* kx = (int32_t*)(ds + OR * OC);  // [n] element offset in xs (E == 0)
  int32_t* ko = kx + kMaxTable;            // [n] element offset in as / ds
  float* kv = (float*)(ko + kMaxTable);    // [n] element value
  int16_t* as = (int16_t*)(kv + kMaxTable);  // [OR][OC] winning element, -1 = not an output
  const int nk = tab.n;
  int rkx[NK], rko[NK];
  float rkv[NK];
  if (CT) {
#pragma unroll
    for (int k = 0; k < NK; ++k) {
      const int k2 = k < nk ? k : 0;
      rkx[k] = tab.o[k2][0] * XC + tab.o[k2][1];
      rko[k] = tab.o[k2][0] * OC + tab.o[k2][1];
      rkv[k] = tab.v[k2];
    }
  } else {
    for (int k = threadIdx.x; k < nk; k += blockDim.x) {
      kx[k] = tab.o[k][0] * XC + tab.o[k][1];
      ko[k] = tab.o[k][0] * OC + tab.o[k][1];
      kv[k] = tab.v[k];
    }
  }
  const int ntx = (W + kBwTX - 1) / kBwTX, nty = (H + kBwTY - 1) / kBwTY;
  const int lb = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = lb % (ntx * nty);
  const int64_t item = lb / (ntx * nty);
  const int x0 = (t % ntx) * kBwTX, y0 = (t / ntx) * kBwTY;
  const int64_t n = (int64_t)H * W;
  const float* vol = in + item * n;
  const float* dyv = dy + item * n;
  // input region starts at (y0 - py - ey, x0 - px - ex), output region at
  // (y0 - py, x0 - px)
  if (CT) {
    // every load of the thread is issued before the first LDS store: one
    // memory round trip for x and dy together
    constexpr int CXR = kBwTY + 4 * E + 2, CXC = kBwTX + 4 * E + 2;
    constexpr int COR = kBwTY + 2 * E + 2, COC = kBwTX + 2 * E + 2;
    constexpr int NX = (CXR * CXC + 255) / 256, NO = (COR * COC + 255) / 256;
    float rx[NX], rd[NO];
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      const int e = (int)threadIdx.x + j * 256;
      rx[j] = 0.f;
      if (e < CXR * CXC) {
        const int r = e / CXC, c = e - r * CXC;
        bool uy, ux;
        const int ym = gmap32(y0 - py - ey + r, H, mode, &uy);
        const int xm = gmap32(x0 - px - ex + c, W, mode, &ux);
        rx[j] = (uy || ux) ? cval : vol[(int64_t)ym * W + xm];
      }
    }
#pragma unroll
    for (int j = 0; j < NO; ++j) {
      const int e = (int)threadIdx.x + j * 256;
      rd[j] = 0.f;
      if (e < COR * COC) {
        const int r = e / COC, c = e - r * COC;
        const int y = y0 - py + r, x = x0 - px + c;
        if ((unsigned)y < (unsigned)H && (unsigned)x < (unsigned)W)
          rd[j] = dyv[(int64_t)y * W + x];
      }
    }
#pragma unroll
    for (int j = 0; j < NX; ++j) {
      const int e = (int)threadIdx.x + j * 256;
      if (e < CXR * CXC) xs[e] = rx[j];
    }
#pragma unroll
    for (int j = 0; j < NO; ++j) {
      const int e = (int)threadIdx.x + j * 256;
      if (e < COR * COC) ds[e] = rd[j];
    }
  } else {
    for (int e = threadIdx.x; e < XR * XC; e += blockDim.x) {
      const int r = e / XC, c = e - r * XC;
      bool uy, ux;
      const int ym = gmap32(y0 - py - ey + r, H, mode, &uy);
      const int xm = gmap32(x0 - px - ex + c, W, mode, &ux);
      xs[e] = (uy || ux) ? cval : vol[(int64_t)ym * W + xm];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < OR * OC; e += blockDim.x) {
    const int r = e / OC, c = e - r * OC;
    const int y = y0 - py + r, x = x0 - px + c;
    if ((unsigned)y >= (unsigned)H || (unsigned)x >= (unsigned)W) {
      as[e] = -1;
      continue;
    }
    const float* base = xs + (r + ey) * XC + (c + ex);
    float best = 0.f;
    int bi = 0;
    if (CT) {
#pragma unroll
      for (int k = 0; k < NK; ++k) {
        if (k < nk) {
          const float v = base[rkx[k]] + rkv[k];
          if (k == 0 || (MAX ? v > best : v < best)) {
            best = v;
            bi = k;
          }
        }
      }
    } else {
      for (int k = 0; k < nk; ++k) {
        const float v = base[kx[k]] + kv[k];
        if (k == 0 || (MAX ? v > best : v < best)) {
          best = v;
          bi = k;
        }
      }
    }
    as[e] = (int16_t)bi;
    if (!CT) ds[e] = dyv[(int64_t)y * W + x];
  }
  __syncthreads();
  for (int e = threadIdx.x; e < kBwTY * kBwTX; e += blockDim.x) {
    const int r = e >> 6, c = e & 63;
    const int qy = y0 + r, qx = x0 + c;
    if (qy >= H || qx >= W) continue;
    if (qy >= py && qy <= H - 1 - py && qx >= px && qx <= W - 1 - px) {
      // interior: q is its own only preimage
      const int s0 = (r + py) * OC + (c + px);
      float acc = 0.f;
      if (CT) {
#pragma unroll
        for (int k = 0; k < NK; ++k) {
          if (k < nk) {
            const int s = s0 - rko[k];
            if (as[s] == k) acc += ds[s];
          }
        }
      } else {
        for (int k = 0; k < nk; ++k) {
          const int s = s0 - ko[k];
          if (as[s] == k) acc += ds[s];
        }
      }
      dx[item * n + (int64_t)qy * W + qx] = acc;
      continue;
    }
    const uint64_t my = grey_preimages(qy, H, py, mode);
    const uint64_t mx = grey_preimages(qx, W, px, mode);
    float acc = 0.f;
    for (int k = 0; k < nk; ++k) {
      for (uint64_t wy = my; wy; wy &= wy - 1) {
        const int pr = grey_pre_cand(__builtin_ctzll(wy), qy, H, py) - tab.o[k][0] - (y0 - py);
        for (uint64_t wx = mx; wx; wx &= wx - 1) {
          const int pc =
              grey_pre_cand(__builtin_ctzll(wx), qx, W, px) - tab.o[k][1] - (x0 - px);
          if (pr < 0 || pr >= OR || pc < 0 || pc >= OC) continue;
          const int s = pr * OC + pc;
          if (as[s] == k) acc += ds[s];
        }
      }
    }
    dx[item * n + (int64_t)qy * W + qx] = acc;
  }
}

// Backward of the flat 3 x 3 element with zero origin on 2-D items (modes
// other than wrap), register streaming: grey_bw2d_kernel's routing and sums,
// term for term, with no LDS.  A wave owns input columns x0 .. x0 + 61
// (lanes 1 .. 62; lane l sits at column X = x0 - 1 + l) and kBwrRows input
// rows.  It loads the mode-extended x rows y0 - 2 .. y0 + kBwrRows + 1 and
// the dy rows y0 - 1 .. y0 + kBwrRows at once, forms the winning element of
// every output of rows y0 - 1 .. y0 + kBwrRows (neighbouring columns by DPP
// lane shifts; the columns past the wave's ends are loaded by lanes 0 and 1),
// and each input then sums, in element order, over the outputs that route to
// it.  Element k of the table (raster order) reads p + o_k: o_k = (1 - k / 3,
// 1 - k % 3) for a dilation (the offsets are negated), (k / 3 - 1, k % 3 - 1)
// for an erosion; its value is +0.0f (dilation) or -0.0f (erosion).  Every
// in-item output that routes to an input q lies within one voxel of q (a
// read outside the item at c maps back to c's reflection, one voxel from
// the border), so border inputs sum over the same register window, with
// grey_preimages' candidates in its order.
constexpr int kBwrRows = 8;


template <bool MAX>
__global__ __launch_bounds__(256) void grey_bwr_kernel(const float* __restrict__ in,
                                                       const float* __restrict__ dy,
                                                       float* __restrict__ dx, int H, int W,
                                                       int mode, float cval) {
  constexpr int NXR = kBwrRows + 4, NOR = kBwrRows + 2;
  const int ntx = (W + 61) / 62, nty = (H + kBwrRows - 1) / kBwrRows;
  const int nyb = (nty + 3) / 4;
  const int lb = xcd_swizzle(blockIdx.x, gridDim.x);
  const int t = lb % (ntx * nyb);
  const int64_t item = lb / (ntx * nyb);
  const int wave = (int)threadIdx.x >> 6, lane = (int)threadIdx.x & 63;
  const int ty = (t / ntx) * 4 + wave;
  if (ty >= nty) return;  // whole wave; no barriers in this kernel
  const int x0 = (t % ntx) * 62, y0 = ty * kBwrRows;
  const int X = x0 - 1 + lane;
  const int64_t n = (int64_t)H * W;
  const float* vol = in + item * n;
  const float* dyv = dy + item * n;
  float* dxv = dx + item * n;
  const float val = MAX ? 0.0f : -0.0f;

  // loads: x rows (this lane's column; lanes 0 / 1 also the columns left of
  // lane 0 / right of lane 63), dy rows; every address in range
  bool ux, uh;
  const int xm = gmap32(X, W, mode, &ux);
  const int xh = gmap32(lane == 1 ? x0 + 63 : x0 - 2, W, mode, &uh);
  const bool xin = (unsigned)X < (unsigned)W;
  const int xd = xin ? X : 0;
  float xv[NXR], xe[NXR], dv[NOR];
  bool ur[NXR];
#pragma unroll
  for (int i = 0; i < NXR; ++i) {
    const int ym = gmap32(y0 - 2 + i, H, mode, &ur[i]);
    xv[i] = vol[(int64_t)ym * W + xm];
    xe[i] = vol[(int64_t)ym * W + xh];
  }
#pragma unroll
  for (int i = 0; i < NOR; ++i) {
    const int py = y0 - 1 + i;
    dv[i] = dyv[(int64_t)((unsigned)py < (unsigned)H ? py : 0) * W + xd];
  }
  const bool hl = __builtin_amdgcn_readlane((int)uh, 0) != 0;
  const bool hr = __builtin_amdgcn_readlane((int)uh, 1) != 0;

  // x rows with the border substitution, and their values one column left
  // and right
  float xc[NXR], xl[NXR], xr[NXR];
#pragma unroll
  for (int i = 0; i < NXR; ++i) {
    xc[i] = (ur[i] || ux) ? cval : xv[i];
    const float el = (ur[i] || hl) ? cval
                                   : __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xe[i]), 0));
    const float er = (ur[i] || hr) ? cval
                                   : __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xe[i]), 1));
    xl[i] = lane_prev(xc[i], el);
    xr[i] = rw_next_lane(xc[i], er);
  }
  // winning element and dy of outputs y0 - 1 + i (-1 / 0 outside the item),
  // and of the outputs one column left and right
  int wc[NOR], wl[NOR], wr[NOR];
  float dc[NOR], dl[NOR], dr[NOR];
#pragma unroll
  for (int i = 0; i < NOR; ++i) {
    const int py = y0 - 1 + i;
    const bool out = xin && (unsigned)py < (unsigned)H;
    float best = 0.f;
    int bi = 0;
#pragma unroll
    for (int k = 0; k < 9; ++k) {
      const int oy = MAX ? 1 - k / 3 : k / 3 - 1, ox = MAX ? 1 - k % 3 : k % 3 - 1;
      const int r = i + 1 + oy;  // x row of p + o_k
      const float v = (ox < 0 ? xl[r] : ox > 0 ? xr[r] : xc[r]) + val;
      if (k == 0 || (MAX ? v > best : v < best)) {
        best = v;
        bi = k;
      }
    }
    wc[i] = out ? bi : -1;
    dc[i] = out ? dv[i] : 0.f;
    wl[i] = lane_prev_i(wc[i], -1);
    wr[i] = lane_next_i(wc[i], -1);
    dl[i] = lane_prev(dc[i], 0.f);
    dr[i] = rw_next_lane(dc[i], 0.f);
  }

  // inputs
  const bool lane_st = lane >= 1 && lane <= 62 && xin;
  const uint64_t mxm = grey_preimages(X < 0 ? 0 : X, W, 2, mode);
  const bool border_lanes = __builtin_amdgcn_ballot_w64(mxm != 1) != 0;  // uniform
#pragma unroll
  for (int j = 0; j < kBwrRows; ++j) {
    const int qy = y0 + j;
    const uint64_t mym = grey_preimages(qy, H, 2, mode);  // uniform
    float acc = 0.f;
    if (mym == 1 && !border_lanes) {
      // q is its own only preimage: p = q - o_k
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const int oy = MAX ? 1 - k / 3 : k / 3 - 1, ox = MAX ? 1 - k % 3 : k % 3 - 1;
        const int r = j + 1 - oy;  // output row of p
        const int w = ox > 0 ? wl[r] : ox < 0 ? wr[r] : wc[r];
        const float d = ox > 0 ? dl[r] : ox < 0 ? dr[r] : dc[r];
        if (w == k) acc += d;
      }
    } else {
      // candidates c of q per axis in grey_preimages' bit order (q, -2, -1,
      // n, n + 1); p = c - o_k routes to q when it lies in the item with
      // winner k.  Of the candidates outside the item only -1 (with o = -1:
      // p = 0) and n (with o = +1: p = n - 1) reach the item, so each
      // element has at most (1 + |oy|)(1 + |ox|) terms, in that order; p then
      // lies within one voxel of q (rows j .. j + 2 of the window).  The
      // window's values pass an empty asm so the selects below stay selects
      // of registers (not of addresses into a stack copy of the arrays).
      int w0 = wl[j], w1 = wc[j], w2 = wr[j], w3 = wl[j + 1], w4 = wc[j + 1], w5 = wr[j + 1],
          w6 = wl[j + 2], w7 = wc[j + 2], w8 = wr[j + 2];
      float e0 = dl[j], e1 = dc[j], e2 = dr[j], e3 = dl[j + 1], e4 = dc[j + 1], e5 = dr[j + 1],
            e6 = dl[j + 2], e7 = dc[j + 2], e8 = dr[j + 2];
      asm volatile("" : "+v"(w0), "+v"(w1), "+v"(w2), "+v"(w3), "+v"(w4), "+v"(w5), "+v"(w6),
                   "+v"(w7), "+v"(w8));
      asm volatile("" : "+v"(e0), "+v"(e1), "+v"(e2), "+v"(e3), "+v"(e4), "+v"(e5), "+v"(e6),
                   "+v"(e7), "+v"(e8));
      const bool yl = (mym >> 2) & 1u, yh = (mym >> 3) & 1u;  // -1 / H map to qy (uniform)
      const bool xlo = (mxm >> 2) & 1u, xhi = (mxm >> 3) & 1u;  // -1 / W map to X
#pragma unroll
      for (int k = 0; k < 9; ++k) {
        const int oy = MAX ? 1 - k / 3 : k / 3 - 1, ox = MAX ? 1 - k % 3 : k % 3 - 1;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          if ((a == 1 && oy != -1) || (a == 2 && oy != 1)) continue;
          const bool vy = a == 0 ? true : a == 1 ? yl : yh;
          const int ry = a == 0 ? -oy : a == 1 ? -qy : H - 1 - qy;
#pragma unroll
          for (int b = 0; b < 3; ++b) {
            if ((b == 1 && ox != -1) || (b == 2 && ox != 1)) continue;
            const bool vx = b == 0 ? true : b == 1 ? xlo : xhi;
            const int rx = b == 0 ? -ox : b == 1 ? -X : W - 1 - X;
            const int w = ry < 0 ? (rx < 0 ? w0 : rx > 0 ? w2 : w1)
                                 : ry > 0 ? (rx < 0 ? w6 : rx > 0 ? w8 : w7)
                                          : (rx < 0 ? w3 : rx > 0 ? w5 : w4);
            const float d = ry < 0 ? (rx < 0 ? e0 : rx > 0 ? e2 : e1)
                                   : ry > 0 ? (rx < 0 ? e6 : rx > 0 ? e8 : e7)
                                            : (rx < 0 ? e3 : rx > 0 ? e5 : e4);
            if (vy && vx && w == k) acc += d;
          }
        }
      }
    }
    if (lane_st && qy < H) dxv[(int64_t)qy * W + X] = acc;
  }
}

static size_t grey_bw2d_lds(const GreyTabN& tab) {
  const int py = tab.ext[0] + 1, px = tab.ext[1] + 1;
  const size_t OR = kBwTY + 2 * py, OC = kBwTX + 2 * px;
  const size_t XR = OR + 2 * tab.ext[0], XC = OC + 2 * tab.ext[1];
  return (XR * XC + OR * OC + 3 * (size_t)kMaxTable) * sizeof(float) + OR * OC * sizeof(int16_t);
}

}  // namespace tmh

extern "C" size_t tm_grey_backward_workspace_bytes(int64_t batch, int ndim, const int64_t* shape) {
  int64_t n = batch;
  for (int j = 0; j < ndim; ++j) n *= shape[j];
  return (size_t)n * sizeof(int16_t) + 256 + 64;
}

extern "C" int tm_grey_backward(const float* input, int64_t batch, int ndim, const int64_t* shape,
                                const float* structure, const uint8_t* footprint,
                                const int64_t* se_shape, const int64_t* origin, int mode,
                                float cval, int dilation, const float* grad_out, float* grad_in,
                                void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(input && grad_out && grad_in, "tm_grey_backward: NULL argument");
  TM_CHECK_ARG(ndim >= 1 && ndim <= kMaxNdim, "tm_grey_backward: ndim must be in 1..8");
  TM_CHECK_ARG(mode >= 0 && mode <= 4, "tm_grey_backward: mode must be in 0..4");
  hipStream_t st = (hipStream_t)stream;
  Shape sh = make_shape(ndim, shape);
  const int64_t total = batch * sh.numel;
  TM_CHECK_ARG(sh.numel < ((int64_t)1 << 31), "tm_grey_backward: items above 2^31 voxels");
  int64_t ne = 1;
  for (int j = 0; j < ndim; ++j) ne *= se_shape[j];
  GreyTabN tab{};
  int32_t ext[kMaxNdim] = {0};
  int n = 0;
  for (int64_t e = 0; e < ne; ++e) {
    if (footprint && !footprint[e]) continue;
    TM_CHECK_ARG(n < kMaxTable, "tm_grey_backward: more than %d active elements", kMaxTable);
    int64_t rem = e;
    int64_t r[kMaxNdim];
    for (int j = ndim - 1; j >= 0; --j) {
      r[j] = rem % se_shape[j];
      rem /= se_shape[j];
    }
    for (int j = 0; j < ndim; ++j) {
      int64_t d = r[j] - (se_shape[j] / 2 + origin[j]);
      if (dilation) d = -d;
      TM_CHECK_ARG(d >= -127 && d <= 127, "tm_grey_backward: element offsets beyond +-127");
      tab.o[n][j] = (int8_t)d;
      const int32_t ad = (int32_t)(d < 0 ? -d : d);
      ext[j] = ad > ext[j] ? ad : ext[j];
    }
    const float s = structure ? structure[e] : 0.0f;
    tab.v[n] = dilation ? s : -s;
    ++n;
  }
  TM_CHECK_ARG(n > 0, "tm_grey_backward: footprint must contain at least one active position");
  tab.n = n;
  for (int j = 0; j < kMaxNdim; ++j) tab.ext[j] = j < ndim ? ext[j] : 0;
  if (ndim == 2 && mode != 4 && ext[0] <= 16 && ext[1] <= 16) {
    const int H = (int)shape[0], W = (int)shape[1];
    const int64_t G = (int64_t)((W + kBwTX - 1) / kBwTX) * ((H + kBwTY - 1) / kBwTY) * batch;
    // the flat 3 x 3 element with zero origin: every element in raster
    // order, offsets +-(r - 1), values +0.0f / -0.0f (grey_bwr_kernel's table)
    bool flat3 = n == 9 && se_shape[0] == 3 && se_shape[1] == 3;
    for (int k = 0; k < n && flat3; ++k) {
      const int oy = dilation ? 1 - k / 3 : k / 3 - 1, ox = dilation ? 1 - k % 3 : k % 3 - 1;
      uint32_t vb, want = dilation ? 0u : 0x80000000u;
      memcpy(&vb, &tab.v[k], 4);
      flat3 = tab.o[k][0] == oy && tab.o[k][1] == ox && vb == want;
    }
    const int64_t Gr = (int64_t)((W + 61) / 62) * (((H + kBwrRows - 1) / kBwrRows + 3) / 4) * batch;
    if (G < ((int64_t)1 << 31) && Gr < ((int64_t)1 << 31)) {
      const size_t lds = grey_bw2d_lds(tab);
      const bool e1 = ext[0] == 1 && ext[1] == 1;
      // launch variant 0: the LDS-tiled kernel; 1: register streaming (flat
      // 3 x 3 only)
      auto run = [&](int v) -> int {
        if (v == 1) {
          if (!flat3) return TM_ERR_ARG;
          if (dilation)
            hipLaunchKernelGGL(grey_bwr_kernel<true>, dim3((unsigned)Gr), dim3(256), 0, st,
                               input, grad_out, grad_in, H, W, mode, cval);
          else
            hipLaunchKernelGGL(grey_bwr_kernel<false>, dim3((unsigned)Gr), dim3(256), 0, st,
                               input, grad_out, grad_in, H, W, mode, cval);
          TM_LAUNCH_CHECK();
          return TM_OK;
        }
#define TM_BW(MX, EE)                                                                           \
  hipLaunchKernelGGL((grey_bw2d_kernel<MX, EE>), dim3((unsigned)G), dim3(256), lds, st, input,     \
                     grad_out, grad_in, H, W, tab, mode, cval)
        if (dilation) {
          if (e1) TM_BW(true, 1);
          else TM_BW(true, 0);
        } else {
          if (e1) TM_BW(false, 1);
          else TM_BW(false, 0);
        }
#undef TM_BW
        TM_LAUNCH_CHECK();
        return TM_OK;
      };
      if (!flat3) return run(0);
      auto p2 = [](int64_t v) { int64_t q = 1; while (q < v) q <<= 1; return q; };
      const int64_t key[4] = {dilation ? 1 : 0, p2(H), p2(W), p2(batch)};
      bool ran = false;
      const int v = tune_pick("greybw", key, 4, 2, 1, run, st, &ran);
      return ran ? TM_OK : run(v);
    }
  }
  // the general path (other ranks, wrap, larger elements) keeps the winning
  // element of every output in the workspace; the 2-D kernels above need none
  const size_t need = tm_grey_backward_workspace_bytes(batch, ndim, shape);
  if (!workspace || workspace_bytes < need) {
    set_error("tm_grey_backward: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, need);
    return TM_ERR_WORKSPACE;
  }
  int16_t* arg = (int16_t*)workspace;
  for (int t = 0; t < tab.n; ++t) {
    int64_t f = 0;
    for (int j = 0; j < ndim; ++j) f += (int64_t)tab.o[t][j] * sh.stride[j];
    tab.flat[t] = (int32_t)f;  // |f| < numel of one item
  }
  const dim3 grid(grid_for(total, 256)), blk(256);
  if (dilation)
    hipLaunchKernelGGL(grey_arg_kernel<true>, grid, blk, 0, st, input, arg, sh, total, tab, mode,
                       cval);
  else
    hipLaunchKernelGGL(grey_arg_kernel<false>, grid, blk, 0, st, input, arg, sh, total, tab, mode,
                       cval);
  TM_LAUNCH_CHECK();
  hipLaunchKernelGGL(grey_grad_gather_kernel, grid, blk, 0, st, arg, grad_out, grad_in, sh, total,
                     tab, mode);
  TM_LAUNCH_CHECK();
  return TM_OK;
}
