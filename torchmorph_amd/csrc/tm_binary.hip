// Binary erosion / dilation (with iterations, mask, border_value) for gfx950.
//
// Semantics follow the reference (binary_kernel.cu:56-127 for one step,
// offsets binary_kernel.cu:199-222; iteration and mask handling
// binary.py:56-88): one step ANDs (erosion) the input at p + d_i or ORs
// (dilation) it at p - d_i over the active structure elements, reading
// border_value outside the item; with a mask, each step's result is kept only
// where mask is set and the step's input is restored elsewhere; iterations < 1
// repeats until an iteration changes nothing.
//
// Layout: every row along the contiguous axis is bit-packed into 64-bit words
// (bit b of word w = voxel 64w + b).  A step computes an output word from the
// shifted neighbour words of the element offsets.  Without a mask, up to 8
// steps run per launch on tiles of rows held in LDS (bin_tile_kernel); the
// first such launch can pack the bool input and the last unpack the bool
// output, so the pipeline touches HBM about once for the input and once for
// the output.  With a mask, or until stable, one step runs per launch
// (bin_step_kernel) on the 1/8-size packed volume, which stays in L2 /
// Infinity Cache.
#include <vector>

#include "tm_common.h"

namespace tmh {

constexpr int kMaxBinTable = 128;
struct BinTable {  // 4-byte multiple: copied to LDS as words
  int n;
  int8_t d[kMaxBinTable][kMaxNdim];  // per element: offsets for every axis
};
static_assert(sizeof(BinTable) % 4 == 0, "BinTable is copied as 32-bit words");

struct Packed {
  int nd;
  int64_t size[kMaxNdim];   // spatial shape
  int64_t rstride[kMaxNdim];  // row stride of the leading axes (in rows)
  int64_t rows;             // rows per item
  int64_t ww;               // words per row
  int64_t W;                // row length in voxels
};

// One packed word per lane.  The 64 loads are unconditional (past the row's
// end they read its last element) and the bits are formed after all of them
// were issued.
template <typename TIn>
__global__ __launch_bounds__(256) void bin_pack_kernel(const TIn* __restrict__ in,
                                                       uint64_t* __restrict__ out, Packed pk,
                                                       int64_t nwords) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nwords) return;
  const int64_t row = g / pk.ww;
  const int64_t w = g - row * pk.ww;
  const TIn* src = in + row * pk.W + w * 64;
  const int nb = (int)min((int64_t)64, pk.W - w * 64);
  TIn v[64];
#pragma unroll
  for (int b = 0; b < 64; ++b) v[b] = src[b < nb ? b : nb - 1];
  uint64_t bits = 0;
#pragma unroll
  for (int b = 0; b < 64; ++b)
    if (b < nb && v[b] != TIn(0)) bits |= 1ull << b;
  out[g] = bits;
}

// Nonzero bytes of a 32-bit word as 4 bits (byte i -> bit i).
__device__ __forceinline__ uint32_t nz4(uint32_t v) {
  const uint32_t hi = (((v & 0x7f7f7f7fu) + 0x7f7f7f7fu) | v) & 0x80808080u;
  return (((hi >> 7) * 0x01020408u) >> 24) & 0xfu;
}

// 64 bytes (16-byte aligned) as 64 bits, nonzero -> 1; four loads issued
// together.
__device__ __forceinline__ uint64_t bin_pack64(const uint8_t* p) {
  const uint4* q4 = (const uint4*)p;
  uint4 q[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) q[k] = q4[k];
  uint64_t bits = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t b16 = nz4(q[k].x) | (nz4(q[k].y) << 4) | (nz4(q[k].z) << 8) |
                         (nz4(q[k].w) << 12);
    bits |= (uint64_t)b16 << (16 * k);
  }
  return bits;
}

// Byte masks whose rows are whole 64-byte words (W % 64 == 0, 16-byte
// aligned input): four 16-byte loads per word.
__global__ __launch_bounds__(256) void bin_pack_bytes_kernel(const uint8_t* __restrict__ in,
                                                             uint64_t* __restrict__ out,
                                                             int64_t nwords) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nwords) return;
  out[g] = bin_pack64(in + g * 64);
}

__global__ __launch_bounds__(256) void bin_unpack_kernel(const uint64_t* __restrict__ in,
                                                         uint8_t* __restrict__ out, Packed pk,
                                                         int64_t total) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total) return;
  const int64_t row = g / pk.W;
  const int64_t x = g - row * pk.W;
  out[g] = (uint8_t)((in[row * pk.ww + (x >> 6)] >> (x & 63)) & 1ull);
}

// 4 bits as 4 bytes of 0 / 1 (bit i -> byte i).
__device__ __forceinline__ uint32_t bytes4(uint32_t b) {
  return (b * 0x00204081u) & 0x01010101u;
}

// 64 bits as 64 bytes of 0 / 1 (16-byte aligned), four 16-byte stores.
__device__ __forceinline__ void bin_unpack64(uint64_t bits, uint8_t* p) {
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const uint32_t b = (uint32_t)(bits >> (16 * k)) & 0xffffu;
    uint4 v;
    v.x = bytes4(b & 0xfu);
    v.y = bytes4((b >> 4) & 0xfu);
    v.z = bytes4((b >> 8) & 0xfu);
    v.w = bytes4((b >> 12) & 0xfu);
    ((uint4*)p)[k] = v;
  }
}

// Rows of whole 16-byte groups (W % 16 == 0, 16-byte aligned output): 16
// output bytes per lane, one 16-byte store.
__global__ __launch_bounds__(256) void bin_unpack_bytes_kernel(const uint64_t* __restrict__ in,
                                                               uint8_t* __restrict__ out,
                                                               Packed pk, int64_t ngroups) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  const int64_t gpr = pk.W / 16;  // groups per row
  const int64_t row = g / gpr;
  const int64_t x = (g - row * gpr) * 16;
  const uint32_t b = (uint32_t)((in[row * pk.ww + (x >> 6)] >> (x & 63)) & 0xffffull);
  uint4 v;
  v.x = bytes4(b & 0xfu);
  v.y = bytes4((b >> 4) & 0xfu);
  v.z = bytes4((b >> 8) & 0xfu);
  v.w = bytes4((b >> 12) & 0xfu);
  *(uint4*)(out + g * 16) = v;
}

// Loaded word v of word index q of a row, with everything outside [0, W)
// reading `border` (q outside the row, bits past W in the last word).
__device__ __forceinline__ uint64_t bin_fix(uint64_t v, const Packed& pk, int64_t q,
                                            uint64_t border) {
  if (q < 0 || q >= pk.ww) return border;
  const int64_t valid = pk.W - q * 64;
  if (valid < 64) {
    const uint64_t m = (1ull << valid) - 1ull;
    v = (v & m) | (border & ~m);
  }
  return v;
}

// Element tables of structures beyond BinTable (more than kMaxBinTable
// active elements, or offsets beyond +-127): a step then runs as one launch
// per chunk of kWideChunk elements, each chunk ANDing (erosion) or ORing
// (dilation) into the previous chunk's partial result, with int16 offsets.
constexpr int kWideChunk = 64;
struct BinTableWide {  // 4-byte multiple: copied to LDS as words
  int n;
  int16_t d[kWideChunk][kMaxNdim];
};
static_assert(sizeof(BinTableWide) % 4 == 0, "BinTableWide is copied as 32-bit words");

// Offset of element t along axis j, read from the LDS copy as 32-bit words
// (8 int8 offsets in two words, or 8 int16 offsets in four).
__device__ __forceinline__ int tab_off(const BinTable& tb, int t, int j) {
  const uint32_t* dw = (const uint32_t*)&tb.d[t][0];
  const uint32_t v = dw[j >> 2] >> (8 * (j & 3));
  return (int)(int8_t)(v & 0xffu);
}
__device__ __forceinline__ int tab_off(const BinTableWide& tb, int t, int j) {
  const uint32_t* dw = (const uint32_t*)&tb.d[t][0];
  const uint32_t v = dw[j >> 1] >> (16 * (j & 1));
  return (int)(int16_t)(v & 0xffffu);
}

// One step (or, for a wide table, one chunk of a step: `partial` then holds
// the previous chunks' result, combined in place, and `last` is set on the
// final chunk).  `invert` reads the complement of `in` (border stays
// border_value): the miss half of a hit-or-miss transform.  `orig`/`mask` (optional, applied on the last chunk)
// implement the reference's torch.where(mask, step(x), x_input).  `changed`
// (optional, last chunk) is set when any output word differs from its input
// word (valid bits only).
template <typename Tab>
__global__ __launch_bounds__(256) void bin_step_kernel(const uint64_t* __restrict__ in,
                                                       uint64_t* __restrict__ out, Packed pk,
                                                       int64_t nwords, Tab tab, int dilation,
                                                       int border_value,
                                                       const uint64_t* __restrict__ orig,
                                                       const uint64_t* __restrict__ mask,
                                                       int* __restrict__ changed, int partial,
                                                       int last, int invert) {
  // the element table is copied to LDS once per workgroup: its offsets at a
  // loop index would otherwise be byte loads from the kernel arguments, each
  // waited for before the next
  __shared__ Tab stab;
  for (int i = threadIdx.x; i < (int)(sizeof(Tab) / 4); i += blockDim.x)
    ((uint32_t*)&stab)[i] = ((const uint32_t*)&tab)[i];
  __syncthreads();
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nwords) return;
  const int64_t per_item = pk.rows * pk.ww;
  const int lead = pk.nd - 1;
  const int64_t item = g / per_item;
  const int64_t rem = g - item * per_item;
  const int64_t row = rem / pk.ww;
  const int64_t w = rem - row * pk.ww;
  // leading coordinates of this row
  int64_t c[kMaxNdim];
  {
    int64_t r = row;
    for (int j = lead - 1; j >= 0; --j) {
      c[j] = r % pk.size[j];
      r /= pk.size[j];
    }
  }
  const uint64_t* src = in + item * per_item;
  const uint64_t border = border_value ? ~0ull : 0ull;
  // elements in groups of 8: every address of a group is formed first
  // (words outside the item read word 0 and are replaced by the border after
  // the load), the 16 loads are issued together, then the terms combine
  uint64_t acc = partial ? out[g] : (dilation ? 0ull : ~0ull);
  for (int t0 = 0; t0 < stab.n; t0 += 8) {
    int64_t a_lo[8], q_lo[8];
    int sh[8];
    bool inb[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int t = t0 + u < stab.n ? t0 + u : stab.n - 1;
      auto off = [&](int j) { return tab_off(stab, t, j); };
      bool ok = true;
      int64_t r2 = row;
#pragma unroll
      for (int j = 0; j < kMaxNdim - 1; ++j) {
        if (j < lead) {
          const int64_t cj = c[j] + off(j);
          ok = ok && cj >= 0 && cj < pk.size[j];
          r2 += (int64_t)off(j) * pk.rstride[j];
        }
      }
      const int64_t tb = w * 64 + off(lead);
      const int64_t q = tb >= 0 ? tb / 64 : -((-tb + 63) / 64);
      sh[u] = (int)(tb - q * 64);
      inb[u] = ok;
      q_lo[u] = q;
      a_lo[u] = ok ? r2 * pk.ww + q : 0;
    }
    uint64_t lo[8], hi[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const bool l_in = inb[u] && q_lo[u] >= 0 && q_lo[u] < pk.ww;
      const bool h_in = inb[u] && q_lo[u] + 1 >= 0 && q_lo[u] + 1 < pk.ww;
      lo[u] = src[l_in ? a_lo[u] : 0];
      hi[u] = src[h_in ? a_lo[u] + 1 : 0];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (t0 + u >= stab.n) break;
      uint64_t term = border;
      if (inb[u]) {
        const uint64_t inv = invert ? ~0ull : 0ull;
        const uint64_t l = bin_fix(lo[u] ^ inv, pk, q_lo[u], border);
        if (sh[u] == 0) {
          term = l;
        } else {
          const uint64_t h = bin_fix(hi[u] ^ inv, pk, q_lo[u] + 1, border);
          term = (l >> sh[u]) | (h << (64 - sh[u]));
        }
      }
      if (dilation) acc |= term;
      else acc &= term;
    }
  }
  if (last && mask) {
    const uint64_t m = mask[g];
    acc = (acc & m) | (orig[g] & ~m);
  }
  out[g] = acc;
  if (last && changed) {
    const int64_t valid = pk.W - w * 64;
    const uint64_t vm = valid >= 64 ? ~0ull : ((1ull << valid) - 1ull);
    // one atomic per wave with a changed word, not one per lane
    const uint64_t b = __ballot(((acc ^ in[g]) & vm) != 0);
    if (b != 0 && (int)(threadIdx.x & 63) == __ffsll((long long)b) - 1) atomicOr(changed, 1);
  }
}

// Several steps of one direction fused in LDS, for items of up to three axes
// (z, y, x; missing leading axes have size 1) without a mask.  A workgroup
// owns a tile of ty x tz rows of an item and loads it with a halo of hy / hz
// rows (hy = steps * max |d_y|, likewise z) into LDS; every step recomputes
// the rows that the remaining steps still need (the region shrinks by the
// structure's reach per step) and the last step leaves the tile exact.  Rows
// outside the item hold `border` in every step, as the step kernel reads
// them.  A row sits in LDS as ww words between two pad words of `border`, so
// the x offsets (|d_x| <= 63) read words w - 1 .. w + 1 without checks; bits
// past W in the last word are reset to `border` after every step.
struct BinTileGeom {
  int Z, Y, ww, W;      // item rows along z, y; words per row; voxels per row
  float inv_pitch, inv_by, inv_ww, inv_ty;  // 1 / divisor for small_div
  int n;                 // structure elements
  // element e reads LDS word (row + eo[e] + w) and the one after it, shifted
  // right by es[e]: eo holds the element's row offset and the word of bit
  // 64 w + d_x relative to w, plus the pad word; entries past n repeat
  // element 0 (plain words, read with scalar loads)
  int eo[kMaxBinTable], es[kMaxBinTable];
  int ty, tz, hy, hz;   // tile rows and halo rows
  int by, bz, pitch;    // LDS rows along y, z (tile + 2 halo); words per LDS row
  int tiles_y, tiles_z;
  int ry, rz;           // reach of the structure along y, z
};

// a / d for 0 <= a < 65536 and 1 <= d <= 1024, with inv = 1.0f / d rounded
// to float: (a + 0.5) * inv lies within 2^-24 relative of (a + 0.5) / d,
// closer than 0.5 / d to its integer part, so truncation gives the quotient
// (checked exhaustively over that range).
__device__ __forceinline__ int small_div(int a, float inv) {
  return (int)(((float)a + 0.5f) * inv);
}

// Bits sh .. sh + 63 of the 128-bit pair (hi:lo), 0 <= sh <= 63, without a
// branch: hi << (64 - sh) taken in two shifts is 0 for sh = 0.
__device__ __forceinline__ uint64_t bin_shift(uint64_t lo, uint64_t hi, int sh) {
  return (lo >> sh) | ((hi << 1) << (63 - sh));
}

// NE > 0: at most NE elements, their LDS offsets in registers and the
// element loop unrolled (every load of a word issued together); NE == 0: any
// count, offsets in LDS.
// IO bit 0: read the bool bytes `in_b` (a row of whole 64-byte words) instead
// of packed words; bit 1: write bool bytes to `out_b` instead of packed words.
// Either replaces a pack / unpack pass over the volume.
template <int NE, int IO>
__global__ __launch_bounds__(256) void bin_tile_kernel(const uint64_t* __restrict__ in,
                                                       uint64_t* __restrict__ out,
                                                       const uint8_t* __restrict__ in_b,
                                                       uint8_t* __restrict__ out_b,
                                                       BinTileGeom g, int dilation,
                                                       int border_value, int steps) {
  constexpr bool IN_B = (IO & 1) != 0, OUT_B = (IO & 2) != 0;
  constexpr int RU = IN_B ? 5 : 8;  // words loaded per lane per round (a 16 x 16-row tile
                                    // of 2-word rows with 4-row halos: one round)
  extern __shared__ uint64_t bin_lds[];
  __shared__ int eo_s[NE > 0 ? 1 : kMaxBinTable];
  __shared__ int es_s[NE > 0 ? 1 : kMaxBinTable];
  const int per = g.by * g.bz * g.pitch;
  uint64_t* A = bin_lds;
  uint64_t* B = bin_lds + per;
  const uint64_t border = border_value ? ~0ull : 0ull;
  const int tile = blockIdx.x;
  const int y0 = (tile % g.tiles_y) * g.ty - g.hy;
  const int z0 = (tile / g.tiles_y) * g.tz - g.hz;
  const int64_t item_rows = (int64_t)g.Z * g.Y;
  const uint64_t* src = in + (int64_t)blockIdx.y * item_rows * g.ww;
  uint64_t* dst = out + (int64_t)blockIdx.y * item_rows * g.ww;
  const int last_valid = g.W - (g.ww - 1) * 64;  // voxels in the last word, 1..64
  const uint64_t last_m = last_valid >= 64 ? ~0ull : ((1ull << last_valid) - 1ull);
  const int n = g.n;
  int eo_r[NE > 0 ? NE : 1], es_r[NE > 0 ? NE : 1];
  if (NE > 0) {
#pragma unroll
    for (int k = 0; k < (NE > 0 ? NE : 1); ++k) {
      eo_r[k] = g.eo[k];
      es_r[k] = g.es[k];
    }
  } else {
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
      eo_s[k] = g.eo[k];
      es_s[k] = g.es[k];
    }
  }
  const int64_t item_w = (int64_t)blockIdx.y * item_rows * g.ww;  // first word of the item
  // the tile's words: RU per lane per round over the rows' real words only
  // (not the pad entries), their loads issued together from addresses
  // clamped into the item, the border and the last word's bits applied
  // afterwards
  const int nw = g.by * g.bz * g.ww;
  for (int i0 = 0; i0 < nw; i0 += RU * 256) {
    uint64_t v[RU];
    int lw[RU], wv[RU];
    bool okv[RU];
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int i = i0 + u * 256 + (int)threadIdx.x;
      const int ic = i < nw ? i : nw - 1;
      const int r = small_div(ic, g.inv_ww), w = ic - r * g.ww;
      const int rz = small_div(r, g.inv_by);
      const int y = y0 + (r - rz * g.by), z = z0 + rz;
      okv[u] = y >= 0 && y < g.Y && z >= 0 && z < g.Z;
      wv[u] = w;
      lw[u] = r * g.pitch + w + 1;
      const int yc = y < 0 ? 0 : (y >= g.Y ? g.Y - 1 : y);
      const int zc = z < 0 ? 0 : (z >= g.Z ? g.Z - 1 : z);
      const int64_t wi = ((int64_t)zc * g.Y + yc) * g.ww + w;
      if (IN_B) v[u] = bin_pack64(in_b + (item_w + wi) * 64);
      else v[u] = src[wi];
    }
#pragma unroll
    for (int u = 0; u < RU; ++u) {
      const int i = i0 + u * 256 + (int)threadIdx.x;
      if (i >= nw) break;
      uint64_t x = okv[u] ? v[u] : border;
      if (okv[u] && wv[u] == g.ww - 1) x = (x & last_m) | (border & ~last_m);
      A[lw[u]] = x;
    }
  }
  // the pads of both buffers (word 0 and words ww + 1 .. pitch - 1 of every
  // row) hold the border
  const int npad = g.pitch - g.ww;
  for (int i = threadIdx.x; i < g.by * g.bz * npad; i += blockDim.x) {
    const int r = i / npad, k = i - r * npad;
    const int at = r * g.pitch + (k == 0 ? 0 : g.ww + k);
    A[at] = border;
    B[at] = border;
  }
  __syncthreads();
  for (int s = 1; s <= steps; ++s) {
    // rows still needed after this step: [s * r, b - s * r) along each axis;
    // one row per lane, all its words
    const int ly = s * g.ry, lz = s * g.rz;
    const int ny = g.by - 2 * ly, nz = g.bz - 2 * lz;
    const float inv_ny = 1.0f / (float)ny;
    for (int r = threadIdx.x; r < ny * nz; r += blockDim.x) {
      const int rz = small_div(r, inv_ny);
      const int byi = ly + (r - rz * ny), bzi = lz + rz;
      const int y = y0 + byi, z = z0 + bzi;
      const int row = (bzi * g.by + byi) * g.pitch;
      const bool inside = y >= 0 && y < g.Y && z >= 0 && z < g.Z;
      for (int w = 0; w < g.ww; ++w) {
        uint64_t acc = border;
        if (inside) {
          acc = dilation ? 0ull : ~0ull;
          if (NE > 0) {
            uint64_t lo[NE > 0 ? NE : 1], hi[NE > 0 ? NE : 1];
#pragma unroll
            for (int k = 0; k < (NE > 0 ? NE : 1); ++k) {
              const uint64_t* p = A + row + w + eo_r[k];
              lo[k] = p[0];
              hi[k] = p[1];
            }
            // entries past n repeat element 0: AND / OR absorb them
#pragma unroll
            for (int k = 0; k < (NE > 0 ? NE : 1); ++k) {
              const uint64_t term = bin_shift(lo[k], hi[k], es_r[k]);
              if (dilation) acc |= term;
              else acc &= term;
            }
          } else {
            for (int t = 0; t < n; ++t) {
              const uint64_t* p = A + row + w + eo_s[t];
              const uint64_t term = bin_shift(p[0], p[1], es_s[t]);
              if (dilation) acc |= term;
              else acc &= term;
            }
          }
          if (w == g.ww - 1) acc = (acc & last_m) | (border & ~last_m);
        }
        B[row + w + 1] = acc;
      }
    }
    __syncthreads();
    uint64_t* t = A;
    A = B;
    B = t;
  }
  // the tile's rows inside the item
  const int tw = g.ty * g.tz * g.ww;
  for (int i = threadIdx.x; i < tw; i += blockDim.x) {
    const int r = small_div(i, g.inv_ww), w = i - r * g.ww;
    const int rz = small_div(r, g.inv_ty);
    const int byi = g.hy + (r - rz * g.ty), bzi = g.hz + rz;
    const int y = y0 + byi, z = z0 + bzi;
    if (y < g.Y && z < g.Z) {
      const uint64_t v = A[(bzi * g.by + byi) * g.pitch + w + 1];
      const int64_t wi = ((int64_t)z * g.Y + y) * g.ww + w;
      if (OUT_B) bin_unpack64(v, out_b + (item_w + wi) * 64);
      else dst[wi] = v;
    }
  }
}

// LDS bytes of the fused steps (two buffers).
inline size_t bin_tile_lds(const BinTileGeom& g) {
  return (size_t)2 * g.by * g.bz * g.pitch * 8;
}

constexpr size_t kBinTileLdsMax = 60 * 1024;

// Geometry of the fused steps for `steps` steps of `tab`, or false when the
// item, the structure or the halo do not fit (the step kernel runs then).
inline bool bin_tile_plan(const Packed& pk, const BinTable& tab, int steps, BinTileGeom* g) {
  if (pk.nd > 3 || steps < 1) return false;
  int rx = 0, ry = 0, rz = 0;
  for (int t = 0; t < tab.n; ++t) {
    const int dx = tab.d[t][pk.nd - 1];
    const int dy = pk.nd >= 2 ? tab.d[t][pk.nd - 2] : 0;
    const int dz = pk.nd == 3 ? tab.d[t][0] : 0;
    rx = max(rx, dx < 0 ? -dx : dx);
    ry = max(ry, dy < 0 ? -dy : dy);
    rz = max(rz, dz < 0 ? -dz : dz);
  }
  if (rx > 63) return false;
  const int64_t Z = pk.nd == 3 ? pk.size[0] : 1, Y = pk.nd >= 2 ? pk.size[pk.nd - 2] : 1;
  if (Z > 65535 * 64 || Y > 65535 * 64 || pk.ww > 64) return false;
  BinTileGeom q{};
  q.Z = (int)Z;
  q.Y = (int)Y;
  q.ww = (int)pk.ww;
  q.W = (int)pk.W;
  q.pitch = (q.ww + 2) | 1;  // odd: neighbouring rows start in different banks
  q.ry = ry;
  q.rz = rz;
  q.hy = steps * ry;
  q.hz = steps * rz;
  // square-ish tiles of at most 16 x 16 rows (3-D) or 64 rows (2-D), halved
  // until the two buffers fit
  q.ty = (int)(Y < (pk.nd == 3 ? 16 : 64) ? Y : (pk.nd == 3 ? 16 : 64));
  q.tz = (int)(Z < 16 ? Z : 16);
  for (;;) {
    q.by = q.ty + 2 * q.hy;
    q.bz = q.tz + 2 * q.hz;
    if (q.hy <= 64 && q.hz <= 64 && bin_tile_lds(q) <= kBinTileLdsMax) break;
    if (q.ty >= q.tz && q.ty > 1) q.ty = (q.ty + 1) / 2;
    else if (q.tz > 1) q.tz = (q.tz + 1) / 2;
    else return false;
  }
  // halo rows should not outnumber the tile's own rows by much
  if ((int64_t)q.by * q.bz > 8 * (int64_t)q.ty * q.tz) return false;
  if (q.pitch > 1024 || q.by > 1024 || q.bz > 1024 || q.by * q.bz * q.pitch >= 65536) return false;
  q.inv_pitch = 1.0f / (float)q.pitch;
  q.inv_by = 1.0f / (float)q.by;
  q.inv_ww = 1.0f / (float)q.ww;
  q.inv_ty = 1.0f / (float)q.ty;
  q.n = tab.n;
  for (int t = 0; t < kMaxBinTable; ++t) {
    const int u = t < tab.n ? t : 0;
    const int dz = pk.nd == 3 ? tab.d[u][0] : 0;
    const int dy = pk.nd >= 2 ? tab.d[u][pk.nd - 2] : 0;
    const int dx = tab.d[u][pk.nd - 1];
    const int qd = dx < 0 ? -1 : 0;  // word of bit 64 w + dx, relative to w
    q.eo[t] = (dz * q.by + dy) * q.pitch + qd + 1;
    q.es[t] = dx - 64 * qd;
  }
  q.tiles_y = (int)((Y + q.ty - 1) / q.ty);
  q.tiles_z = (int)((Z + q.tz - 1) / q.tz);
  if ((int64_t)q.tiles_y * q.tiles_z > 2147483647) return false;
  *g = q;
  return true;
}

template <int NE>
static void bin_tile_launch(int io, dim3 grid, size_t lds, hipStream_t st, const uint64_t* in,
                            uint64_t* out, const uint8_t* in_b, uint8_t* out_b,
                            const BinTileGeom& g, int dil, int border_value, int steps) {
  const dim3 blk(256);
  switch (io) {
    case 0: hipLaunchKernelGGL((bin_tile_kernel<NE, 0>), grid, blk, lds, st, in, out, in_b, out_b,
                               g, dil, border_value, steps); break;
    case 1: hipLaunchKernelGGL((bin_tile_kernel<NE, 1>), grid, blk, lds, st, in, out, in_b, out_b,
                               g, dil, border_value, steps); break;
    case 2: hipLaunchKernelGGL((bin_tile_kernel<NE, 2>), grid, blk, lds, st, in, out, in_b, out_b,
                               g, dil, border_value, steps); break;
    default: hipLaunchKernelGGL((bin_tile_kernel<NE, 3>), grid, blk, lds, st, in, out, in_b,
                                out_b, g, dil, border_value, steps); break;
  }
}

__global__ void bin_clear_flags(int* flags, int n) {
  if ((int)threadIdx.x < n) flags[threadIdx.x] = 0;
}

template <typename T>
static int bin_run(const T* input, int64_t batch, int ndim, const int64_t* shape,
                   const BinTable& tab, const BinTableWide* wide, int nwide, int border_value,
                   int nops, const int* op_dilation, const int* op_iterations,
                   const uint8_t* mask_bytes, uint8_t* output, void* ws, size_t ws_bytes,
                   hipStream_t st);

}  // namespace tmh

extern "C" size_t tm_binary_workspace_bytes(int64_t batch, int ndim, const int64_t* shape) {
  if (ndim < 1 || ndim > tmh::kMaxNdim || batch < 1) return 0;
  int64_t rows = batch;
  for (int j = 0; j < ndim - 1; ++j) rows *= shape[j];
  const int64_t ww = (shape[ndim - 1] + 63) / 64;
  // ping, pong, orig, mask (packed) + flags
  return (size_t)rows * ww * 8 * 4 + 64 * sizeof(int);  // + 64 change flags
}

namespace tmh {

template <typename T>
static int bin_run(const T* input, int64_t batch, int ndim, const int64_t* shape,
                   const BinTable& tab, const BinTableWide* wide, int nwide, int border_value,
                   int nops, const int* op_dilation, const int* op_iterations,
                   const uint8_t* mask_bytes, uint8_t* output, void* ws, size_t ws_bytes,
                   hipStream_t st) {
  Packed pk{};
  pk.nd = ndim;
  for (int j = 0; j < ndim; ++j) pk.size[j] = shape[j];
  pk.W = shape[ndim - 1];
  pk.ww = (pk.W + 63) / 64;
  pk.rows = 1;
  for (int j = ndim - 2; j >= 0; --j) {
    pk.rstride[j] = pk.rows;
    pk.rows *= shape[j];
  }
  const int64_t nwords = batch * pk.rows * pk.ww;
  const int64_t total = batch * pk.rows * pk.W;
  const size_t need = tm_binary_workspace_bytes(batch, ndim, shape);
  if (!ws || ws_bytes < need) {
    set_error("tm_binary: workspace of %zu bytes is smaller than the %zu required", ws_bytes,
              need);
    return TM_ERR_WORKSPACE;
  }
  uint64_t* bufA = (uint64_t*)ws;
  uint64_t* bufB = bufA + nwords;
  uint64_t* orig = bufB + nwords;
  uint64_t* mpk = orig + nwords;
  int* flags = (int*)(mpk + nwords);
  const dim3 wgrid(grid_for(nwords, 256)), blk(256);
  // one step from src into dst: one launch, or one per chunk of a wide table
  auto step = [&](const uint64_t* src, uint64_t* dst, int dil, const uint64_t* org, int* chg) {
    if (!wide) {
      hipLaunchKernelGGL(bin_step_kernel<BinTable>, wgrid, blk, 0, st, src, dst, pk, nwords, tab,
                         dil, border_value, org, mask_bytes ? mpk : nullptr, chg, 0, 1, 0);
      return;
    }
    for (int c = 0; c < nwide; ++c)
      hipLaunchKernelGGL(bin_step_kernel<BinTableWide>, wgrid, blk, 0, st, src, dst, pk, nwords,
                         wide[c], dil, border_value, org, mask_bytes ? mpk : nullptr, chg,
                         c > 0 ? 1 : 0, c == nwide - 1 ? 1 : 0, 0);
  };
  auto bytes_ok = [&](const void* p) { return pk.W % 64 == 0 && ((uintptr_t)p % 16) == 0; };
  // one launch of up to 8 steps in LDS tiles (g filled; fewer steps when the
  // halo would not fit), or one step per launch (with a mask, or when no
  // tile fits): returns the steps, sets *tiled
  auto chunk_plan = [&](int left, BinTileGeom* g, bool* tiled) {
    int k = left < 8 ? left : 8;
    *tiled = !mask_bytes && batch <= 65535 && !wide;
    while (*tiled && !bin_tile_plan(pk, tab, k, g)) {
      if (k == 1) *tiled = false;
      else k /= 2;
    }
    return *tiled ? k : 1;
  };
  // the first tiled launch packs the bool input itself, the last one writes
  // the bool output itself (rows of whole 64-byte words, 16-byte aligned).
  // Not when input and output overlap (the second pass of an opening or
  // closing reads the output of the first): a launch that both reads and
  // writes the bytes would let one tile overwrite another tile's halo.
  const uintptr_t ib = (uintptr_t)input, ob = (uintptr_t)output;
  const bool overlap = ib < ob + (uintptr_t)total && ob < ib + (uintptr_t)total * sizeof(T);
  bool fuse_in = false;
  if (sizeof(T) == 1 && !overlap && bytes_ok(input) && op_iterations[0] >= 1) {
    BinTileGeom g0{};
    chunk_plan(op_iterations[0], &g0, &fuse_in);
  }
  const bool out_bytes_ok = bytes_ok(output);
  bool fused_out = false;
  if (fuse_in) {
    // no pack pass
  } else if (sizeof(T) == 1 && bytes_ok(input))
    hipLaunchKernelGGL(bin_pack_bytes_kernel, wgrid, blk, 0, st, (const uint8_t*)input, bufA,
                       nwords);
  else
    hipLaunchKernelGGL((bin_pack_kernel<T>), wgrid, blk, 0, st, input, bufA, pk, nwords);
  TM_LAUNCH_CHECK();
  if (mask_bytes) {
    if (bytes_ok(mask_bytes))
      hipLaunchKernelGGL(bin_pack_bytes_kernel, wgrid, blk, 0, st, mask_bytes, mpk, nwords);
    else
      hipLaunchKernelGGL((bin_pack_kernel<uint8_t>), wgrid, blk, 0, st, mask_bytes, mpk, pk,
                         nwords);
    TM_LAUNCH_CHECK();
  }
  uint64_t* cur = bufA;
  uint64_t* nxt = bufB;
  for (int op = 0; op < nops; ++op) {
    const int dil = op_dilation[op];
    const int iters = op_iterations[op];
    if (mask_bytes) {
      TM_CHECK_HIP(hipMemcpyAsync(orig, cur, nwords * 8, hipMemcpyDeviceToDevice, st));
    }
    if (iters >= 1) {
      for (int done = 0; done < iters;) {
        BinTileGeom g{};
        bool tiled = false;
        const int k = chunk_plan(iters - done, &g, &tiled);
        if (tiled) {
          const bool first = op == 0 && done == 0;
          const bool last = op == nops - 1 && done + k == iters;
          const int io = (first && fuse_in ? 1 : 0) | (last && out_bytes_ok ? 2 : 0);
          const dim3 tg(g.tiles_y * g.tiles_z, (unsigned)batch);
          const uint8_t* in_b = (io & 1) ? (const uint8_t*)input : nullptr;
          uint8_t* out_b = (io & 2) ? output : nullptr;
          if (tab.n <= 8)
            bin_tile_launch<8>(io, tg, bin_tile_lds(g), st, cur, nxt, in_b, out_b, g, dil,
                               border_value, k);
          else if (tab.n <= 32)
            bin_tile_launch<32>(io, tg, bin_tile_lds(g), st, cur, nxt, in_b, out_b, g, dil,
                                border_value, k);
          else
            bin_tile_launch<0>(io, tg, bin_tile_lds(g), st, cur, nxt, in_b, out_b, g, dil,
                               border_value, k);
          fused_out = (io & 2) != 0;
        } else {
          step(cur, nxt, dil, mask_bytes ? orig : nullptr, nullptr);
        }
        done += k;
        TM_LAUNCH_CHECK();
        uint64_t* t = cur;
        cur = nxt;
        nxt = t;
      }
    } else {
      // until an iteration changes nothing (at least two iterations, as the
      // reference compares consecutive results); extra iterations past the
      // fixed point are exact no-ops, so the host checks the last step's
      // change flag after every chunk of steps, the chunk doubling from 8 to
      // 64 (the flag words of the workspace): long propagations wait on the
      // host about 7x less often; the steps run past the fixed point are at
      // most one chunk.
      constexpr int kMaxChunk = 64;
      int chunk = 8;
      int64_t done = 0;
      for (;;) {
        hipLaunchKernelGGL(bin_clear_flags, dim3(1), dim3(kMaxChunk), 0, st, flags, chunk);
        TM_LAUNCH_CHECK();
        for (int k = 0; k < chunk; ++k) {
          step(cur, nxt, dil, mask_bytes ? orig : nullptr, flags + k);
          TM_LAUNCH_CHECK();
          uint64_t* t = cur;
          cur = nxt;
          nxt = t;
        }
        done += chunk;
        int last = 0;
        TM_CHECK_HIP(hipMemcpyAsync(&last, flags + chunk - 1, sizeof(int), hipMemcpyDeviceToHost,
                                    st));
        TM_CHECK_HIP(hipStreamSynchronize(st));
        if (last == 0) break;
        if (done > (int64_t)1 << 40) break;
        chunk = chunk * 2 < kMaxChunk ? chunk * 2 : kMaxChunk;
      }
    }
  }
  if (fused_out) return TM_OK;  // the last tiled launch wrote the bytes
  if (pk.W % 16 == 0 && ((uintptr_t)output % 16) == 0)
    hipLaunchKernelGGL(bin_unpack_bytes_kernel, dim3(grid_for(total / 16, 256)), blk, 0, st, cur,
                       output, pk, total / 16);
  else
    hipLaunchKernelGGL(bin_unpack_kernel, dim3(grid_for(total, 256)), blk, 0, st, cur, output,
                       pk, total);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

// Element offsets of the active positions of `structure` (C order over
// se_shape, centre se_shape / 2 + origin): erosion reads p + d_e, dilation
// (`negate`) p - d_e (reference binary_kernel.cu:199-229).  Tables that fit
// BinTable (at most kMaxBinTable elements, offsets within +-127) fill *tab
// and set *narrow; larger ones fill *wide in chunks of kWideChunk.
static int bin_tables(const uint8_t* structure, const int64_t* se_shape, const int64_t* origin,
                      int ndim, bool negate, BinTable* tab, std::vector<BinTableWide>* wide,
                      bool* narrow, int* count) {
  int64_t ne = 1;
  for (int j = 0; j < ndim; ++j) ne *= se_shape[j];
  std::vector<int> offs;  // element-major, kMaxNdim per element
  int n = 0;
  bool fits = true;
  for (int64_t e = 0; e < ne; ++e) {
    if (!structure[e]) continue;
    int64_t rem = e;
    int64_t r[kMaxNdim];
    for (int j = ndim - 1; j >= 0; --j) {
      r[j] = rem % se_shape[j];
      rem /= se_shape[j];
    }
    for (int j = 0; j < kMaxNdim; ++j) {
      int64_t dd = j < ndim ? r[j] - (se_shape[j] / 2 + origin[j]) : 0;
      if (negate) dd = -dd;
      TM_CHECK_ARG(dd >= -32767 && dd <= 32767, "tm_binary: structure offsets beyond +-32767");
      fits = fits && dd >= -127 && dd <= 127;
      offs.push_back((int)dd);
    }
    ++n;
  }
  *count = n;
  *narrow = fits && n <= kMaxBinTable;
  *tab = BinTable{};
  wide->clear();
  if (*narrow) {
    tab->n = n;
    for (int e = 0; e < n; ++e)
      for (int j = 0; j < kMaxNdim; ++j) tab->d[e][j] = (int8_t)offs[e * kMaxNdim + j];
    return TM_OK;
  }
  wide->assign((n + kWideChunk - 1) / kWideChunk, BinTableWide{});
  for (int e = 0; e < n; ++e) {
    BinTableWide& w = (*wide)[e / kWideChunk];
    for (int j = 0; j < kMaxNdim; ++j) w.d[e % kWideChunk][j] = (int16_t)offs[e * kMaxNdim + j];
    w.n = e % kWideChunk + 1;
  }
  return TM_OK;
}

}  // namespace tmh

extern "C" int tm_binary(const void* input, int input_dtype, int64_t batch, int ndim,
                         const int64_t* shape, const uint8_t* structure, const int64_t* se_shape,
                         const int64_t* origin, int border_value, int nops, const int* op_dilation,
                         const int* op_iterations, const uint8_t* mask, uint8_t* output,
                         void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(input && output && structure, "tm_binary: NULL input/output/structure");
  TM_CHECK_ARG(ndim >= 1 && ndim <= kMaxNdim, "tm_binary: ndim must be in 1..8, got %d", ndim);
  TM_CHECK_ARG(batch >= 1, "tm_binary: batch must be >= 1");
  TM_CHECK_ARG(nops >= 1 && op_dilation && op_iterations, "tm_binary: no operation given");
  for (int j = 0; j < ndim; ++j)
    TM_CHECK_ARG(shape[j] >= 1 && se_shape[j] >= 1, "tm_binary: empty axis");
  const hipStream_t st = (hipStream_t)stream;
  int status = TM_OK;
  BinTable tabs[2]{};
  std::vector<BinTableWide> wides[2];
  bool narrow = true;
  int nwide = 0;
  for (int dir = 0; dir < 2; ++dir) {
    int n = 0;
    status = bin_tables(structure, se_shape, origin, ndim, dir != 0, &tabs[dir], &wides[dir],
                        &narrow, &n);
    if (status != TM_OK) return status;
    TM_CHECK_ARG(n > 0, "tm_binary: structure must contain at least one active position");
    nwide = (int)wides[dir].size();
  }
  // Single-direction sequences run in one packed pipeline; mixed sequences
  // (opening/closing) chain pipelines through the output buffer.
  int start = 0;
  const void* cur_in = input;
  int cur_dtype = input_dtype;
  while (start < nops) {
    int end = start + 1;
    while (end < nops && op_dilation[end] == op_dilation[start]) ++end;
    const int dir = op_dilation[start] ? 1 : 0;
    const BinTable& t = tabs[dir];
    const BinTableWide* w = narrow ? nullptr : wides[dir].data();
    if (cur_dtype == TM_U8)
      status = bin_run((const uint8_t*)cur_in, batch, ndim, shape, t, w, nwide, border_value,
                       end - start, op_dilation + start, op_iterations + start, mask, output,
                       workspace, workspace_bytes, st);
    else if (cur_dtype == TM_F32)
      status = bin_run((const float*)cur_in, batch, ndim, shape, t, w, nwide, border_value,
                       end - start, op_dilation + start, op_iterations + start, mask, output,
                       workspace, workspace_bytes, st);
    else {
      set_error("tm_binary: input dtype must be uint8/bool or float32");
      return TM_ERR_ARG;
    }
    if (status != TM_OK) return status;
    cur_in = output;
    cur_dtype = TM_U8;
    start = end;
  }
  return TM_OK;
}

namespace tmh {

__global__ void bin_fill_kernel(uint64_t* out, int64_t n, uint64_t v) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) out[g] = v;
}

// One step of the packed `src` into `dst` with a narrow or a chunked wide
// table: `invert` reads the complement of src, `partial` combines into dst,
// `orig`/`mask`/`changed` as bin_step_kernel.
static int bin_step_into(const uint64_t* src, uint64_t* dst, const Packed& pk, int64_t nwords,
                         bool narrow, const BinTable& tab, const std::vector<BinTableWide>& wide,
                         int dil, int border, const uint64_t* orig, const uint64_t* mask,
                         int* changed, int invert, int partial, hipStream_t st) {
  const dim3 grid(grid_for(nwords, 256)), blk(256);
  if (narrow) {
    hipLaunchKernelGGL(bin_step_kernel<BinTable>, grid, blk, 0, st, src, dst, pk, nwords, tab,
                       dil, border, orig, mask, changed, partial, 1, invert);
  } else {
    for (size_t c = 0; c < wide.size(); ++c)
      hipLaunchKernelGGL(bin_step_kernel<BinTableWide>, grid, blk, 0, st, src, dst, pk, nwords,
                         wide[c], dil, border, orig, mask, changed, (partial || c > 0) ? 1 : 0,
                         c + 1 == wide.size() ? 1 : 0, invert);
  }
  TM_LAUNCH_CHECK();
  return TM_OK;
}

static int bin_erode_into(const uint64_t* src, uint64_t* dst, const Packed& pk, int64_t nwords,
                          bool narrow, const BinTable& tab, const std::vector<BinTableWide>& wide,
                          int invert, int partial, hipStream_t st) {
  return bin_step_into(src, dst, pk, nwords, narrow, tab, wide, 0, 0, nullptr, nullptr, nullptr,
                       invert, partial, st);
}

__global__ void bin_not_kernel(const uint64_t* __restrict__ in, uint64_t* __restrict__ out,
                               int64_t n) {
  const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g < n) out[g] = ~in[g];
}

static Packed bin_packed_shape(int ndim, const int64_t* shape) {
  Packed pk{};
  pk.nd = ndim;
  for (int j = 0; j < ndim; ++j) pk.size[j] = shape[j];
  pk.W = shape[ndim - 1];
  pk.ww = (pk.W + 63) / 64;
  pk.rows = 1;
  for (int j = ndim - 2; j >= 0; --j) {
    pk.rstride[j] = pk.rows;
    pk.rows *= shape[j];
  }
  return pk;
}

static void bin_pack_any(const void* input, int dtype, uint64_t* out, const Packed& pk,
                         int64_t nwords, hipStream_t st) {
  const dim3 grid(grid_for(nwords, 256)), blk(256);
  if (dtype == TM_U8 && pk.W % 64 == 0 && ((uintptr_t)input % 16) == 0)
    hipLaunchKernelGGL(bin_pack_bytes_kernel, grid, blk, 0, st, (const uint8_t*)input, out,
                       nwords);
  else if (dtype == TM_U8)
    hipLaunchKernelGGL((bin_pack_kernel<uint8_t>), grid, blk, 0, st, (const uint8_t*)input, out,
                       pk, nwords);
  else
    hipLaunchKernelGGL((bin_pack_kernel<float>), grid, blk, 0, st, (const float*)input, out, pk,
                       nwords);
}

static void bin_unpack_any(const uint64_t* in, uint8_t* output, const Packed& pk, int64_t total,
                           hipStream_t st) {
  const dim3 blk(256);
  if (pk.W % 16 == 0 && ((uintptr_t)output % 16) == 0)
    hipLaunchKernelGGL(bin_unpack_bytes_kernel, dim3(grid_for(total / 16, 256)), blk, 0, st, in,
                       output, pk, total / 16);
  else
    hipLaunchKernelGGL(bin_unpack_kernel, dim3(grid_for(total, 256)), blk, 0, st, in, output, pk,
                       total);
}

}  // namespace tmh

extern "C" int tm_binary_hit_or_miss(const void* input, int input_dtype, int64_t batch, int ndim,
                                     const int64_t* shape, const uint8_t* structure1,
                                     const int64_t* se1_shape, const int64_t* origin1,
                                     const uint8_t* structure2, const int64_t* se2_shape,
                                     const int64_t* origin2, uint8_t* output, void* workspace,
                                     size_t workspace_bytes, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(input && output && structure1 && structure2,
               "tm_binary_hit_or_miss: NULL input/output/structure");
  TM_CHECK_ARG(ndim >= 1 && ndim <= kMaxNdim, "tm_binary_hit_or_miss: ndim must be in 1..8");
  TM_CHECK_ARG(batch >= 1, "tm_binary_hit_or_miss: batch must be >= 1");
  TM_CHECK_ARG(input_dtype == TM_U8 || input_dtype == TM_F32,
               "tm_binary_hit_or_miss: input dtype must be uint8/bool or float32");
  for (int j = 0; j < ndim; ++j)
    TM_CHECK_ARG(shape[j] >= 1 && se1_shape[j] >= 1 && se2_shape[j] >= 1,
                 "tm_binary_hit_or_miss: empty axis");
  const size_t need = tm_binary_workspace_bytes(batch, ndim, shape);
  if (!workspace || workspace_bytes < need) {
    set_error("tm_binary_hit_or_miss: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, need);
    return TM_ERR_WORKSPACE;
  }
  BinTable t1{}, t2{};
  std::vector<BinTableWide> w1, w2;
  bool nar1 = true, nar2 = true;
  int n1 = 0, n2 = 0;
  int status = bin_tables(structure1, se1_shape, origin1, ndim, false, &t1, &w1, &nar1, &n1);
  if (status != TM_OK) return status;
  status = bin_tables(structure2, se2_shape, origin2, ndim, false, &t2, &w2, &nar2, &n2);
  if (status != TM_OK) return status;
  const Packed pk = bin_packed_shape(ndim, shape);
  const int64_t nwords = batch * pk.rows * pk.ww;
  const int64_t total = batch * pk.rows * pk.W;
  uint64_t* bufA = (uint64_t*)workspace;
  uint64_t* bufB = bufA + nwords;
  const hipStream_t st = (hipStream_t)stream;
  const dim3 wgrid(grid_for(nwords, 256)), blk(256);
  bin_pack_any(input, input_dtype, bufA, pk, nwords, st);
  TM_LAUNCH_CHECK();
  // hit: erosion of x by structure1 (border 0); an empty structure1 hits
  // everywhere.  miss: erosion of not-x by structure2 (border 0), ANDed in.
  if (n1 > 0) {
    status = bin_erode_into(bufA, bufB, pk, nwords, nar1, t1, w1, 0, 0, st);
  } else {
    hipLaunchKernelGGL(bin_fill_kernel, wgrid, blk, 0, st, bufB, nwords, ~0ull);
    TM_LAUNCH_CHECK();
  }
  if (status != TM_OK) return status;
  if (n2 > 0) status = bin_erode_into(bufA, bufB, pk, nwords, nar2, t2, w2, 1, 1, st);
  if (status != TM_OK) return status;
  bin_unpack_any(bufB, output, pk, total, st);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

extern "C" int tm_binary_fill_holes(const void* input, int input_dtype, int64_t batch, int ndim,
                                    const int64_t* shape, const uint8_t* structure,
                                    const int64_t* se_shape, const int64_t* origin,
                                    uint8_t* output, void* workspace, size_t workspace_bytes,
                                    void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(input && output && structure, "tm_binary_fill_holes: NULL input/output/structure");
  TM_CHECK_ARG(ndim >= 1 && ndim <= kMaxNdim, "tm_binary_fill_holes: ndim must be in 1..8");
  TM_CHECK_ARG(batch >= 1, "tm_binary_fill_holes: batch must be >= 1");
  TM_CHECK_ARG(input_dtype == TM_U8 || input_dtype == TM_F32,
               "tm_binary_fill_holes: input dtype must be uint8/bool or float32");
  for (int j = 0; j < ndim; ++j)
    TM_CHECK_ARG(shape[j] >= 1 && se_shape[j] >= 1, "tm_binary_fill_holes: empty axis");
  const size_t need = tm_binary_workspace_bytes(batch, ndim, shape);
  if (!workspace || workspace_bytes < need) {
    set_error("tm_binary_fill_holes: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, need);
    return TM_ERR_WORKSPACE;
  }
  BinTable tab{};
  std::vector<BinTableWide> wide;
  bool narrow = true;
  int n = 0;
  int status = bin_tables(structure, se_shape, origin, ndim, true, &tab, &wide, &narrow, &n);
  if (status != TM_OK) return status;
  TM_CHECK_ARG(n > 0, "tm_binary_fill_holes: structure must contain at least one active position");
  const Packed pk = bin_packed_shape(ndim, shape);
  const int64_t nwords = batch * pk.rows * pk.ww;
  const int64_t total = batch * pk.rows * pk.W;
  uint64_t* cur = (uint64_t*)workspace;
  uint64_t* nxt = cur + nwords;
  uint64_t* zero = nxt + nwords;  // the seed: where(mask, step, seed) restores 0
  uint64_t* bg = zero + nwords;   // the mask: background voxels
  int* flags = (int*)(bg + nwords);
  const hipStream_t st = (hipStream_t)stream;
  const dim3 wgrid(grid_for(nwords, 256)), blk(256);
  bin_pack_any(input, input_dtype, cur, pk, nwords, st);
  hipLaunchKernelGGL(bin_not_kernel, wgrid, blk, 0, st, cur, bg, nwords);
  hipLaunchKernelGGL(bin_fill_kernel, wgrid, blk, 0, st, zero, nwords, 0ull);
  hipLaunchKernelGGL(bin_fill_kernel, wgrid, blk, 0, st, cur, nwords, 0ull);
  TM_LAUNCH_CHECK();
  // dilation of the empty seed from outside the item (border 1), restricted
  // to the background, until a step changes nothing (at least two steps, as
  // the reference compares consecutive results); the host reads the last
  // step's change flag after chunks of 8, 16, 32, then 64 steps
  constexpr int kMaxChunk = 64;
  int chunk = 8;
  for (int64_t done = 0;;) {
    hipLaunchKernelGGL(bin_clear_flags, dim3(1), dim3(kMaxChunk), 0, st, flags, chunk);
    TM_LAUNCH_CHECK();
    for (int k = 0; k < chunk; ++k) {
      status = bin_step_into(cur, nxt, pk, nwords, narrow, tab, wide, 1, 1, zero, bg, flags + k,
                             0, 0, st);
      if (status != TM_OK) return status;
      uint64_t* t = cur;
      cur = nxt;
      nxt = t;
    }
    done += chunk;
    int last = 0;
    TM_CHECK_HIP(hipMemcpyAsync(&last, flags + chunk - 1, sizeof(int), hipMemcpyDeviceToHost, st));
    TM_CHECK_HIP(hipStreamSynchronize(st));
    if (last == 0 || done > ((int64_t)1 << 40)) break;
    chunk = chunk * 2 < kMaxChunk ? chunk * 2 : kMaxChunk;
  }
  // holes and objects: everything the outside did not reach
  hipLaunchKernelGGL(bin_not_kernel, wgrid, blk, 0, st, cur, nxt, nwords);
  TM_LAUNCH_CHECK();
  bin_unpack_any(nxt, output, pk, total, st);
  TM_LAUNCH_CHECK();
  return TM_OK;
}
