// BF16 MFMA building blocks for the warm-start point-set network (gfx950).
//
// No reference counterpart: the reference has no learned component.  The
// network predicts Sinkhorn dual potentials (see tm_warmstart.hip); this file
// holds the hand-written kernels it is built from:
//  * gemm_bf16_kernel: C = A . B with bf16 operands and fp32 accumulation on
//    gfx950's v_mfma_f32_16x16x32_bf16 (64 x 64 output tile per 256-lane
//    workgroup, each wave a 32 x 32 quadrant = 2 x 2 MFMA tiles; K staged
//    through LDS in steps of 32, one MFMA per tile and step).  Both operands
//    sit K-contiguous in LDS (B transposed while staged), so a lane reads its
//    eight K values of a row or column as one 16-byte LDS read; 80-byte rows
//    keep a quarter-wave's reads on distinct banks.  Operands are addressed
//    through (row, col) strides, so the
//    same kernel computes A.B, A.B^T and A^T.B (weight gradients, with the
//    long reduction split over workgroups and summed in a fixed order).
//    Fused epilogues: fp32 store, fp32 + bias, bias (+ per-row-group add) +
//    ReLU to bf16.
//  * small fp32 kernels around it (casts, ReLU masks, column sums, segment
//    means, Adam).
#include "tm_mlp.h"

namespace tmh {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  // round to nearest even (finite inputs; NaN stays NaN)
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

enum GemmEpi {
  EPI_F32 = 0,           // C (f32) = acc  (+ beta * C)
  EPI_F32_BIAS = 1,      // C (f32) = acc + bias[n]
  EPI_BF16_BIAS_RELU = 2 // C (bf16) = relu(acc + bias[n] + rowadd[group(m)][n])
};

constexpr int kBM = 64, kBN = 64, kBK = 32;

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmArgs g) {
  __shared__ __align__(16) uint16_t As[kBM][kBK + 8];  // [m][k]
  __shared__ __align__(16) uint16_t Bs[kBN][kBK + 8];  // [n][k]: B transposed
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int m0 = blockIdx.y * kBM, n0 = blockIdx.x * kBN;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  // K range of this split
  const int kper = ((g.K + g.ksplit - 1) / g.ksplit + kBK - 1) / kBK * kBK;
  const int kb = blockIdx.z * kper;
  const int ke = min(g.K, kb + kper);
  f32x4 acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += kBK) {
    // stage A (64 x 32) and B (32 x 64): 8 elements per lane each
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int idx = tid + e * 256;
      const int r = idx / kBK, c = idx % kBK;
      const int gm = m0 + r, gk = k0 + c;
      As[r][c] = (gm < g.M && gk < ke) ? g.A[gm * g.sam + gk * g.sak] : (uint16_t)0;
      const int rb = idx / kBN, cb = idx % kBN;
      const int gk2 = k0 + rb, gn = n0 + cb;
      Bs[cb][rb] = (gk2 < ke && gn < g.N) ? g.B[gk2 * g.sbk + gn * g.sbn] : (uint16_t)0;
    }
    __syncthreads();
    {
      // lane: row / column lane % 16 of its 16 x 16 tile, K values
      // 8 (lane / 16) .. 8 (lane / 16) + 7 of this step
      bf16x8 af[2], bfr[2];
      const int kk = 8 * (lane >> 4);
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        af[t] = *(const bf16x8*)&As[wm + t * 16 + (lane & 15)][kk];
        bfr[t] = *(const bf16x8*)&Bs[wn + t * 16 + (lane & 15)][kk];
      }
#pragma unroll
      for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[a], bfr[b], acc[a][b], 0, 0, 0);
    }
    __syncthreads();
  }
  // epilogue: lane holds rows 4*(lane/16)+r, column lane%16 of each tile
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int gm = m0 + wm + a * 16 + 4 * (lane >> 4) + r;
        const int gn = n0 + wn + b * 16 + (lane & 15);
        if (gm >= g.M || gn >= g.N) continue;
        float v = acc[a][b][r];
        if (g.ksplit > 1) {
          ((float*)g.C)[(int64_t)blockIdx.z * g.M * g.ldc + (int64_t)gm * g.ldc + gn] = v;
          continue;
        }
        if (g.epi == EPI_F32) {
          float* c = (float*)g.C + (int64_t)gm * g.ldc + gn;
          *c = g.beta != 0.f ? v + g.beta * *c : v;
        } else if (g.epi == EPI_F32_BIAS) {
          ((float*)g.C)[(int64_t)gm * g.ldc + gn] = v + (g.bias ? g.bias[gn] : 0.f);
        } else {
          v += g.bias ? g.bias[gn] : 0.f;
          if (g.rowadd) v += g.rowadd[(int64_t)g.rowgroup[gm] * g.N + gn];
          ((uint16_t*)g.C)[(int64_t)gm * g.ldc + gn] = f2bf(v > 0.f ? v : 0.f);
        }
      }
}

// Sums ksplit partial products in split order: C = sum_z P[z] (+ beta C).
__global__ __launch_bounds__(256) void splitk_reduce_kernel(const float* __restrict__ P,
                                                            float* __restrict__ C, int64_t mn,
                                                            int ksplit, float beta) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= mn) return;
  float s = 0.f;
  for (int z = 0; z < ksplit; ++z) s += P[(int64_t)z * mn + i];
  C[i] = beta != 0.f ? s + beta * C[i] : s;
}

int launch_gemm(const GemmArgs& g0, float* splitk_ws, size_t splitk_ws_elems, hipStream_t st) {
  GemmArgs g = g0;
  const dim3 blk(256);
  const unsigned gx = (g.N + kBN - 1) / kBN, gy = (g.M + kBM - 1) / kBM;
  // long reductions with few output tiles: split K over workgroups
  int ks = 1;
  if (g.epi == EPI_F32 && splitk_ws) {
    const int tiles = gx * gy;
    while (ks < 64 && tiles * ks < 512 && g.K / (ks * 2) >= 256) ks *= 2;
    if ((size_t)ks * g.M * g.N > splitk_ws_elems) ks = 1;
  }
  if (ks > 1 && g.ldc != g.N) ks = 1;  // split-K reduces into a dense output
  if (ks > 1) {
    GemmArgs p = g;
    p.C = splitk_ws;
    p.ldc = g.N;
    p.ksplit = ks;
    hipLaunchKernelGGL(gemm_bf16_kernel, dim3(gx, gy, ks), blk, 0, st, p);
    TM_LAUNCH_CHECK();
    const int64_t mn = (int64_t)g.M * g.N;
    hipLaunchKernelGGL(splitk_reduce_kernel, dim3(grid_for(mn, 256)), blk, 0, st, splitk_ws,
                       (float*)g.C, mn, ks, g.beta);
    TM_LAUNCH_CHECK();
    return TM_OK;
  }
  g.ksplit = 1;
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3(gx, gy, 1), blk, 0, st, g);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

__global__ void cast_bf16_kernel(const float* __restrict__ x, uint16_t* __restrict__ y,
                                 int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

}  // namespace tmh

// Test/utility entry: C(f32, M x N, row-major) = A (bf16) . B (bf16) with
// strides; used by the tests to check the MFMA lane layout against torch.
extern "C" int tm_mlp_gemm_bf16(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B,
                                int64_t sbk, int64_t sbn, float* C, int M, int N, int K,
                                float* splitk_ws, size_t splitk_ws_elems, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(A && B && C && M > 0 && N > 0 && K > 0, "tm_mlp_gemm_bf16: bad arguments");
  GemmArgs g{A, sam, sak, B, sbk, sbn, C, N, M, N, K, EPI_F32, nullptr, nullptr, nullptr, 0.f, 1};
  return launch_gemm(g, splitk_ws, splitk_ws_elems, (hipStream_t)stream);
}

extern "C" int tm_cast_bf16(const float* x, uint16_t* y, int64_t n, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(x && y && n >= 0, "tm_cast_bf16: bad arguments");
  if (n == 0) return TM_OK;
  hipLaunchKernelGGL(cast_bf16_kernel, dim3(grid_for(n, 256)), dim3(256), 0,
                     (hipStream_t)stream, x, y, n);
  TM_LAUNCH_CHECK();
  return TM_OK;
}
