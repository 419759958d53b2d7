// Warm-start network for point-cloud Sinkhorn (gfx950, BF16 MFMA, fp32
// accumulation).  No reference counterpart.
//
// A DeepSets model predicts the dual potential f of every x point of a pair
// from both clouds; plain Sinkhorn then starts from it and still runs to the
// same marginal tolerance (it decides convergence and is the cold-start fallback).
//
// Normalisation (per pair, computed on the device): c = mean of y,
// s2 = mean_ij |x_i - y_j|^2 = mean|x - c|^2 + mean|y - c|^2 (closed form),
// u = (p - c) / sqrt(s2), r = log(eps / s2).  The network predicts f / s2.
//
//   phi  : [u, flag] (4)  -> 128 -> 128          (ReLU, shared by x and y points)
//   pool : ex = mean_i phi(x_i), ey = mean_j phi(y_j)
//   head : [phi(x_i), ex, ey, r] (385) -> 256 -> 256 -> 1   (ReLU hidden)
// The pair-constant part of the head's first layer (ex, ey, r, bias) is one
// 256-vector per pair added in the GEMM epilogue.  182,017 parameters.
//
// Training loss (per batch of pairs): mean over pairs of w_p times the mean
// squared error of (prediction - target) / s2 after removing its per-pair
// mean (the dual potentials are defined up to a constant shift between f and
// g), w_p = (s2 / eps)^alpha.
#include <vector>

#include "tm_mlp.h"

namespace tmh {

constexpr int kIn = 4, kH = 128, kG = 256;

// parameter layout (fp32, flat)
constexpr int64_t oW0 = 0, oB0 = oW0 + kIn * kH;
constexpr int64_t oW1 = oB0 + kH, oB1 = oW1 + kH * kH;
constexpr int64_t oW2a = oB1 + kH, oW2x = oW2a + kH * kG, oW2y = oW2x + kH * kG;
constexpr int64_t ow2l = oW2y + kH * kG, oB2 = ow2l + kG;
constexpr int64_t oW3 = oB2 + kG, oB3 = oW3 + kG * kG;
constexpr int64_t ow4 = oB3 + kG, oB4 = ow4 + kG;
constexpr int64_t kNParams = oB4 + 1;

__device__ __forceinline__ uint16_t f2bf_ws(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f_ws(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

struct WsPair {
  int64_t xoff, yoff;
  int32_t n, m;
  float eps;
  int32_t pad;
};

// per pair: centre (3, padded to 4), s2, r = log(eps / s2)
__global__ __launch_bounds__(256) void ws_stats_kernel(const float* __restrict__ x,
                                                       const float* __restrict__ y,
                                                       const WsPair* __restrict__ pairs,
                                                       float* __restrict__ stats) {
  __shared__ float red[256];
  const WsPair pr = pairs[blockIdx.x];
  float c[3];
  for (int d = 0; d < 3; ++d) {
    float s = 0.f;
    for (int j = threadIdx.x; j < pr.m; j += blockDim.x) s += y[(pr.yoff + j) * 3 + d];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
      __syncthreads();
    }
    c[d] = red[0] / (float)pr.m;
    __syncthreads();
  }
  float vx = 0.f, vy = 0.f;
  for (int i = threadIdx.x; i < pr.n; i += blockDim.x) {
    for (int d = 0; d < 3; ++d) {
      const float t = x[(pr.xoff + i) * 3 + d] - c[d];
      vx = fmaf(t, t, vx);
    }
  }
  for (int j = threadIdx.x; j < pr.m; j += blockDim.x) {
    for (int d = 0; d < 3; ++d) {
      const float t = y[(pr.yoff + j) * 3 + d] - c[d];
      vy = fmaf(t, t, vy);
    }
  }
  red[threadIdx.x] = vx / (float)pr.n + vy / (float)pr.m;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float s2 = fmaxf(red[0], 1e-30f);
    float* st = stats + blockIdx.x * 8;
    st[0] = c[0];
    st[1] = c[1];
    st[2] = c[2];
    st[3] = s2;
    st[4] = logf(pr.eps / s2);
  }
}

// network inputs [u, flag] in bf16 and the pair index of every point
__global__ void ws_prep_kernel(const float* __restrict__ p, int64_t npts,
                               const int32_t* __restrict__ group, const float* __restrict__ stats,
                               float flag, uint16_t* __restrict__ U) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  const float* st = stats + group[i] * 8;
  const float inv = rsqrtf(st[3]);
  for (int d = 0; d < 3; ++d) U[i * kIn + d] = f2bf_ws((p[i * 3 + d] - st[d]) * inv);
  U[i * kIn + 3] = f2bf_ws(flag);
}

// E[p][c] = mean over the pair's rows of H (bf16, [rows][kH]); one block per pair
__global__ __launch_bounds__(kH) void ws_pool_kernel(const uint16_t* __restrict__ H,
                                                     const WsPair* __restrict__ pairs, int use_y,
                                                     float* __restrict__ E) {
  const WsPair pr = pairs[blockIdx.x];
  const int64_t off = use_y ? pr.yoff : pr.xoff;
  const int cnt = use_y ? pr.m : pr.n;
  const int c = threadIdx.x;
  float s = 0.f;
  for (int r = 0; r < cnt; ++r) s += bf2f_ws(H[(off + r) * kH + c]);
  E[blockIdx.x * kH + c] = s / (float)cnt;
}

// T[p][n] = ex.W2x[:, n] + ey.W2y[:, n] + r w2l[n] + b2[n]
__global__ __launch_bounds__(kG) void ws_pairterm_kernel(const float* __restrict__ Ex,
                                                         const float* __restrict__ Ey,
                                                         const float* __restrict__ stats,
                                                         const float* __restrict__ P,
                                                         float* __restrict__ T) {
  const int p = blockIdx.x, n = threadIdx.x;
  float s = P[oB2 + n] + stats[p * 8 + 4] * P[ow2l + n];
  for (int k = 0; k < kH; ++k) {
    s = fmaf(Ex[p * kH + k], P[oW2x + (int64_t)k * kG + n], s);
    s = fmaf(Ey[p * kH + k], P[oW2y + (int64_t)k * kG + n], s);
  }
  T[p * kG + n] = s;
}

// out_i = A2_i . w4 + b4 (one wave per row)
__global__ __launch_bounds__(256) void ws_out_kernel(const uint16_t* __restrict__ A2,
                                                     const float* __restrict__ P, int64_t rows,
                                                     float* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= rows) return;
  float s = 0.f;
  for (int k = lane; k < kG; k += 64) s = fmaf(bf2f_ws(A2[row * kG + k]), P[ow4 + k], s);
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
  if (lane == 0) out[row] = s + P[oB4];
}

// predicted potentials in cost units: f = out * s2
__global__ void ws_scale_kernel(const float* __restrict__ out, const int32_t* __restrict__ group,
                                const float* __restrict__ stats, int64_t n, float* __restrict__ f) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f[i] = out[i] * stats[group[i] * 8 + 3];
}

// loss and its gradient, one block per pair (deterministic)
__global__ __launch_bounds__(256) void ws_loss_kernel(const float* __restrict__ out,
                                                      const float* __restrict__ target,
                                                      const WsPair* __restrict__ pairs,
                                                      const float* __restrict__ stats, int npairs,
                                                      float alpha, float* __restrict__ dout,
                                                      float* __restrict__ lpair) {
  __shared__ float red[256];
  const WsPair pr = pairs[blockIdx.x];
  const float s2 = stats[blockIdx.x * 8 + 3];
  float s = 0.f;
  for (int i = threadIdx.x; i < pr.n; i += blockDim.x)
    s += out[pr.xoff + i] - target[pr.xoff + i] / s2;
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const float mean = red[0] / (float)pr.n;
  __syncthreads();
  float l = 0.f;
  // pair weight (s2 / eps)^alpha: alpha = 1 measures errors in units of eps
  // (what the Sinkhorn iteration count depends on), alpha = 0 in units of s2
  const float wp = powf(s2 / pr.eps, alpha);
  const float gs = 2.f * wp / ((float)pr.n * (float)npairs);
  for (int i = threadIdx.x; i < pr.n; i += blockDim.x) {
    const float rr = out[pr.xoff + i] - target[pr.xoff + i] / s2 - mean;
    l = fmaf(rr, rr, l);
    dout[pr.xoff + i] = gs * rr;
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) lpair[blockIdx.x] = wp * red[0] / (float)pr.n / (float)npairs;
}

__global__ void ws_sum_kernel(const float* __restrict__ v, int n, float* __restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    float s = 0.f;
    for (int i = 0; i < n; ++i) s += v[i];
    *out = s;
  }
}

// dZ = dA (f32) * (A > 0) with A the bf16 post-activation; writes f32 and bf16
__global__ void ws_relu_bwd_kernel(const float* __restrict__ dA, const uint16_t* __restrict__ A,
                                   int64_t n, float* __restrict__ dZ, uint16_t* __restrict__ dZb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float v = (A[i] & 0x7fffu) && !(A[i] & 0x8000u) ? dA[i] : 0.f;
  dZ[i] = v;
  dZb[i] = f2bf_ws(v);
}

// dA2[i][n] = dout_i w4[n]
__global__ void ws_out_bwd_kernel(const float* __restrict__ dout, const float* __restrict__ P,
                                  int64_t rows, float* __restrict__ dA) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * kG) return;
  dA[t] = dout[t / kG] * P[ow4 + t % kG];
}

// Column sums with optional row weights, two levels for determinism:
// part[b][c] = sum over rows of block b of w_r * X[r][c]; out[c] (+)= sum_b part.
template <typename T>
__device__ __forceinline__ float ldv(const T* p, int64_t i);
template <>
__device__ __forceinline__ float ldv<float>(const float* p, int64_t i) { return p[i]; }
template <>
__device__ __forceinline__ float ldv<uint16_t>(const uint16_t* p, int64_t i) {
  return bf2f_ws(p[i]);
}

constexpr int kColRows = 512;

template <typename T>
__global__ __launch_bounds__(256) void ws_colsum_part_kernel(const T* __restrict__ X,
                                                             const float* __restrict__ w,
                                                             int64_t rows, int cols,
                                                             float* __restrict__ part) {
  const int64_t r0 = (int64_t)blockIdx.x * kColRows;
  const int64_t r1 = r0 + kColRows < rows ? r0 + kColRows : rows;
  for (int c = threadIdx.x; c < cols; c += blockDim.x) {
    float s = 0.f;
    for (int64_t r = r0; r < r1; ++r) s += (w ? w[r] : 1.f) * ldv<T>(X, r * cols + c);
    part[(int64_t)blockIdx.x * cols + c] = s;
  }
}

__global__ void ws_colsum_final_kernel(const float* __restrict__ part, int nb, int cols,
                                       float* __restrict__ out, int accumulate) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += part[(int64_t)b * cols + c];
  out[c] = accumulate ? out[c] + s : s;
}

// dT[p][n] = sum over the pair's x rows of dZ1[i][n]
__global__ __launch_bounds__(kG) void ws_segsum_kernel(const float* __restrict__ dZ,
                                                       const WsPair* __restrict__ pairs,
                                                       float* __restrict__ dT) {
  const WsPair pr = pairs[blockIdx.x];
  const int n = threadIdx.x;
  float s = 0.f;
  for (int i = 0; i < pr.n; ++i) s += dZ[(pr.xoff + i) * kG + n];
  dT[blockIdx.x * kG + n] = s;
}

// pair-level weight gradients and dEx / dEy (P is small)
__global__ __launch_bounds__(kG) void ws_pair_bwd_kernel(const float* __restrict__ Ex,
                                                         const float* __restrict__ Ey,
                                                         const float* __restrict__ stats,
                                                         const float* __restrict__ dT, int npairs,
                                                         float* __restrict__ grad) {
  // block k < kH: row k of dW2x and dW2y; block kH: dw2l and db2
  const int k = blockIdx.x, n = threadIdx.x;
  if (k < kH) {
    float sx = 0.f, sy = 0.f;
    for (int p = 0; p < npairs; ++p) {
      sx = fmaf(Ex[p * kH + k], dT[p * kG + n], sx);
      sy = fmaf(Ey[p * kH + k], dT[p * kG + n], sy);
    }
    grad[oW2x + (int64_t)k * kG + n] = sx;
    grad[oW2y + (int64_t)k * kG + n] = sy;
  } else {
    float sl = 0.f, sb = 0.f;
    for (int p = 0; p < npairs; ++p) {
      sl = fmaf(stats[p * 8 + 4], dT[p * kG + n], sl);
      sb += dT[p * kG + n];
    }
    grad[ow2l + n] = sl;
    grad[oB2 + n] = sb;
  }
}

__global__ __launch_bounds__(kH) void ws_dE_kernel(const float* __restrict__ dT,
                                                   const float* __restrict__ P,
                                                   float* __restrict__ dEx,
                                                   float* __restrict__ dEy) {
  const int p = blockIdx.x, k = threadIdx.x;
  float sx = 0.f, sy = 0.f;
  for (int n = 0; n < kG; ++n) {
    sx = fmaf(dT[p * kG + n], P[oW2x + (int64_t)k * kG + n], sx);
    sy = fmaf(dT[p * kG + n], P[oW2y + (int64_t)k * kG + n], sy);
  }
  dEx[p * kH + k] = sx;
  dEy[p * kH + k] = sy;
}

// dH[i][c] += dE[group(i)][c] / count(group(i))
__global__ void ws_unpool_kernel(float* __restrict__ dH, const float* __restrict__ dE,
                                 const int32_t* __restrict__ group,
                                 const WsPair* __restrict__ pairs, int use_y, int64_t rows,
                                 int accumulate) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= rows * kH) return;
  const int64_t i = t / kH;
  const int c = (int)(t % kH);
  const int p = group[i];
  const float cnt = (float)(use_y ? pairs[p].m : pairs[p].n);
  const float v = dE[p * kH + c] / cnt;
  dH[t] = accumulate ? dH[t] + v : v;
}

__global__ void ws_adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                               float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                               float b1, float b2, float eps, float c1, float c2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * ((mi / c1) / (sqrtf(vi / c2) + eps) + wd * p[i]);
}

__global__ void ws_cast_kernel(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf_ws(x[i]);
}

// -------------------------------------------------------------- workspace --
struct WsLayout {
  size_t bytes = 0;
  size_t pairs, stats, gx, gy, Ux, Uy, H1x, H1y, H2x, H2y, Ex, Ey, T, A1, A2, out, Wb;
  size_t dout, lpair, dA, dZf, dZb, dH2x, dH2y, dZ2x, dZ2y, dEx, dEy, dT, part, splitk;
  int64_t partElems, splitkElems;
  void build(int64_t nx, int64_t ny, int np) {
    auto take = [&](size_t b) {
      const size_t o = bytes;
      bytes += (b + 255) & ~(size_t)255;
      return o;
    };
    const int64_t nmax = nx > ny ? nx : ny;
    pairs = take(sizeof(WsPair) * np);
    stats = take(sizeof(float) * 8 * np);
    gx = take(sizeof(int32_t) * nx);
    gy = take(sizeof(int32_t) * ny);
    Ux = take(2 * kIn * nx);
    Uy = take(2 * kIn * ny);
    H1x = take(2 * kH * nx);
    H1y = take(2 * kH * ny);
    H2x = take(2 * kH * nx);
    H2y = take(2 * kH * ny);
    Ex = take(4 * kH * np);
    Ey = take(4 * kH * np);
    T = take(4 * kG * np);
    A1 = take(2 * kG * nx);
    A2 = take(2 * kG * nx);
    out = take(4 * nx);
    Wb = take(2 * kNParams);
    dout = take(4 * nx);
    lpair = take(4 * (np + 1));
    dA = take(4 * kG * nmax);
    dZf = take(4 * kG * nmax);
    dZb = take(2 * kG * nmax);
    dH2x = take(4 * kH * nx);
    dH2y = take(4 * kH * ny);
    dZ2x = take(2 * kH * nmax);
    dZ2y = take(2 * kH * nmax);
    dEx = take(4 * kH * np);
    dEy = take(4 * kH * np);
    dT = take(4 * kG * np);
    partElems = ((nmax + kColRows - 1) / kColRows) * kG;
    part = take(4 * partElems);
    splitkElems = 64 * kG * kG;
    splitk = take(4 * splitkElems);
  }
};

struct WsCtx {
  char* base;
  WsLayout L;
  int np;
  int64_t nx, ny;
  hipStream_t st;
  template <typename T>
  T* at(size_t off) const { return (T*)(base + off); }
};

static int ws_setup(WsCtx& c, const float* x, const float* y, int np, const int32_t* n,
                    const int32_t* m, const float* eps, void* ws, size_t ws_bytes, hipStream_t st) {
  int64_t nx = 0, ny = 0;
  std::vector<WsPair> hp(np);
  for (int p = 0; p < np; ++p) {
    TM_CHECK_ARG(n[p] >= 1 && m[p] >= 1, "warm start: empty cloud in pair %d", p);
    TM_CHECK_ARG(eps[p] > 0.f, "warm start: eps must be > 0");
    hp[p] = WsPair{nx, ny, n[p], m[p], eps[p], 0};
    nx += n[p];
    ny += m[p];
  }
  c.L.build(nx, ny, np);
  if (!ws || ws_bytes < c.L.bytes) {
    set_error("warm start: workspace of %zu bytes is smaller than the %zu required", ws_bytes,
              c.L.bytes);
    return TM_ERR_WORKSPACE;
  }
  c.base = (char*)ws;
  c.np = np;
  c.nx = nx;
  c.ny = ny;
  c.st = st;
  std::vector<int32_t> gxh(nx), gyh(ny);
  for (int p = 0; p < np; ++p) {
    for (int i = 0; i < n[p]; ++i) gxh[hp[p].xoff + i] = p;
    for (int j = 0; j < m[p]; ++j) gyh[hp[p].yoff + j] = p;
  }
  TM_CHECK_HIP(hipMemcpyAsync(c.at<WsPair>(c.L.pairs), hp.data(), sizeof(WsPair) * np,
                              hipMemcpyHostToDevice, st));
  TM_CHECK_HIP(hipMemcpyAsync(c.at<int32_t>(c.L.gx), gxh.data(), sizeof(int32_t) * nx,
                              hipMemcpyHostToDevice, st));
  TM_CHECK_HIP(hipMemcpyAsync(c.at<int32_t>(c.L.gy), gyh.data(), sizeof(int32_t) * ny,
                              hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(ws_stats_kernel, dim3(np), dim3(256), 0, st, x, y, c.at<WsPair>(c.L.pairs),
                     c.at<float>(c.L.stats));
  TM_LAUNCH_CHECK();
  return TM_OK;
}

static GemmArgs gemm_nn(const uint16_t* A, int64_t lda, const uint16_t* B, int64_t ldb, void* C,
                        int M, int N, int K, int epi, const float* bias, const float* rowadd,
                        const int32_t* rowgroup) {
  return GemmArgs{A, lda, 1, B, ldb, 1, C, N, M, N, K, epi, bias, rowadd, rowgroup, 0.f, 1};
}

static int ws_forward(WsCtx& c, const float* P, const float* x, const float* y) {
  hipStream_t st = c.st;
  const WsLayout& L = c.L;
  uint16_t* Wb = c.at<uint16_t>(L.Wb);
  hipLaunchKernelGGL(ws_cast_kernel, dim3(grid_for(kNParams, 256)), dim3(256), 0, st, P, Wb,
                     kNParams);
  hipLaunchKernelGGL(ws_prep_kernel, dim3(grid_for(c.nx, 256)), dim3(256), 0, st, x, c.nx,
                     c.at<int32_t>(L.gx), c.at<float>(L.stats), 0.0f, c.at<uint16_t>(L.Ux));
  hipLaunchKernelGGL(ws_prep_kernel, dim3(grid_for(c.ny, 256)), dim3(256), 0, st, y, c.ny,
                     c.at<int32_t>(L.gy), c.at<float>(L.stats), 1.0f, c.at<uint16_t>(L.Uy));
  TM_LAUNCH_CHECK();
  const int nx = (int)c.nx, ny = (int)c.ny;
  int s = TM_OK;
  // phi
  s |= launch_gemm(gemm_nn(c.at<uint16_t>(L.Ux), kIn, Wb + oW0, kH, c.at<uint16_t>(L.H1x), nx, kH,
                           kIn, 2, P + oB0, nullptr, nullptr), nullptr, 0, st);
  s |= launch_gemm(gemm_nn(c.at<uint16_t>(L.Uy), kIn, Wb + oW0, kH, c.at<uint16_t>(L.H1y), ny, kH,
                           kIn, 2, P + oB0, nullptr, nullptr), nullptr, 0, st);
  s |= launch_gemm(gemm_nn(c.at<uint16_t>(L.H1x), kH, Wb + oW1, kH, c.at<uint16_t>(L.H2x), nx, kH,
                           kH, 2, P + oB1, nullptr, nullptr), nullptr, 0, st);
  s |= launch_gemm(gemm_nn(c.at<uint16_t>(L.H1y), kH, Wb + oW1, kH, c.at<uint16_t>(L.H2y), ny, kH,
                           kH, 2, P + oB1, nullptr, nullptr), nullptr, 0, st);
  // pool + pair term
  hipLaunchKernelGGL(ws_pool_kernel, dim3(c.np), dim3(kH), 0, st, c.at<uint16_t>(L.H2x),
                     c.at<WsPair>(L.pairs), 0, c.at<float>(L.Ex));
  hipLaunchKernelGGL(ws_pool_kernel, dim3(c.np), dim3(kH), 0, st, c.at<uint16_t>(L.H2y),
                     c.at<WsPair>(L.pairs), 1, c.at<float>(L.Ey));
  hipLaunchKernelGGL(ws_pairterm_kernel, dim3(c.np), dim3(kG), 0, st, c.at<float>(L.Ex),
                     c.at<float>(L.Ey), c.at<float>(L.stats), P, c.at<float>(L.T));
  TM_LAUNCH_CHECK();
  // head
  s |= launch_gemm(gemm_nn(c.at<uint16_t>(L.H2x), kH, Wb + oW2a, kG, c.at<uint16_t>(L.A1), nx, kG,
                           kH, 2, nullptr, c.at<float>(L.T), c.at<int32_t>(L.gx)), nullptr, 0, st);
  s |= launch_gemm(gemm_nn(c.at<uint16_t>(L.A1), kG, Wb + oW3, kG, c.at<uint16_t>(L.A2), nx, kG,
                           kG, 2, P + oB3, nullptr, nullptr), nullptr, 0, st);
  hipLaunchKernelGGL(ws_out_kernel, dim3((unsigned)((c.nx + 3) / 4)), dim3(256), 0, st,
                     c.at<uint16_t>(L.A2), P, c.nx, c.at<float>(L.out));
  TM_LAUNCH_CHECK();
  return s ? TM_ERR_HIP : TM_OK;
}

static int colsum(WsCtx& c, const void* X, bool bf16, const float* w, int64_t rows, int cols,
                  float* out, int accumulate) {
  const int nb = (int)((rows + kColRows - 1) / kColRows);
  float* part = c.at<float>(c.L.part);
  if (bf16)
    hipLaunchKernelGGL(ws_colsum_part_kernel<uint16_t>, dim3(nb), dim3(256), 0, c.st,
                       (const uint16_t*)X, w, rows, cols, part);
  else
    hipLaunchKernelGGL(ws_colsum_part_kernel<float>, dim3(nb), dim3(256), 0, c.st,
                       (const float*)X, w, rows, cols, part);
  hipLaunchKernelGGL(ws_colsum_final_kernel, dim3((cols + 255) / 256), dim3(256), 0, c.st, part,
                     nb, cols, out, accumulate);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

// weight gradient dW[k][n] (+)= sum_r X[r][k] dZ[r][n] (X, dZ bf16, row-major)
static int wgrad(WsCtx& c, const uint16_t* X, int kin, const uint16_t* dZ, int nout, int64_t rows,
                 float* dW, bool accumulate) {
  GemmArgs g{X, 1, kin, dZ, nout, 1, dW, nout, kin, nout, (int)rows, 0, nullptr, nullptr,
             nullptr, accumulate ? 1.f : 0.f, 1};
  return launch_gemm(g, c.at<float>(c.L.splitk), (size_t)c.L.splitkElems, c.st);
}

// dX[r][k] = sum_n dZ[r][n] W[k][n]  (dZ bf16 [rows][nout], W bf16 [kin][nout])
static int dgrad(WsCtx& c, const uint16_t* dZ, int nout, const uint16_t* W, int kin, int64_t rows,
                 float* dX, bool accumulate) {
  GemmArgs g{dZ, nout, 1, W, 1, nout, dX, kin, (int)rows, kin, nout, 0, nullptr, nullptr, nullptr,
             accumulate ? 1.f : 0.f, 1};
  return launch_gemm(g, nullptr, 0, c.st);
}

static int ws_backward(WsCtx& c, const float* P, const float* target, float alpha, float* grad,
                       float* loss) {
  hipStream_t st = c.st;
  const WsLayout& L = c.L;
  const uint16_t* Wb = c.at<uint16_t>(L.Wb);
  const int64_t nx = c.nx, ny = c.ny;
  int s = TM_OK;
  float* dout = c.at<float>(L.dout);
  hipLaunchKernelGGL(ws_loss_kernel, dim3(c.np), dim3(256), 0, st, c.at<float>(L.out), target,
                     c.at<WsPair>(L.pairs), c.at<float>(L.stats), c.np, alpha, dout,
                     c.at<float>(L.lpair));
  hipLaunchKernelGGL(ws_sum_kernel, dim3(1), dim3(64), 0, st, c.at<float>(L.lpair), c.np, loss);
  TM_LAUNCH_CHECK();
  // output layer: dw4 = sum_i dout_i A2_i, db4 = sum dout, dA2 = dout w4
  s |= colsum(c, c.at<uint16_t>(L.A2), true, dout, nx, kG, grad + ow4, 0);
  s |= colsum(c, dout, false, nullptr, nx, 1, grad + oB4, 0);
  float* dA = c.at<float>(L.dA);
  float* dZf = c.at<float>(L.dZf);
  uint16_t* dZb = c.at<uint16_t>(L.dZb);
  hipLaunchKernelGGL(ws_out_bwd_kernel, dim3(grid_for(nx * kG, 256)), dim3(256), 0, st, dout, P,
                     nx, dA);
  // layer 3: A2 = relu(A1 W3 + b3)
  hipLaunchKernelGGL(ws_relu_bwd_kernel, dim3(grid_for(nx * kG, 256)), dim3(256), 0, st, dA,
                     c.at<uint16_t>(L.A2), nx * kG, dZf, dZb);
  TM_LAUNCH_CHECK();
  s |= colsum(c, dZf, false, nullptr, nx, kG, grad + oB3, 0);
  s |= wgrad(c, c.at<uint16_t>(L.A1), kG, dZb, kG, nx, grad + oW3, false);
  s |= dgrad(c, dZb, kG, Wb + oW3, kG, nx, dA, false);
  // layer 2 (head input): A1 = relu(H2x W2a + T[pair])
  hipLaunchKernelGGL(ws_relu_bwd_kernel, dim3(grid_for(nx * kG, 256)), dim3(256), 0, st, dA,
                     c.at<uint16_t>(L.A1), nx * kG, dZf, dZb);
  TM_LAUNCH_CHECK();
  s |= wgrad(c, c.at<uint16_t>(L.H2x), kH, dZb, kG, nx, grad + oW2a, false);
  float* dH2x = c.at<float>(L.dH2x);
  float* dH2y = c.at<float>(L.dH2y);
  s |= dgrad(c, dZb, kG, Wb + oW2a, kH, nx, dH2x, false);
  hipLaunchKernelGGL(ws_segsum_kernel, dim3(c.np), dim3(kG), 0, st, dZf, c.at<WsPair>(L.pairs),
                     c.at<float>(L.dT));
  hipLaunchKernelGGL(ws_pair_bwd_kernel, dim3(kH + 1), dim3(kG), 0, st, c.at<float>(L.Ex),
                     c.at<float>(L.Ey), c.at<float>(L.stats), c.at<float>(L.dT), c.np, grad);
  hipLaunchKernelGGL(ws_dE_kernel, dim3(c.np), dim3(kH), 0, st, c.at<float>(L.dT), P,
                     c.at<float>(L.dEx), c.at<float>(L.dEy));
  hipLaunchKernelGGL(ws_unpool_kernel, dim3(grid_for(nx * kH, 256)), dim3(256), 0, st, dH2x,
                     c.at<float>(L.dEx), c.at<int32_t>(L.gx), c.at<WsPair>(L.pairs), 0, nx, 1);
  hipLaunchKernelGGL(ws_unpool_kernel, dim3(grid_for(ny * kH, 256)), dim3(256), 0, st, dH2y,
                     c.at<float>(L.dEy), c.at<int32_t>(L.gy), c.at<WsPair>(L.pairs), 1, ny, 0);
  TM_LAUNCH_CHECK();
  // phi layer 2 (shared by x and y): H2 = relu(H1 W1 + b1)
  uint16_t* dZx = c.at<uint16_t>(L.dZ2x);
  uint16_t* dZy = c.at<uint16_t>(L.dZ2y);
  float* dZfy = dZf;  // fp32 copies only feed the bias sums
  hipLaunchKernelGGL(ws_relu_bwd_kernel, dim3(grid_for(nx * kH, 256)), dim3(256), 0, st, dH2x,
                     c.at<uint16_t>(L.H2x), nx * kH, dZf, dZx);
  TM_LAUNCH_CHECK();
  s |= colsum(c, dZf, false, nullptr, nx, kH, grad + oB1, 0);
  hipLaunchKernelGGL(ws_relu_bwd_kernel, dim3(grid_for(ny * kH, 256)), dim3(256), 0, st, dH2y,
                     c.at<uint16_t>(L.H2y), ny * kH, dZfy, dZy);
  TM_LAUNCH_CHECK();
  s |= colsum(c, dZfy, false, nullptr, ny, kH, grad + oB1, 1);
  s |= wgrad(c, c.at<uint16_t>(L.H1x), kH, dZx, kH, nx, grad + oW1, false);
  s |= wgrad(c, c.at<uint16_t>(L.H1y), kH, dZy, kH, ny, grad + oW1, true);
  // dH1 (reuse dH2 buffers)
  s |= dgrad(c, dZx, kH, Wb + oW1, kH, nx, dH2x, false);
  s |= dgrad(c, dZy, kH, Wb + oW1, kH, ny, dH2y, false);
  // phi layer 1: H1 = relu(U W0 + b0)
  hipLaunchKernelGGL(ws_relu_bwd_kernel, dim3(grid_for(nx * kH, 256)), dim3(256), 0, st, dH2x,
                     c.at<uint16_t>(L.H1x), nx * kH, dZf, dZx);
  TM_LAUNCH_CHECK();
  s |= colsum(c, dZf, false, nullptr, nx, kH, grad + oB0, 0);
  hipLaunchKernelGGL(ws_relu_bwd_kernel, dim3(grid_for(ny * kH, 256)), dim3(256), 0, st, dH2y,
                     c.at<uint16_t>(L.H1y), ny * kH, dZfy, dZy);
  TM_LAUNCH_CHECK();
  s |= colsum(c, dZfy, false, nullptr, ny, kH, grad + oB0, 1);
  s |= wgrad(c, c.at<uint16_t>(L.Ux), kIn, dZx, kH, nx, grad + oW0, false);
  s |= wgrad(c, c.at<uint16_t>(L.Uy), kIn, dZy, kH, ny, grad + oW0, true);
  return s ? TM_ERR_HIP : TM_OK;
}

}  // namespace tmh

extern "C" int64_t tm_warmstart_param_count(void) { return tmh::kNParams; }

extern "C" size_t tm_warmstart_workspace_bytes(int npairs, const int32_t* n, const int32_t* m) {
  int64_t nx = 0, ny = 0;
  for (int p = 0; p < npairs; ++p) {
    nx += n[p];
    ny += m[p];
  }
  tmh::WsLayout L;
  L.build(nx, ny, npairs);
  return L.bytes;
}

// Predicted f (cost units) for every x point: f = net(x, y, eps) * s2.
extern "C" int tm_warmstart_predict(const float* params, const float* x, const float* y,
                                    int npairs, const int32_t* n, const int32_t* m,
                                    const float* eps, float* f, void* workspace,
                                    size_t workspace_bytes, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(params && x && y && f && n && m && eps, "tm_warmstart_predict: NULL argument");
  TM_CHECK_ARG(npairs >= 1, "tm_warmstart_predict: npairs must be >= 1");
  WsCtx c{};
  int s = ws_setup(c, x, y, npairs, n, m, eps, workspace, workspace_bytes, (hipStream_t)stream);
  if (s != TM_OK) return s;
  if ((s = ws_forward(c, params, x, y)) != TM_OK) return s;
  hipLaunchKernelGGL(ws_scale_kernel, dim3(grid_for(c.nx, 256)), dim3(256), 0, c.st,
                     c.at<float>(c.L.out), c.at<int32_t>(c.L.gx), c.at<float>(c.L.stats), c.nx, f);
  TM_LAUNCH_CHECK();
  return TM_OK;
}

// Loss and parameter gradient for targets f_target (cost units) of every x
// point; grad has tm_warmstart_param_count() floats, loss is one float.
extern "C" int tm_warmstart_grad(const float* params, const float* x, const float* y, int npairs,
                                 const int32_t* n, const int32_t* m, const float* eps,
                                 const float* f_target, float loss_alpha, float* grad, float* loss,
                                 void* workspace, size_t workspace_bytes, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(params && x && y && f_target && grad && loss, "tm_warmstart_grad: NULL argument");
  WsCtx c{};
  int s = ws_setup(c, x, y, npairs, n, m, eps, workspace, workspace_bytes, (hipStream_t)stream);
  if (s != TM_OK) return s;
  if ((s = ws_forward(c, params, x, y)) != TM_OK) return s;
  return ws_backward(c, params, f_target, loss_alpha, grad, loss);
}

// Adam step with bias correction and decoupled weight decay.
extern "C" int tm_warmstart_adam(float* params, const float* grad, float* m1, float* m2,
                                 int64_t n, int step, float lr, float beta1, float beta2,
                                 float eps, float weight_decay, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(params && grad && m1 && m2 && n > 0 && step >= 1, "tm_warmstart_adam: bad args");
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(ws_adam_kernel, dim3(grid_for(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     params, grad, m1, m2, n, lr, beta1, beta2, eps, c1, c2, weight_decay);
  TM_LAUNCH_CHECK();
  return TM_OK;
}
