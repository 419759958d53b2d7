// Shared declarations of the BF16 MFMA GEMM (tm_mlp.hip).
#pragma once
#include "tm_common.h"

namespace tmh {

struct GemmArgs {
  const uint16_t* A;
  int64_t sam, sak;  // A(m, k) = A[m * sam + k * sak]
  const uint16_t* B;
  int64_t sbk, sbn;  // B(k, n) = B[k * sbk + n * sbn]
  void* C;
  int64_t ldc;
  int M, N, K;
  int epi;                  // 0: fp32 (+ beta C), 1: fp32 + bias, 2: bf16 relu(+bias +rowadd)
  const float* bias;
  const float* rowadd;      // [groups][N] fp32 or NULL
  const int32_t* rowgroup;  // group of each row or NULL
  float beta;
  int ksplit;
};

// Launches C = A.B; long reductions with few output tiles are split over
// workgroups (partials in splitk_ws, summed in split order) when a workspace
// is given.
int launch_gemm(const GemmArgs& g, float* splitk_ws, size_t splitk_ws_elems, hipStream_t st);

}  // namespace tmh
