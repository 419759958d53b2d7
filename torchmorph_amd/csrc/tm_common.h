// Shared host/device helpers for the torchmorph HIP library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/torchmorph_hip.h"

namespace tmh {

// Last error message, per host thread.
void set_error(const char* fmt, ...);
const char* get_error();

constexpr int kMaxNdim = 8;

// Small by-value shape descriptor passed to kernels (kernel arguments live in
// the constant/kernarg segment, so no H2D copy is needed per call).
struct Shape {
  int ndim;
  int64_t size[kMaxNdim];
  int64_t stride[kMaxNdim];  // element stride of each spatial axis
  int64_t numel;             // prod(size)
};

inline Shape make_shape(int ndim, const int64_t* s) {
  Shape sh{};
  sh.ndim = ndim;
  int64_t st = 1;
  for (int d = ndim - 1; d >= 0; --d) {
    sh.size[d] = s[d];
    sh.stride[d] = st;
    st *= s[d];
  }
  sh.numel = st;
  return sh;
}

#define TM_CHECK_ARG(cond, ...)        \
  do {                                 \
    if (!(cond)) {                     \
      ::tmh::set_error(__VA_ARGS__);    \
      return TM_ERR_ARG;               \
    }                                  \
  } while (0)

#define TM_CHECK_HIP(expr)                                                   \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) {                                                  \
      ::tmh::set_error("%s failed: %s", #expr, hipGetErrorString(_e));        \
      return TM_ERR_HIP;                                                     \
    }                                                                        \
  } while (0)

#define TM_LAUNCH_CHECK()                                                    \
  do {                                                                       \
    hipError_t _e = hipGetLastError();                                       \
    if (_e != hipSuccess) {                                                  \
      ::tmh::set_error("kernel launch failed: %s", hipGetErrorString(_e));    \
      return TM_ERR_HIP;                                                     \
    }                                                                        \
  } while (0)

inline size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Pinned host staging for descriptor uploads, one per host thread and call
// site: a buffer is reused only after its previous copy has completed.
struct PinnedStage {
  void* host = nullptr;
  size_t cap = 0;
  hipEvent_t ev = nullptr;
  bool pending = false;
};

// Waits for the stage's previous copy and grows the buffer to `need` bytes.
inline int stage_acquire(PinnedStage& s, size_t need) {
  if (s.pending) {
    TM_CHECK_HIP(hipEventSynchronize(s.ev));
    s.pending = false;
  }
  if (s.cap < need) {
    if (s.host) TM_CHECK_HIP(hipHostFree(s.host));
    s.host = nullptr;
    s.cap = 0;
    TM_CHECK_HIP(hipHostMalloc(&s.host, need, hipHostMallocDefault));
    s.cap = need;
  }
  if (!s.ev) TM_CHECK_HIP(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
  return TM_OK;
}

// Copies the first `n` staged bytes to `dst` on `st`; the stage stays busy
// until that copy completes.
inline int stage_upload(PinnedStage& s, void* dst, size_t n, hipStream_t st) {
  TM_CHECK_HIP(hipMemcpyAsync(dst, s.host, n, hipMemcpyHostToDevice, st));
  TM_CHECK_HIP(hipEventRecord(s.ev, st));
  s.pending = true;
  return TM_OK;
}

inline unsigned grid_for(int64_t n, int block) {
  int64_t g = (n + block - 1) / block;
  return (unsigned)(g < 1 ? 1 : g);
}

// MI355X dispatches workgroup i of a launch to XCD i % 8, and each XCD has
// its own L2.  xcd_swizzle maps the launch-order id to a logical block index
// so that every XCD receives one contiguous range of logical blocks: blocks
// that share cache lines (neighbouring tiles of one image) then share an L2
// instead of being fetched once per XCD.  A bijection of [0, G).
__device__ __forceinline__ int xcd_swizzle(int id, int G) {
  const int x = id & 7, s = id >> 3, q = G >> 3, r = G & 7;
  return x * q + (x < r ? x : r) + s;
}

// Reads a mask element: nonzero means foreground (reference feeds
// input.float(); any nonzero value is foreground).
template <typename T>
__device__ __forceinline__ bool is_fg(const T* p, int64_t i) {
  return p[i] != T(0);
}

}  // namespace tmh
