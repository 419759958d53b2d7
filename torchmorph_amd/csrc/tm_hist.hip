// Histogram Sinkhorn for gfx950: batches of (nb, d) histograms sharing one (d, d)
// cost (reference torchmorph/optimal_transport.py:71-150, csrc/ot_kernel.cu:83-283).
//
// One half-iteration, for every batch row b and output index i:
//   log form:     out[b,i] = num[b,i] - LSE_j(scale[b,j] - M[i,j] / eps)    (M = cost)
//   scaling form: out[b,i] = num[b,i] / (sum_j M[i,j] scale[b,j] + 1e-12)  (M = K)
// The u-update passes (a, M, v), the v-update (b, M^T, u).
//
// Layout: one wave owns one output row i and a batch tile of NB rows.  Lanes
// stride j with V-wide vector loads, so a wave reads row i of M once, coalesced
// (1 KB per load instruction for float4), and reuses every element for its NB
// batch rows from registers; the NB scale rows are read the same way and shared
// by the workgroup's four waves through L1.  No LDS and no barriers: each wave
// folds its 64 lane-partial states with cross-lane butterflies (fixed order, so
// results are deterministic).  The log-sum-exp is online over chunks of V terms:
// chunk maximum, one rescale of the running sum, V exponentials.
//
// Fused epilogues (nothing of size nb x d x d is ever formed):
//  * value: the last v-update also accumulates S2 = sum_j e^(z_j - m) C[i,j]
//    (scaling: sum_j K[i,j] u_j C[i,j]); with the final v the row's share of
//    <P, C> is b_i S2 / S1 (log form) or v_i S2 (scaling form);
//  * threshold (early stop as the reference's torch iterations, checked at
//    iterations i % 10 == 0): the u-update writes |u_new - u_old|, a check
//    kernel sums it per batch row in a fixed order and sets a stop word that
//    every later update reads first.  The loop stays capturable: no host sync
//    inside a chunk.
//  * finish: centred potentials f = eps log u - mean, g likewise, and the
//    per-row value, one wave per batch row.
#include <cmath>
#include <cstdlib>

#include "tm_common.h"

namespace tmh {
namespace {

template <typename T, int V>
struct Vec;
template <>
struct Vec<float, 4> { using type = float4; };
template <>
struct Vec<float, 1> { using type = float; };
template <>
struct Vec<double, 2> { using type = double2; };
template <>
struct Vec<double, 1> { using type = double; };

template <typename T, int V>
__device__ __forceinline__ void load_vec(const T* __restrict__ p, T (&o)[V]) {
  union {
    typename Vec<T, V>::type v;
    T s[V];
  } u;
  u.v = *reinterpret_cast<const typename Vec<T, V>::type*>(p);
#pragma unroll
  for (int k = 0; k < V; ++k) o[k] = u.s[k];
}

// float follows the reference build's fast-math exponent and logarithm;
// double uses the correctly rounded library functions.
__device__ __forceinline__ float hexp(float x) { return __expf(x); }
__device__ __forceinline__ double hexp(double x) { return exp(x); }
__device__ __forceinline__ float hlog(float x) { return __logf(x); }
__device__ __forceinline__ double hlog(double x) { return log(x); }

template <typename T>
struct HistArgs {
  const T* num;    // (nb, d): log a / log b (log form) or a / b
  const T* M;      // (d, d): cost (log form) or K (scaling form), row i = output i
  const T* C;      // (d, d): cost rows for the value epilogue (scaling form)
  const T* scale;  // (nb, d): opposite (log-)scaling
  T* out;          // (nb, d): updated in place
  T* delta;        // (nb, d): |new - old| when non-null
  T* valpart;      // (nb, d): value shares when the epilogue is on
  const int* stop; // stop word: nonzero = converged, the kernel does nothing
  int64_t nb;
  int d;
  T inv_eps;
};

// Merges the lane state (m, s, s2) with the one `off` lanes away.
template <typename T, bool LOG, bool VAL>
__device__ __forceinline__ void merge_lanes(T& m, T& s, T& s2, int off) {
  const T om = __shfl_xor(m, off, 64);
  const T os = __shfl_xor(s, off, 64);
  const T os2 = VAL ? __shfl_xor(s2, off, 64) : T(0);
  if (LOG) {
    const T hi = m > om ? m : om;
    if (hi == -INFINITY) return;  // both empty: s == s2 == 0 already
    const T ra = hexp(m - hi), rb = hexp(om - hi);
    s = s * ra + os * rb;
    if (VAL) s2 = s2 * ra + os2 * rb;
    m = hi;
  } else {
    s += os;
    if (VAL) s2 += os2;
  }
}

template <typename T, int NB, int V, bool LOG, bool VAL>
__global__ __launch_bounds__(256) void hist_update_kernel(HistArgs<T> a) {
  if (a.stop && *a.stop) return;
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= a.d) return;  // whole wave: no barriers in this kernel
  const int64_t b0 = (int64_t)blockIdx.y * NB;
  const int d = a.d;
  const T* __restrict__ mrow = a.M + (int64_t)i * d;
  const T* __restrict__ crow = (VAL && !LOG) ? a.C + (int64_t)i * d : nullptr;
  T m[NB], s[NB], s2[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    m[b] = LOG ? T(-INFINITY) : T(0);
    s[b] = T(0);
    s2[b] = T(0);
  }
  for (int j = lane * V; j < d; j += 64 * V) {
    T mv[V], cv[V];
    load_vec<T, V>(mrow + j, mv);
    if (VAL && !LOG) load_vec<T, V>(crow + j, cv);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      if (b0 + b >= a.nb) break;  // uniform over the wave
      T sv[V];
      load_vec<T, V>(a.scale + (b0 + b) * d + j, sv);
      if (LOG) {
        T z[V];
        T cm = T(-INFINITY);
#pragma unroll
        for (int k = 0; k < V; ++k) {
          z[k] = sv[k] - mv[k] * a.inv_eps;
          cm = z[k] > cm ? z[k] : cm;
        }
        const T hi = m[b] > cm ? m[b] : cm;
        if (hi != -INFINITY) {
          const T r = hexp(m[b] - hi);
          s[b] *= r;
          if (VAL) s2[b] *= r;
#pragma unroll
          for (int k = 0; k < V; ++k) {
            const T e = hexp(z[k] - hi);
            s[b] += e;
            if (VAL) s2[b] += e * mv[k];
          }
          m[b] = hi;
        }
      } else {
#pragma unroll
        for (int k = 0; k < V; ++k) {
          const T t = mv[k] * sv[k];
          s[b] += t;
          if (VAL) s2[b] += t * cv[k];
        }
      }
    }
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) merge_lanes<T, LOG, VAL>(m[b], s[b], s2[b], off);
  }
  if (lane != 0) return;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (b0 + b >= a.nb) break;
    const int64_t idx = (b0 + b) * d + i;
    const T num = a.num[idx];
    T nv;
    if (LOG)  // an all -inf row (zero opposite marginal) pins the potential
      nv = m[b] == -INFINITY ? T(-INFINITY) : num - m[b] - hlog(s[b]);
    else
      nv = num / (s[b] + T(1e-12));
    if (a.delta) a.delta[idx] = fabs(nv - a.out[idx]);
    a.out[idx] = nv;
    if (VAL) {
      if (LOG)
        a.valpart[idx] = m[b] == -INFINITY ? T(0) : hexp(num) * (s2[b] / s[b]);
      else
        a.valpart[idx] = nv * s2[b];
    }
  }
}

// Sums |u_new - u_old| per batch row (fixed order) and takes the maximum over
// rows; sets *stop = 10 * k + 1 (iterations run) on the k-th check when the
// maximum is <= threshold.  A NaN anywhere never stops (the reference compares
// a NaN maximum, which is false).  One workgroup of 1024 lanes.
template <typename T>
__global__ __launch_bounds__(1024) void hist_check_kernel(const T* __restrict__ delta,
                                                          int64_t nb, int d, T threshold,
                                                          int* stop, int* nchecks) {
  if (*stop) return;
  __shared__ T wmax[16];
  __shared__ int wnan[16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  T mx = T(0);
  int nan = 0;
  for (int64_t b = wave; b < nb; b += 16) {
    T acc = T(0);
    for (int j = lane; j < d; j += 64) acc += delta[b * d + j];
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (acc != acc) nan = 1;
    else mx = acc > mx ? acc : mx;
  }
  if (lane == 0) {
    wmax[wave] = mx;
    wnan[wave] = nan;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 16; ++w) {
      mx = wmax[w] > mx ? wmax[w] : mx;
      nan |= wnan[w];
    }
    const int k = *nchecks;
    *nchecks = k + 1;
    if (!nan && mx <= threshold) *stop = 10 * k + 1;
  }
}

__global__ void hist_reset_kernel(int* words) {
  if (threadIdx.x < 4) words[threadIdx.x] = 0;
}

// Centred potentials and the value of every batch row; one wave per row.
template <typename T>
__global__ __launch_bounds__(256) void hist_finish_kernel(const T* __restrict__ u,
                                                          const T* __restrict__ v,
                                                          const T* __restrict__ valpart,
                                                          T* f, T* g, T* dist, int64_t nb,
                                                          int d, int log_space, T eps) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= nb) return;
  const T* ur = u + b * d;
  const T* vr = v + b * d;
  T su = T(0), sv = T(0), sd = T(0);
  for (int j = lane; j < d; j += 64) {
    su += eps * (log_space ? ur[j] : hlog(ur[j]));
    sv += eps * (log_space ? vr[j] : hlog(vr[j]));
    if (valpart) sd += valpart[b * d + j];
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    su += __shfl_xor(su, off, 64);
    sv += __shfl_xor(sv, off, 64);
    sd += __shfl_xor(sd, off, 64);
  }
  const T mu = su / T(d), mv = sv / T(d);
  if (f)
    for (int j = lane; j < d; j += 64) f[b * d + j] = eps * (log_space ? ur[j] : hlog(ur[j])) - mu;
  if (g)
    for (int j = lane; j < d; j += 64) g[b * d + j] = eps * (log_space ? vr[j] : hlog(vr[j])) - mv;
  if (dist && lane == 0) dist[b] = sd;
}

// plan[b, i, j] = exp(log u[b,i] - C[i,j] / eps + log v[b,j]), the reference's
// reconstruction (optimal_transport.py:228-233), one lane per entry, coalesced
// along j.
template <typename T>
__global__ __launch_bounds__(256) void hist_plan_kernel(const T* __restrict__ u,
                                                        const T* __restrict__ v,
                                                        const T* __restrict__ C, T* plan,
                                                        int64_t nb, int d, int log_space,
                                                        T eps) {
  const int64_t dd = (int64_t)d * d;
  const int64_t total = nb * dd;
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = t / dd;
    const int64_t r = t - b * dd;
    const int i = (int)(r / d), j = (int)(r - (int64_t)i * d);
    const T lu = log_space ? u[b * d + i] : hlog(u[b * d + i]);
    const T lv = log_space ? v[b * d + j] : hlog(v[b * d + j]);
    plan[t] = hexp(lu - C[r] / eps + lv);
  }
}

// ------------------------------------------------------------------ host ---

template <typename T, int NB, int V>
int launch_update(bool log_space, bool value, const HistArgs<T>& a, hipStream_t st) {
  const dim3 grid((unsigned)((a.d + 3) / 4), (unsigned)((a.nb + NB - 1) / NB)), blk(256);
  if (log_space) {
    if (value) hipLaunchKernelGGL((hist_update_kernel<T, NB, V, true, true>), grid, blk, 0, st, a);
    else hipLaunchKernelGGL((hist_update_kernel<T, NB, V, true, false>), grid, blk, 0, st, a);
  } else {
    if (value) hipLaunchKernelGGL((hist_update_kernel<T, NB, V, false, true>), grid, blk, 0, st, a);
    else hipLaunchKernelGGL((hist_update_kernel<T, NB, V, false, false>), grid, blk, 0, st, a);
  }
  TM_LAUNCH_CHECK();
  return TM_OK;
}

template <typename T, int V>
int dispatch_nb(bool log_space, bool value, const HistArgs<T>& a, hipStream_t st) {
  if (a.nb <= 1) return launch_update<T, 1, V>(log_space, value, a, st);
  if (a.nb <= 2) return launch_update<T, 2, V>(log_space, value, a, st);
  if (a.nb <= 4) return launch_update<T, 4, V>(log_space, value, a, st);
  return launch_update<T, 8, V>(log_space, value, a, st);
}

// Vector loads need every row start on a 16-byte boundary: d a multiple of the
// vector width and 16-byte aligned base pointers (torch views may be offset).
template <typename T>
int update(bool log_space, bool value, const HistArgs<T>& a, hipStream_t st) {
  constexpr int kV = sizeof(T) == 4 ? 4 : 2;
  auto al = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  const bool vec = a.d % kV == 0 && al(a.M) && al(a.scale) && (!value || !a.C || al(a.C));
  if (vec) return dispatch_nb<T, kV>(log_space, value, a, st);
  return dispatch_nb<T, 1>(log_space, value, a, st);
}

constexpr int64_t kMaxGridY = 65535;

// Runs `launch(k)` (k iterations enqueued on *lst) n_iter times in total; for
// n_iter >= 100 a chunk of `chunk` iterations is captured once as a hipGraph
// and replayed (reference optimal_transport.py:34-68), and `poll()` (when
// given) is asked after replays whether to stop.  A capture failure falls back
// to plain launches; extra iterations are never run.
template <typename F, typename P>
int run_chunked(F&& launch, P&& poll, int n_iter, int chunk, hipStream_t st, hipStream_t* lst,
                int* fallback_out) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  const bool capturing = hipStreamIsCapturing(st, &cs) == hipSuccess &&
                         cs != hipStreamCaptureStatusNone;
  const char* force = getenv("TORCHMORPH_AMD_FORCE_GRAPH_FAILURE");
  if (n_iter < 100 || capturing) return launch(0, n_iter);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  bool ok = hipStreamCreateWithFlags(&cap, hipStreamNonBlocking) == hipSuccess;
  if (ok) ok = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal) == hipSuccess;
  int status = TM_OK;
  if (ok) {
    *lst = cap;
    status = launch(0, chunk);
    *lst = st;
    ok = hipStreamEndCapture(cap, &graph) == hipSuccess && status == TM_OK;
  }
  if (ok && force && force[0] == '1') ok = false;
  if (ok) ok = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess;
  if (!ok) {
    if (graph) (void)hipGraphDestroy(graph);
    if (cap) (void)hipStreamDestroy(cap);
    (void)hipGetLastError();
    if (fallback_out) *fallback_out = 1;
    return launch(0, n_iter);
  }
  const int reps = n_iter / chunk;
  int done = 0;
  for (int r = 0; r < reps && status == TM_OK; ++r) {
    if (hipGraphLaunch(exec, st) != hipSuccess) status = TM_ERR_HIP;
    if (status == TM_OK && poll(r + 1) ) {
      done = 1;
      break;
    }
  }
  if (status == TM_OK && !done && n_iter % chunk) status = launch(reps * chunk, n_iter % chunk);
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  (void)hipStreamDestroy(cap);
  return status;
}

template <typename T>
int solve(int log_space, const T* num_a, const T* num_b, const T* M, const T* M_T, const T* C,
          const T* C_T, T* u, T* v, int64_t nb, int d, int n_iter, double epsilon,
          double threshold, T* f, T* g, T* dist, int* iters_run, int* used_fallback,
          char* ws, hipStream_t st) {
  int* words = (int*)ws;  // [0] stop, [1] number of checks
  T* delta = (T*)(ws + 256);
  T* valpart = delta + (threshold > 0 ? nb * d : 0);
  const bool value = dist != nullptr;
  const bool thr = threshold > 0;
  hipStream_t lst = st;
  if (used_fallback) *used_fallback = 0;
  hipLaunchKernelGGL(hist_reset_kernel, dim3(1), dim3(64), 0, st, words);
  TM_LAUNCH_CHECK();
  HistArgs<T> ua{num_a, M, C, v, u, nullptr, nullptr, thr ? words : nullptr, nb, d,
                 (T)(1.0 / epsilon)};
  HistArgs<T> va{num_b, M_T, C_T, u, v, nullptr, nullptr, thr ? words : nullptr, nb, d,
                 (T)(1.0 / epsilon)};
  // value epilogue on the last v-update of a fixed-count run
  const bool fuse_last = value && !thr && n_iter >= 1;
  const int loop_iters = fuse_last ? n_iter - 1 : n_iter;
  auto launch = [&](int first, int k) -> int {
    for (int it = first; it < first + k; ++it) {
      const bool check = thr && it % 10 == 0;
      HistArgs<T> a = ua;
      a.delta = check ? delta : nullptr;
      int s = update<T>(log_space, false, a, lst);
      if (s != TM_OK) return s;
      s = update<T>(log_space, false, va, lst);
      if (s != TM_OK) return s;
      if (check) {
        hipLaunchKernelGGL(hist_check_kernel<T>, dim3(1), dim3(1024), 0, lst, delta, nb, d,
                           (T)threshold, words, words + 1);
        TM_LAUNCH_CHECK();
      }
    }
    return TM_OK;
  };
  int stop_host = 0;
  // Host polls of the stop word (threshold runs only): after graph replays
  // 1, 2, 4, 8, then every 8 replays of 10 iterations.
  auto poll = [&](int replays) -> bool {
    if (!thr) return false;
    if (replays > 8 ? replays % 8 : (replays & (replays - 1))) return false;
    if (hipMemcpyAsync(&stop_host, words, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess)
      return false;
    if (hipStreamSynchronize(st) != hipSuccess) return false;
    return stop_host != 0;
  };
  int status = run_chunked(launch, poll, loop_iters, thr ? 10 : 25, st, &lst, used_fallback);
  if (status != TM_OK) return status;
  if (value) {
    // the last v-update with the value epilogue: for a fixed count it is the
    // n_iter-th iteration's; after an early stop it recomputes v from the
    // final u, which reproduces the stored v bit for bit
    if (fuse_last) {
      int s = update<T>(log_space, false, ua, st);
      if (s != TM_OK) return s;
    }
    HistArgs<T> a = va;
    a.valpart = valpart;
    a.stop = nullptr;
    int s = update<T>(log_space, true, a, st);
    if (s != TM_OK) return s;
  }
  if (f || g || dist) {
    hipLaunchKernelGGL(hist_finish_kernel<T>, dim3((unsigned)((nb + 3) / 4)), dim3(256), 0, st,
                       (const T*)u, (const T*)v, value ? (const T*)valpart : (const T*)nullptr,
                       f, g, dist, nb, d, log_space, (T)epsilon);
    TM_LAUNCH_CHECK();
  }
  if (iters_run) {
    *iters_run = n_iter;
    if (thr) {
      TM_CHECK_HIP(hipMemcpyAsync(&stop_host, words, sizeof(int), hipMemcpyDeviceToHost, st));
      TM_CHECK_HIP(hipStreamSynchronize(st));
      if (stop_host) *iters_run = stop_host;
    }
  }
  return TM_OK;
}

}  // namespace
}  // namespace tmh

// ------------------------------------------------------------------ C ABI ---

extern "C" int tm_sinkhorn_hist(int dtype, int log_space, const void* num_a, const void* num_b,
                                const void* M, const void* M_T, void* u, void* v, int64_t nb,
                                int d, int n_iter, double epsilon, int* used_fallback,
                                void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(num_a && num_b && M && M_T && u && v, "tm_sinkhorn_hist: NULL argument");
  TM_CHECK_ARG(nb >= 1 && d >= 1 && n_iter >= 0, "tm_sinkhorn_hist: bad sizes");
  TM_CHECK_ARG(nb <= kMaxGridY, "tm_sinkhorn_hist: at most %lld batch rows",
               (long long)kMaxGridY);
  TM_CHECK_ARG(dtype == TM_F32 || dtype == TM_F64, "tm_sinkhorn_hist: dtype must be f32 or f64");
  TM_CHECK_ARG(epsilon > 0, "tm_sinkhorn_hist: epsilon must be positive, got %g", epsilon);
  hipStream_t st = (hipStream_t)stream;
  hipStream_t lst = st;
  if (used_fallback) *used_fallback = 0;
  auto body = [&](auto zero) -> int {
    using T = decltype(zero);
    HistArgs<T> ua{(const T*)num_a, (const T*)M, nullptr, (const T*)v, (T*)u, nullptr, nullptr,
                   nullptr, nb, d, (T)(1.0 / epsilon)};
    HistArgs<T> va{(const T*)num_b, (const T*)M_T, nullptr, (const T*)u, (T*)v, nullptr,
                   nullptr, nullptr, nb, d, (T)(1.0 / epsilon)};
    auto launch = [&](int, int k) -> int {
      for (int it = 0; it < k; ++it) {
        int s = update<T>(log_space, false, ua, lst);
        if (s != TM_OK) return s;
        s = update<T>(log_space, false, va, lst);
        if (s != TM_OK) return s;
      }
      return TM_OK;
    };
    auto never = [](int) { return false; };
    return run_chunked(launch, never, n_iter, 25, st, &lst, used_fallback);
  };
  return dtype == TM_F32 ? body(0.0f) : body(0.0);
}

extern "C" size_t tm_sinkhorn_hist_workspace_bytes(int dtype, int64_t nb, int d,
                                                   double threshold, int want_value) {
  const size_t el = dtype == TM_F64 ? 8 : 4;
  size_t b = 256;
  if (threshold > 0) b += tmh::align256(el * (size_t)nb * d);
  if (want_value) b += tmh::align256(el * (size_t)nb * d);
  return b;
}

extern "C" int tm_sinkhorn_hist_solve(int dtype, int log_space, const void* num_a,
                                      const void* num_b, const void* M, const void* M_T,
                                      const void* C, const void* C_T, void* u, void* v,
                                      int64_t nb, int d, int max_iter, double epsilon,
                                      double threshold, void* f, void* g, void* dist,
                                      int* iters_run, int* used_fallback, void* workspace,
                                      size_t workspace_bytes, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(num_a && num_b && M && M_T && u && v, "tm_sinkhorn_hist_solve: NULL argument");
  TM_CHECK_ARG(nb >= 1 && d >= 1 && max_iter >= 0, "tm_sinkhorn_hist_solve: bad sizes");
  TM_CHECK_ARG(nb <= kMaxGridY, "tm_sinkhorn_hist_solve: at most %lld batch rows",
               (long long)kMaxGridY);
  TM_CHECK_ARG(dtype == TM_F32 || dtype == TM_F64,
               "tm_sinkhorn_hist_solve: dtype must be f32 or f64");
  TM_CHECK_ARG(epsilon > 0, "tm_sinkhorn_hist_solve: epsilon must be positive, got %g",
               epsilon);
  TM_CHECK_ARG(threshold >= 0, "tm_sinkhorn_hist_solve: threshold must be non-negative");
  TM_CHECK_ARG(!dist || (C && C_T), "tm_sinkhorn_hist_solve: the value needs C and C_T");
  const size_t need = tm_sinkhorn_hist_workspace_bytes(dtype, nb, d, threshold, dist != nullptr);
  if (!workspace || workspace_bytes < need) {
    set_error("tm_sinkhorn_hist_solve: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, need);
    return TM_ERR_WORKSPACE;
  }
  hipStream_t st = (hipStream_t)stream;
  if (dtype == TM_F32)
    return solve<float>(log_space, (const float*)num_a, (const float*)num_b, (const float*)M,
                        (const float*)M_T, (const float*)C, (const float*)C_T, (float*)u,
                        (float*)v, nb, d, max_iter, epsilon, threshold, (float*)f, (float*)g,
                        (float*)dist, iters_run, used_fallback, (char*)workspace, st);
  return solve<double>(log_space, (const double*)num_a, (const double*)num_b, (const double*)M,
                       (const double*)M_T, (const double*)C, (const double*)C_T, (double*)u,
                       (double*)v, nb, d, max_iter, epsilon, threshold, (double*)f, (double*)g,
                       (double*)dist, iters_run, used_fallback, (char*)workspace, st);
}

extern "C" int tm_sinkhorn_hist_plan(int dtype, int log_space, const void* u, const void* v,
                                     const void* C, int64_t nb, int d, double epsilon,
                                     void* plan, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(u && v && C && plan, "tm_sinkhorn_hist_plan: NULL argument");
  TM_CHECK_ARG(nb >= 1 && d >= 1, "tm_sinkhorn_hist_plan: bad sizes");
  TM_CHECK_ARG(dtype == TM_F32 || dtype == TM_F64,
               "tm_sinkhorn_hist_plan: dtype must be f32 or f64");
  hipStream_t st = (hipStream_t)stream;
  const int64_t total = nb * (int64_t)d * d;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (dtype == TM_F32)
    hipLaunchKernelGGL(hist_plan_kernel<float>, dim3((unsigned)blocks), dim3(256), 0, st,
                       (const float*)u, (const float*)v, (const float*)C, (float*)plan, nb, d,
                       log_space, (float)epsilon);
  else
    hipLaunchKernelGGL(hist_plan_kernel<double>, dim3((unsigned)blocks), dim3(256), 0, st,
                       (const double*)u, (const double*)v, (const double*)C, (double*)plan, nb,
                       d, log_space, epsilon);
  TM_LAUNCH_CHECK();
  return TM_OK;
}
