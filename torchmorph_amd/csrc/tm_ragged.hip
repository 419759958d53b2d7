// Ragged batches of distance transforms in one persistent launch.
//
// The reference takes dense (B, C, spatial...) tensors only, so a list of
// masks of different shapes costs one call (and its passes) per shape
// (reference distance_transform.py:10-102, edt_kernel.cu:904-958,
// cdt_kernel.cu:269-463).  Here a list of jobs -- (item, op) pairs, each with
// its own shape, sampling and output offsets -- becomes one device work-list:
//
//  * every separable pass of every job is cut into tasks of up to 256 x m
//    lines (one lane per line, m lines per lane, about 64k voxels per task);
//    a chamfer job with indices is one task per raster pass instead (whole
//    item, tm_cdt_dev.h);
//  * tasks are ordered level-major (all first passes, then all second passes,
//    ...), largest jobs first within a level; a task of level l of a job
//    depends only on that job's tasks of levels < l, which come earlier in the
//    list;
//  * a persistent grid of 256-thread workgroups dequeues tasks in list order
//    with one atomic counter.  Before running a task a workgroup waits until
//    the job's completion counter reaches the number of tasks it depends on
//    (acquire load, bounded spin); after it, every thread fences its stores
//    and one thread increments the counter (release).  Because dependencies
//    always point to tasks dequeued earlier, by workgroups that are already
//    running, the wait always ends; the spin bound only turns a broken
//    invariant into an error flag (read back by tm_ragged_dt_check) instead of
//    a hang, and a raised flag makes every workgroup leave at its next
//    dequeue or poll.
//
// schedule = 1 runs the same tasks with one launch per level instead (stream
// order provides the dependencies, no waiting) for comparison.
//
// The per-line code is the dense kernels' (tm_dt_lines.h), so every job's
// result is bit-identical to the dense entry point on that item alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "tm_cdt_dev.h"
#include "tm_dt_lines.h"

namespace tmh {

enum RgKind : int32_t {
  RG_EDT0 = 0,
  RG_EDT_STACK = 1,
  RG_CDT0 = 2,
  RG_CDT_TAXICAB = 3,
  RG_CDT_CHESSBOARD = 4,
  RG_CDT_RASTER = 5
};

constexpr int kRgLevels = kMaxNdim;     // levels 0..ndim-1
constexpr int kRgErr = kRgLevels;       // counter slot of the abort flag
constexpr int kRgJob0 = kRgLevels + 1;  // first job completion counter
constexpr int kRgThreads = 256;
constexpr int64_t kRgSpinLimit = (int64_t)1 << 22;

struct RgJob {
  Shape sh;
  Samp sp;
  int64_t in_off;   // elements of the input
  int64_t out_off;  // elements of the distances, -1 = none
  int64_t idx_off;  // elements of the indices (ndim planes of numel), -1 = none
  int64_t ws_off;   // bytes into the work area
  int32_t op;       // TM_RG_*
  int32_t raster;   // chamfer with indices: raster passes
};

struct RgTask {
  int32_t job;
  int32_t kind;
  int32_t d;       // axis of a line pass, pass index of a raster task
  int32_t m;       // lines per lane
  int64_t line0;   // first line of the task
  int64_t nlines;  // lines in the task
  int32_t need;    // job counter value required before starting
  int32_t pad;
};

struct RgArgs {
  float* dist;
  void* idx;
  int idx64;
  char* work;
  const RgJob* jobs;
  const RgTask* tasks;
  int* ctr;
  int qslot;   // dequeue counter used by this launch
  int lo, hi;  // task range of this launch
  int wait;    // 1: wait for dependencies (persistent schedule)
};

__global__ void rg_zero_kernel(int* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

template <typename TIn>
__device__ void rg_run_task(const TIn* __restrict__ in_all, const RgArgs& a, const RgTask& tk) {
  const RgJob& jb = a.jobs[tk.job];
  const Shape& sh = jb.sh;
  const TIn* in = in_all + jb.in_off;
  char* wk = a.work + jb.ws_off;
  const int64_t numel = sh.numel;
  const int nd = sh.ndim;

  if (tk.kind == RG_CDT_RASTER) {
    CdtOut o;
    o.dist = (a.dist && jb.out_off >= 0) ? a.dist + jb.out_off : nullptr;
    o.idx = (a.idx && jb.idx_off >= 0) ? (char*)a.idx + jb.idx_off * (a.idx64 ? 8 : 4) : nullptr;
    o.idx64 = a.idx64;
    o.plane = numel;
    int32_t* dt = (int32_t*)wk;
    cdt_item_pass(in, dt, dt + numel, sh, nd == 3 ? 1 : 0, jb.op == TM_RG_CDT_TAXICAB ? 1 : 0,
                  tk.d, o);
    return;
  }

  const int d = tk.d;
  const int64_t nl_axis = numel / sh.size[d];  // lines of this axis in the item
  const bool fin = d == nd - 1;
  for (int r = 0; r < tk.m; ++r) {
    const int64_t li = (int64_t)r * kRgThreads + threadIdx.x;
    if (li >= tk.nlines) break;
    const int64_t t = tk.line0 + li;
    switch (tk.kind) {
      case RG_EDT0:
      case RG_EDT_STACK: {
        EdtOut o;
        o.dist = (a.dist && jb.out_off >= 0) ? a.dist + jb.out_off : nullptr;
        o.idx = (a.idx && jb.idx_off >= 0) ? (char*)a.idx + jb.idx_off * (a.idx64 ? 8 : 4)
                                           : nullptr;
        o.idx64 = a.idx64;
        o.plane = numel;
        const int nbuf = nd > 1 ? nd - 1 : 1;
        int32_t* A = (int32_t*)wk;
        int32_t* B = A + (int64_t)nbuf * numel;
        int32_t* gpos = B + (int64_t)nbuf * numel;
        double* gh = (double*)(((uintptr_t)(gpos + numel) + 255) & ~(uintptr_t)255);
        if (tk.kind == RG_EDT0) {
          if (fin) edt_line_axis0<TIn, true>(in, A, sh, t, jb.sp, o);
          else edt_line_axis0<TIn, false>(in, A, sh, t, jb.sp, o);
        } else {
          const int32_t* src = (d & 1) ? A : B;
          int32_t* dst = (d & 1) ? B : A;
          if (fin)
            edt_line_stack<true>(src, dst, numel, sh, d, t, jb.sp, gpos + t, gh + t, nl_axis, o);
          else
            edt_line_stack<false>(src, dst, numel, sh, d, t, jb.sp, gpos + t, gh + t, nl_axis, o);
        }
        break;
      }
      case RG_CDT0:
      case RG_CDT_TAXICAB:
      case RG_CDT_CHESSBOARD: {
        int32_t* w = (int32_t*)(a.dist + jb.out_off);
        int64_t c[kMaxNdim], base, b;
        decode_line(sh, d, t, c, &base, &b);
        const int64_t step = sh.stride[d], L = sh.size[d];
        if (tk.kind == RG_CDT0) {
          if (fin) cdt_line_axis0<TIn, true>(in, w, base, step, L);
          else cdt_line_axis0<TIn, false>(in, w, base, step, L);
        } else if (tk.kind == RG_CDT_TAXICAB) {
          if (fin) cdt_line_taxicab<true>(w, base, step, L);
          else cdt_line_taxicab<false>(w, base, step, L);
        } else {
          uint16_t* S = (uint16_t*)wk;
          uint16_t* T = S + numel;
          int32_t* G = (int32_t*)(T + numel);
          if (fin) cdt_line_chessboard<true>(w, base, step, L, S + t, T + t, G + t, nl_axis);
          else cdt_line_chessboard<false>(w, base, step, L, S + t, T + t, G + t, nl_axis);
        }
        break;
      }
      default:
        break;
    }
  }
}

template <typename TIn>
__global__ __launch_bounds__(kRgThreads) void rg_kernel(const TIn* __restrict__ in, RgArgs a) {
  __shared__ int s_task;
  int* next = a.ctr + a.qslot;
  int* err = a.ctr + kRgErr;
  for (;;) {
    if (threadIdx.x == 0) {
      int k = a.lo + atomicAdd(next, 1);
      if (k >= a.hi || __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        k = -1;
      } else if (a.wait) {
        const RgTask& tk = a.tasks[k];
        int* done = a.ctr + kRgJob0 + tk.job;
        int64_t polls = 0;
        while (__hip_atomic_load(done, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < tk.need) {
          if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
            k = -1;
            break;
          }
          if (++polls > kRgSpinLimit) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            k = -1;
            break;
          }
          __builtin_amdgcn_s_sleep(4);
        }
      }
      s_task = k;
    }
    __syncthreads();
    const int k = __builtin_amdgcn_readfirstlane(s_task);
    if (k < 0) return;
    __threadfence();  // acquire for every thread: the job's earlier passes are visible
    const RgTask tk = a.tasks[k];
    rg_run_task(in, a, tk);
    __threadfence();  // release this thread's stores
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_fetch_add(a.ctr + kRgJob0 + tk.job, 1, __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_AGENT);
  }
}

// ------------------------------------------------------------------ host --

struct RgPlan {
  std::vector<RgJob> jobs;
  std::vector<RgTask> tasks;
  int level_lo[kRgLevels + 1];
  int nlevels;
  size_t ctr_bytes, jobs_off, tasks_off, work_off, total;
};

static size_t job_work_bytes(const RgJob& j) {
  const int64_t n = j.sh.numel;
  const int nd = j.sh.ndim;
  if (j.op == TM_RG_EDT) {
    // feature planes A and B (ndim-1 each, one for 1-D), stack positions, stack values
    const int nbuf = nd > 1 ? nd - 1 : 1;
    return align256((size_t)n * 4 * (2 * nbuf + 1)) + (size_t)n * 8;
  }
  if (j.raster) return (size_t)n * 8;
  return j.op == TM_RG_CDT_CHESSBOARD ? (size_t)n * 8 : 0;
}

static int build_plan(const tm_ragged_job* in, int64_t njobs, bool want_dist, bool want_idx,
                      RgPlan& p) {
  TM_CHECK_ARG(njobs >= 1 && njobs <= (1 << 24), "tm_ragged_dt: njobs out of range");
  p.jobs.resize((size_t)njobs);
  size_t work = 0;
  for (int64_t i = 0; i < njobs; ++i) {
    const tm_ragged_job& q = in[i];
    TM_CHECK_ARG(q.op == TM_RG_EDT || q.op == TM_RG_CDT_CHESSBOARD || q.op == TM_RG_CDT_TAXICAB,
                 "tm_ragged_dt: job %lld: unknown op %d", (long long)i, q.op);
    TM_CHECK_ARG(q.ndim >= 1 && q.ndim <= kMaxNdim, "tm_ragged_dt: job %lld: ndim %d out of 1..8",
                 (long long)i, q.ndim);
    int64_t n = 1;
    for (int j = 0; j < q.ndim; ++j) {
      TM_CHECK_ARG(q.shape[j] >= 1 && q.shape[j] <= 65535,
                   "tm_ragged_dt: job %lld: axis %d size out of range 1..65535", (long long)i, j);
      n *= q.shape[j];
    }
    TM_CHECK_ARG(n < ((int64_t)1 << 31), "tm_ragged_dt: job %lld: more than 2^31 voxels",
                 (long long)i);
    TM_CHECK_ARG(q.in_offset >= 0, "tm_ragged_dt: job %lld: negative input offset", (long long)i);
    const bool d = q.out_offset >= 0, x = q.idx_offset >= 0;
    TM_CHECK_ARG(d || x, "tm_ragged_dt: job %lld requests no output", (long long)i);
    TM_CHECK_ARG(!d || want_dist, "tm_ragged_dt: job %lld has a distance offset but distances is NULL",
                 (long long)i);
    TM_CHECK_ARG(!x || want_idx, "tm_ragged_dt: job %lld has an index offset but indices is NULL",
                 (long long)i);
    RgJob& j = p.jobs[(size_t)i];
    j.sh = make_shape(q.ndim, q.shape);
    for (int a = 0; a < kMaxNdim; ++a) j.sp.s[a] = 1.0;
    if (q.op == TM_RG_EDT)
      for (int a = 0; a < q.ndim; ++a) {
        const double s = q.sampling[a];
        TM_CHECK_ARG(s > 0.0 && s < 1e300, "tm_ragged_dt: job %lld: sampling must be finite > 0",
                     (long long)i);
        j.sp.s[a] = s;
      }
    j.in_off = q.in_offset;
    j.out_off = d ? q.out_offset : -1;
    j.idx_off = x ? q.idx_offset : -1;
    j.op = q.op;
    j.raster = (q.op != TM_RG_EDT && x) ? 1 : 0;
    j.ws_off = (int64_t)work;
    work += align256(job_work_bytes(j));
  }

  // tasks, level-major, larger jobs first within a level
  std::vector<int32_t> order((size_t)njobs);
  for (int64_t i = 0; i < njobs; ++i) order[(size_t)i] = (int32_t)i;
  auto cost = [&](int32_t i) {
    const RgJob& j = p.jobs[(size_t)i];
    return (double)j.sh.numel * (j.raster ? 8.0 : 1.0);
  };
  std::stable_sort(order.begin(), order.end(),
                   [&](int32_t x, int32_t y) { return cost(x) > cost(y); });
  std::vector<int32_t> done_before((size_t)njobs, 0);
  p.tasks.clear();
  p.nlevels = 0;
  for (int lev = 0; lev < kRgLevels; ++lev) {
    p.level_lo[lev] = (int)p.tasks.size();
    bool any = false;
    for (int32_t ji : order) {
      const RgJob& j = p.jobs[(size_t)ji];
      const int levels = j.raster ? 2 : j.sh.ndim;
      if (lev >= levels) continue;
      any = true;
      int ntask = 0;
      if (j.raster) {
        RgTask t{};
        t.job = ji;
        t.kind = RG_CDT_RASTER;
        t.d = lev;
        t.m = 1;
        t.line0 = 0;
        t.nlines = 1;
        t.need = done_before[(size_t)ji];
        p.tasks.push_back(t);
        ntask = 1;
      } else {
        const int64_t L = j.sh.size[lev];
        const int64_t nl = j.sh.numel / L;
        // about 64k voxels per task, at most 8 lines per lane
        int64_t m = (65536 / kRgThreads + L - 1) / L;
        m = std::max<int64_t>(1, std::min<int64_t>(8, m));
        const int64_t per = m * kRgThreads;
        int32_t kind;
        if (j.op == TM_RG_EDT) kind = lev == 0 ? RG_EDT0 : RG_EDT_STACK;
        else if (lev == 0) kind = RG_CDT0;
        else kind = j.op == TM_RG_CDT_TAXICAB ? RG_CDT_TAXICAB : RG_CDT_CHESSBOARD;
        for (int64_t l0 = 0; l0 < nl; l0 += per) {
          RgTask t{};
          t.job = ji;
          t.kind = kind;
          t.d = lev;
          t.line0 = l0;
          t.nlines = std::min(per, nl - l0);
          t.m = (int32_t)((t.nlines + kRgThreads - 1) / kRgThreads);
          t.need = done_before[(size_t)ji];
          p.tasks.push_back(t);
          ++ntask;
        }
      }
      done_before[(size_t)ji] += ntask;
    }
    if (!any) break;
    p.nlevels = lev + 1;
  }
  p.level_lo[p.nlevels] = (int)p.tasks.size();
  TM_CHECK_ARG(p.tasks.size() < (size_t)1 << 31, "tm_ragged_dt: too many tasks");

  p.ctr_bytes = align256((size_t)(kRgJob0 + njobs) * sizeof(int));
  p.jobs_off = p.ctr_bytes;
  p.tasks_off = p.jobs_off + align256(p.jobs.size() * sizeof(RgJob));
  p.work_off = p.tasks_off + align256(p.tasks.size() * sizeof(RgTask));
  p.total = p.work_off + work;
  return TM_OK;
}

// Pinned staging for the descriptor upload, one per host thread.
static thread_local PinnedStage g_stage;

struct RgDevInfo {
  int cus = 0;
  int occ[2] = {0, 0};
};
static std::mutex g_dev_mu;
static RgDevInfo g_dev[64];

static int grid_limit(int dtype_slot, const void* kern) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 1024;
  std::lock_guard<std::mutex> lk(g_dev_mu);
  RgDevInfo& d = g_dev[dev];
  if (d.cus == 0) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c < 1)
      c = 256;
    d.cus = c;
  }
  if (d.occ[dtype_slot] == 0) {
    int o = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, kern, kRgThreads, 0) != hipSuccess || o < 1)
      o = 2;
    d.occ[dtype_slot] = o;
  }
  return d.cus * d.occ[dtype_slot];
}

template <typename TIn>
static int rg_launch(const TIn* in, const RgPlan& p, RgArgs a, int schedule, int dslot,
                     hipStream_t st) {
  const int nctr = kRgJob0 + (int)p.jobs.size();
  hipLaunchKernelGGL(rg_zero_kernel, dim3(grid_for(nctr, 256)), dim3(256), 0, st, a.ctr, nctr);
  TM_LAUNCH_CHECK();
  const int cap = grid_limit(dslot, (const void*)rg_kernel<TIn>);
  if (schedule == 0) {
    a.qslot = 0;
    a.lo = 0;
    a.hi = (int)p.tasks.size();
    a.wait = 1;
    const int g = std::max(1, std::min(cap, a.hi));
    hipLaunchKernelGGL((rg_kernel<TIn>), dim3(g), dim3(kRgThreads), 0, st, in, a);
    TM_LAUNCH_CHECK();
  } else {
    for (int l = 0; l < p.nlevels; ++l) {
      a.qslot = l;
      a.lo = p.level_lo[l];
      a.hi = p.level_lo[l + 1];
      a.wait = 0;
      const int g = std::max(1, std::min(cap, a.hi - a.lo));
      hipLaunchKernelGGL((rg_kernel<TIn>), dim3(g), dim3(kRgThreads), 0, st, in, a);
      TM_LAUNCH_CHECK();
    }
  }
  return TM_OK;
}

}  // namespace tmh

extern "C" size_t tm_ragged_dt_workspace_bytes(const tm_ragged_job* jobs, int64_t njobs) {
  tmh::RgPlan p;
  bool anyd = false, anyx = false;
  for (int64_t i = 0; jobs && i < njobs; ++i) {
    anyd = anyd || jobs[i].out_offset >= 0;
    anyx = anyx || jobs[i].idx_offset >= 0;
  }
  if (!jobs || tmh::build_plan(jobs, njobs, anyd, anyx, p) != TM_OK) return 0;
  return p.total;
}

extern "C" int tm_ragged_dt(const tm_ragged_job* jobs, int64_t njobs, const void* input,
                            int input_dtype, float* distances, void* indices, int index_dtype,
                            int schedule, int upload, void* workspace, size_t workspace_bytes,
                            void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(jobs != nullptr, "tm_ragged_dt: jobs is NULL");
  TM_CHECK_ARG(input != nullptr, "tm_ragged_dt: input is NULL");
  TM_CHECK_ARG(input_dtype == TM_U8 || input_dtype == TM_F32,
               "tm_ragged_dt: input dtype must be uint8/bool or float32");
  TM_CHECK_ARG(!indices || index_dtype == TM_I32 || index_dtype == TM_I64,
               "tm_ragged_dt: index dtype must be int32 or int64");
  TM_CHECK_ARG(schedule == 0 || schedule == 1,
               "tm_ragged_dt: schedule must be 0 (persistent) or 1 (one launch per level)");
  RgPlan p;
  const int rc = build_plan(jobs, njobs, distances != nullptr, indices != nullptr, p);
  if (rc != TM_OK) return rc;
  if (!workspace || workspace_bytes < p.total) {
    set_error("tm_ragged_dt: workspace of %zu bytes is smaller than the %zu required",
              workspace_bytes, p.total);
    return TM_ERR_WORKSPACE;
  }
  TM_CHECK_ARG(((uintptr_t)workspace % 256) == 0, "tm_ragged_dt: workspace must be 256-byte aligned");
  hipStream_t st = (hipStream_t)stream;
  char* ws = (char*)workspace;

  if (upload) {
    const size_t jb = p.jobs.size() * sizeof(RgJob);
    const size_t tb = p.tasks.size() * sizeof(RgTask);
    const size_t need = p.work_off - p.jobs_off;
    PinnedStage& s = g_stage;
    int rc2 = stage_acquire(s, need);
    if (rc2 != TM_OK) return rc2;
    char* h = (char*)s.host;
    memcpy(h, p.jobs.data(), jb);
    memcpy(h + (p.tasks_off - p.jobs_off), p.tasks.data(), tb);
    rc2 = stage_upload(s, ws + p.jobs_off, need, st);
    if (rc2 != TM_OK) return rc2;
  }

  RgArgs a;
  a.dist = distances;
  a.idx = indices;
  a.idx64 = index_dtype == TM_I64 ? 1 : 0;
  a.work = ws + p.work_off;
  a.jobs = (const RgJob*)(ws + p.jobs_off);
  a.tasks = (const RgTask*)(ws + p.tasks_off);
  a.ctr = (int*)ws;
  a.qslot = 0;
  a.lo = a.hi = 0;
  a.wait = 0;
  if (input_dtype == TM_U8) return rg_launch((const uint8_t*)input, p, a, schedule, 0, st);
  return rg_launch((const float*)input, p, a, schedule, 1, st);
}

extern "C" int tm_ragged_dt_check(const void* workspace, void* stream) {
  using namespace tmh;
  TM_CHECK_ARG(workspace != nullptr, "tm_ragged_dt_check: workspace is NULL");
  int flag = 0;
  hipStream_t st = (hipStream_t)stream;
  TM_CHECK_HIP(hipMemcpyAsync(&flag, (const int*)workspace + kRgErr, sizeof(int),
                              hipMemcpyDeviceToHost, st));
  TM_CHECK_HIP(hipStreamSynchronize(st));
  if (flag != 0) {
    set_error("tm_ragged_dt: a task's dependency wait exceeded its spin limit; outputs are invalid");
    return TM_ERR_HIP;
  }
  return TM_OK;
}
