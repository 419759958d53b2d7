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
// schedule = 2 has no work-list: one ordinary grid launch per (level, kind
// of pass), one line per lane, and a per-block table naming each block's job
// and first line, longest lines first.  Each launch is specialised to its
// kind of pass (EDT axis 0 and chamfer axis 0 from per-chunk background
// bitmasks (tm_dt_axis0.h), EDT envelope, taxicab, chessboard, raster); the
// EDT envelope is the prefetching form of tm_dt_fast.h with its stacks in
// the blocked global layout.  Passes along an item's contiguous last axis
// (distances-only EDT of 2-D and 3-D items, taxicab chamfer) run row-staged
// instead (tm_dt_chunk.h): a workgroup owns 256 consecutive rows and moves
// along them in chunks staged through LDS.  Chessboard passes along every
// axis >= 1 are computed per voxel by exact expanding windows
// (cheb_window), with no stack.  Stream order provides the dependencies.
//
// The per-line code is the dense kernels' (tm_dt_lines.h, tm_dt_fast.h), so
// every job's result is bit-identical to the dense entry point on that item
// alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "tm_cdt_dev.h"
#include "tm_dt_axis0.h"
#include "tm_dt_chunk.h"
#include "tm_dt_fast.h"
#include "tm_dt_lines.h"
#include "tm_tune.h"

namespace tmh {

enum RgKind : int32_t {
  RG_EDT0 = 0,
  RG_EDT_STACK = 1,
  RG_CDT0 = 2,
  RG_CDT_TAXICAB = 3,
  RG_CDT_CHESSBOARD = 4,
  RG_CDT_RASTER = 5,
  // schedule 2 only: row-staged passes along the contiguous last axis
  RG_EDT_ROWS2 = 6,  // EDT, 2-D item, distances only
  RG_EDT_ROWS3 = 7,  // EDT, 3-D item, distances only
  RG_CDT_TAXICAB_ROWS = 8,
  RG_CDT_CHEB_WINDOW = 9  // chessboard axis >= 1, one voxel per lane
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
  int64_t w0_off;   // bytes into the job's work: axis-0 chunk words (schedule 2)
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

// Schedule 2: one block of lines of one job (lines line0 .. line0 + nlines - 1
// of the job's pass at the launch's level).
struct RgBlk {
  int32_t job;
  int32_t nlines;
  int64_t line0;
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

// Schedule 2: one launch per (level d, kind); lane = line.  The block table
// of a launch lists the longest lines first and is read in block id order.
template <typename TIn, int KIND>
__global__ __launch_bounds__(kRgThreads) void rg_level_kernel(const TIn* __restrict__ in_all,
                                                              RgArgs a,
                                                              const RgBlk* __restrict__ blks,
                                                              int d) {
  const RgBlk bk = blks[blockIdx.x];
  if ((int)threadIdx.x >= bk.nlines) return;
  const RgJob& jb = a.jobs[bk.job];
  const Shape& sh = jb.sh;
  const TIn* in = in_all + jb.in_off;
  char* wk = a.work + jb.ws_off;
  const int64_t numel = sh.numel;
  const int nd = sh.ndim;
  const bool fin = d == nd - 1;
  const int64_t nl_axis = numel / sh.size[d];
  const int64_t t = bk.line0 + threadIdx.x;
  if (KIND == RG_EDT0 || KIND == RG_EDT_STACK) {
    EdtOut o;
    o.dist = (a.dist && jb.out_off >= 0) ? a.dist + jb.out_off : nullptr;
    o.idx = (a.idx && jb.idx_off >= 0) ? (char*)a.idx + jb.idx_off * (a.idx64 ? 8 : 4) : nullptr;
    o.idx64 = a.idx64;
    o.plane = numel;
    const int nbuf = nd > 1 ? nd - 1 : 1;
    int32_t* A = (int32_t*)wk;
    int32_t* B = A + (int64_t)nbuf * numel;
    if (KIND == RG_EDT0) {
      uint32_t* words = (uint32_t*)(wk + jb.w0_off);
      if (fin) edt_axis0_words<TIn, true>(in, A, sh, t, jb.sp, o, words);
      else edt_axis0_words<TIn, false>(in, A, sh, t, jb.sp, o, words);
    } else {
      const int32_t* src = (d & 1) ? A : B;
      int32_t* dst = (d & 1) ? B : A;
      int32_t* gpos = B + (int64_t)nbuf * numel;
      double* gh = (double*)(((uintptr_t)(gpos + numel) + 255) & ~(uintptr_t)255);
      if (nd <= 3) {  // compact entries, blocked global stacks (tm_dt_fast.h)
        StkAddr<true> A;
        A.E = rows_stack_e(gpos);
        A.F1 = rows_stack_f1(gpos, sh.size[d], nl_axis);
        A.ss = 0;
        A.nl8 = nl_axis * 8;
        A.t8 = t * 8;
        if (d == 1) {
          if (fin) edt_line_stack_c<true, 1, 8, true>(src, dst, numel, sh, t, jb.sp, A, o);
          else edt_line_stack_c<false, 1, 8, true>(src, dst, numel, sh, t, jb.sp, A, o);
        } else {
          if (fin) edt_line_stack_c<true, 2, 8, true>(src, dst, numel, sh, t, jb.sp, A, o);
          else edt_line_stack_c<false, 2, 8, true>(src, dst, numel, sh, t, jb.sp, A, o);
        }
      } else {
        if (fin)
          edt_line_stack_fast<true, 8>(src, dst, numel, sh, d, t, jb.sp, gpos + t, gh + t,
                                       nl_axis, o);
        else
          edt_line_stack_fast<false, 8>(src, dst, numel, sh, d, t, jb.sp, gpos + t, gh + t,
                                        nl_axis, o);
      }
    }
  } else {
    int32_t* w = (int32_t*)(a.dist + jb.out_off);
    // chessboard passes alternate between the work area and the output so
    // that the last one lands in the output (cheb_window is not in place)
    if (KIND == RG_CDT0 && jb.op == TM_RG_CDT_CHESSBOARD && ((nd - 1) & 1)) w = (int32_t*)wk;
    int64_t c[kMaxNdim], base, b;
    decode_line(sh, d, t, c, &base, &b);
    const int64_t step = sh.stride[d], L = sh.size[d];
    if (KIND == RG_CDT0) {
      uint32_t* words = (uint32_t*)(wk + jb.w0_off);
      if (fin) cdt_axis0_words<TIn, true>(in, w, base, step, (int)L, t, nl_axis, words);
      else cdt_axis0_words<TIn, false>(in, w, base, step, (int)L, t, nl_axis, words);
    } else if (KIND == RG_CDT_TAXICAB) {
      if (fin) cdt_taxicab_chunked<true>(w, base, step, (int)L);
      else cdt_taxicab_chunked<false>(w, base, step, (int)L);
    } else {
      uint16_t* S = (uint16_t*)wk;
      uint16_t* T = S + numel;
      int32_t* G = (int32_t*)(T + numel);
      if (fin) cdt_line_chessboard<true>(w, base, step, L, S + t, T + t, G + t, nl_axis);
      else cdt_line_chessboard<false>(w, base, step, L, S + t, T + t, G + t, nl_axis);
    }
  }
}

// Schedule 2, chessboard pass d >= 1 of a job without indices: a block is
// kRgWinVox consecutive voxels of the item (kRgWinVox / kRgThreads per lane),
// each computed by cheb_window from the previous pass's values.
constexpr int kRgWinVox = 4 * kRgThreads;

__global__ __launch_bounds__(kRgThreads) void rg_cheb_window_kernel(
    RgArgs a, const RgBlk* __restrict__ blks, int d) {
  const RgBlk bk = blks[xcd_swizzle(blockIdx.x, gridDim.x)];
  const RgJob& jb = a.jobs[bk.job];
  const Shape& sh = jb.sh;
  const int nd = sh.ndim;
  int32_t* out = (int32_t*)(a.dist + jb.out_off);
  int32_t* buf = (int32_t*)(a.work + jb.ws_off);
  int32_t* dst = ((nd - 1 - d) & 1) ? buf : out;
  const int32_t* src = ((nd - d) & 1) ? buf : out;
  const bool fin = d == nd - 1;
  const int64_t step = sh.stride[d];
  const int L = (int)sh.size[d];
  for (int j = 0; j < kRgWinVox / kRgThreads; ++j) {
    const int e = j * kRgThreads + (int)threadIdx.x;
    if (e >= bk.nlines) break;
    const int64_t v = bk.line0 + e;
    const int64_t hi = v / step;
    const int i = (int)(hi % L);
    const int64_t base = v - (int64_t)i * step;
    const int32_t r = cheb_window(src, base, step, L, i);
    dst[v] = fin ? cdt_final_bits(r) : r;
  }
}

// Schedule 2, row-staged passes along the last axis (d = ndim - 1): each
// workgroup owns kCLanes consecutive rows of one job and walks them in chunks
// of CH voxels (a launch variant: more voxels per chunk mean fewer barriers
// and more LDS and registers per workgroup).
template <int KIND, int CH>
__global__ __launch_bounds__(kCLanes) void rg_rows_kernel(RgArgs a,
                                                          const RgBlk* __restrict__ blks) {
  extern __shared__ int32_t lds_rows[];
  const RgBlk bk = blks[blockIdx.x];  // longest rows first
  const RgJob& jb = a.jobs[bk.job];
  const Shape& sh = jb.sh;
  char* wk = a.work + jb.ws_off;
  const int64_t numel = sh.numel;
  const int nd = sh.ndim;
  const int64_t nl_axis = numel / sh.size[nd - 1];
  if (KIND == RG_EDT_ROWS2 || KIND == RG_EDT_ROWS3) {
    const int nbuf = nd > 1 ? nd - 1 : 1;
    int32_t* A = (int32_t*)wk;
    int32_t* B = A + (int64_t)nbuf * numel;
    const int32_t* src = ((nd - 1) & 1) ? A : B;
    int32_t* area = B + (int64_t)nbuf * numel;
    const int64_t L = sh.size[nd - 1];
    uint32_t* E = rows_stack_e(area);
    uint16_t* F1 = rows_stack_f1(area, L, nl_axis);
    if (KIND == RG_EDT_ROWS2)
      edt_rows_final<1, CH>(src, numel, sh, bk.line0, bk.nlines, jb.sp, E, F1, nl_axis,
                            a.dist + jb.out_off, lds_rows);
    else
      edt_rows_final<2, CH>(src, numel, sh, bk.line0, bk.nlines, jb.sp, E, F1, nl_axis,
                            a.dist + jb.out_off, lds_rows);
  } else {
    cdt_rows_taxicab<CH>((int32_t*)(a.dist + jb.out_off), sh, bk.line0, bk.nlines, lds_rows);
  }
}

// LDS bytes of rg_rows_kernel<KIND, CH>.
static size_t rows_lds(int kind, int ch) {
  if (kind == RG_EDT_ROWS3) return (size_t)2 * kCLanes * (ch + 1) * sizeof(int32_t);
  return (size_t)kCLanes * (ch + 1) * sizeof(int32_t);
}

// Chunk depths (2-D EDT rows, 3-D EDT rows, taxicab rows) of per-pass launch
// variant v; variant 0 is the default.
constexpr int kRgRowVariants = 5;
static const int kRgRowCH[kRgRowVariants][3] = {{32, 16, 32}, {16, 8, 16}, {64, 32, 64},
                                                {16, 16, 16}, {32, 16, 64}};

// Schedule 2, chamfer with indices: one workgroup per job runs raster pass d
// (every thread takes part: the pass has workgroup barriers).
template <typename TIn>
__global__ __launch_bounds__(kRgThreads) void rg_raster_kernel(const TIn* __restrict__ in_all,
                                                               RgArgs a,
                                                               const RgBlk* __restrict__ blks,
                                                               int d) {
  const RgJob& jb = a.jobs[blks[blockIdx.x].job];
  const int64_t numel = jb.sh.numel;
  CdtOut o;
  o.dist = (a.dist && jb.out_off >= 0) ? a.dist + jb.out_off : nullptr;
  o.idx = (a.idx && jb.idx_off >= 0) ? (char*)a.idx + jb.idx_off * (a.idx64 ? 8 : 4) : nullptr;
  o.idx64 = a.idx64;
  o.plane = numel;
  int32_t* dt = (int32_t*)(a.work + jb.ws_off);
  cdt_item_pass(in_all + jb.in_off, dt, dt + numel, jb.sh, jb.sh.ndim == 3 ? 1 : 0,
                jb.op == TM_RG_CDT_TAXICAB ? 1 : 0, d, o);
}

// ------------------------------------------------------------------ host --

// One schedule-2 launch: blocks [lo, hi) of the block table, all of one
// kind of pass at one level.
// chain: the group of jobs whose pass chain the launch belongs to (schedule
// 2 runs each chain on its own stream): 0 EDT of 3-D items, 1 EDT of other
// items, 2 taxicab, 3 chessboard and raster (chamfer with indices, on the
// caller's stream).
struct RgLaunch {
  int level, kind, lo, hi, chain;
};

struct RgPlan {
  std::vector<RgJob> jobs;
  std::vector<RgTask> tasks;
  std::vector<RgBlk> blks;
  std::vector<RgLaunch> launches;
  int level_lo[kRgLevels + 1];
  int nlevels;
  size_t ctr_bytes, jobs_off, tasks_off, blks_off, work_off, total;
};

// Whether the job's last EDT pass runs row-staged (rg_rows_kernel).
static bool rows_final(const RgJob& j) {
  const int nd = j.sh.ndim;
  return j.op == TM_RG_EDT && !j.raster && j.idx_off < 0 && nd >= 2 && nd <= 3;
}

static size_t job_pass_bytes(const RgJob& j) {
  const int64_t n = j.sh.numel;
  const int nd = j.sh.ndim;
  if (j.op == TM_RG_EDT) {
    // feature planes A and B (ndim-1 each, one for 1-D), stack positions, stack values
    const int nbuf = nd > 1 ? nd - 1 : 1;
    size_t b = align256((size_t)n * 4 * (2 * nbuf + 1)) + (size_t)n * 8;
    if (nd <= 3)  // the blocked compact stacks (tm_dt_fast.h) after the planes, per pass
      for (int d = 1; d < nd; ++d)
        b = std::max(b, (size_t)n * 8 * nbuf + (size_t)rows_stack_bytes(j.sh.size[d],
                                                                          n / j.sh.size[d]));
    return b;
  }
  if (j.raster) return (size_t)n * 8;
  return j.op == TM_RG_CDT_CHESSBOARD ? (size_t)n * 8 : 0;
}

// The pass buffers, then the chunk words of the first pass (tm_dt_axis0.h).
static size_t job_work_bytes(const RgJob& j) {
  if (j.raster) return job_pass_bytes(j);
  return align256(job_pass_bytes(j)) +
         (size_t)axis0_words_bytes(j.sh.size[0], j.sh.numel / j.sh.size[0]);
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
    j.w0_off = (int64_t)align256(job_pass_bytes(j));
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

  // schedule 2: per level, per kind, blocks of one line per lane
  p.blks.clear();
  p.launches.clear();
  auto chain_of = [&](const RgJob& j) {
    if (j.raster) return 3;
    if (j.op == TM_RG_EDT) return j.sh.ndim == 3 ? 0 : 1;
    return j.op == TM_RG_CDT_TAXICAB ? 2 : 3;
  };
  for (int lev = 0; lev < p.nlevels; ++lev)
  for (int chain = 0; chain < 4; ++chain) {
    for (int g = 0; g < 10; ++g) {
      // g: 0 EDT0, 1 EDT envelope, 2 CDT0 of taxicab jobs, 3 taxicab, 4
      //    chessboard (every axis >= 1, voxel blocks, cheb_window), 5 raster;
      //    along the contiguous last axis, row-staged: 6 EDT 2-D, 7 EDT 3-D,
      //    8 taxicab; 9 CDT0 of chessboard jobs
      RgLaunch ln{lev, 0, (int)p.blks.size(), 0, 0};
      for (int32_t ji : order) {
        const RgJob& j = p.jobs[(size_t)ji];
        if (chain_of(j) != chain) continue;
        if (j.raster) {
          if (g != 5 || lev >= 2) continue;
          p.blks.push_back(RgBlk{ji, 1, 0});
          continue;
        }
        if (lev >= j.sh.ndim) continue;
        const int64_t L = j.sh.size[lev];
        const bool last = lev >= 1 && lev == j.sh.ndim - 1;
        int grp;
        if (j.op == TM_RG_EDT) {
          if (lev == 0) grp = 0;
          else if (last && rows_final(j)) grp = j.sh.ndim == 2 ? 6 : 7;
          else grp = 1;
        } else if (lev == 0) {
          grp = j.op == TM_RG_CDT_TAXICAB ? 2 : 9;
        } else if (j.op == TM_RG_CDT_TAXICAB) {
          grp = last ? 8 : 3;
        } else {
          grp = 4;
        }
        if (grp != g) continue;
        if (g == 4) {  // voxel blocks
          for (int64_t v0 = 0; v0 < j.sh.numel; v0 += kRgWinVox)
            p.blks.push_back(RgBlk{ji, (int32_t)std::min<int64_t>(kRgWinVox, j.sh.numel - v0), v0});
          continue;
        }
        const int64_t nl = j.sh.numel / L;
        for (int64_t l0 = 0; l0 < nl; l0 += kRgThreads)
          p.blks.push_back(RgBlk{ji, (int32_t)std::min<int64_t>(kRgThreads, nl - l0), l0});
      }
      ln.hi = (int)p.blks.size();
      if (g != 4 && g != 5) {
        // a lane walks its whole line: longest lines first, so that the
        // longest blocks start in the first round of workgroups (launched
        // without the XCD swizzle, block id order = dispatch order)
        std::stable_sort(p.blks.begin() + ln.lo, p.blks.begin() + ln.hi,
                         [&](const RgBlk& x, const RgBlk& y) {
                           return p.jobs[(size_t)x.job].sh.size[lev] >
                                  p.jobs[(size_t)y.job].sh.size[lev];
                         });
      }
      static const int kinds[10] = {RG_EDT0,       RG_EDT_STACK,  RG_CDT0,
                                    RG_CDT_TAXICAB, RG_CDT_CHEB_WINDOW, RG_CDT_RASTER,
                                    RG_EDT_ROWS2,  RG_EDT_ROWS3,  RG_CDT_TAXICAB_ROWS,
                                    RG_CDT0};
      ln.kind = kinds[g];
      ln.chain = chain;
      if (ln.hi > ln.lo) p.launches.push_back(ln);
    }
  }
  TM_CHECK_ARG(p.blks.size() < (size_t)1 << 31, "tm_ragged_dt: too many line blocks");

  p.ctr_bytes = align256((size_t)(kRgJob0 + njobs) * sizeof(int));
  p.jobs_off = p.ctr_bytes;
  p.tasks_off = p.jobs_off + align256(p.jobs.size() * sizeof(RgJob));
  p.blks_off = p.tasks_off + align256(p.tasks.size() * sizeof(RgTask));
  p.work_off = p.blks_off + align256(p.blks.size() * sizeof(RgBlk));
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

// Side streams of schedule 2, one set per host thread and device: the EDT,
// taxicab and chessboard pass chains are independent (their jobs share no
// memory), so each runs on its own stream and the long last blocks of one
// chain's launches overlap the other chains' work.  The chains fork from
// the caller's stream and join it again through events, which also holds
// while the caller's stream is being captured into a graph.  Streams at the
// default priority; never destroyed (a captured graph may refer to them).
struct RgSideStreams {
  bool ok = false, tried = false;
  hipStream_t s[3] = {nullptr, nullptr, nullptr};
  hipEvent_t fork = nullptr, join[3] = {nullptr, nullptr, nullptr};
};
constexpr int kRgMaxDevices = 16;
static thread_local RgSideStreams g_rg_side[kRgMaxDevices];

static RgSideStreams* rg_side_streams() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kRgMaxDevices) return nullptr;
  RgSideStreams& r = g_rg_side[dev];
  if (!r.tried) {
    r.tried = true;
    bool ok = hipEventCreateWithFlags(&r.fork, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; i < 3 && ok; ++i)
      ok = hipStreamCreateWithFlags(&r.s[i], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&r.join[i], hipEventDisableTiming) == hipSuccess;
    r.ok = ok;
  }
  return r.ok ? &r : nullptr;
}

template <typename TIn>
static int rg_launch(const TIn* in, const RgPlan& p, RgArgs a, int schedule, int dslot,
                     hipStream_t st, int variant = 0) {
  const int nctr = kRgJob0 + (int)p.jobs.size();
  hipLaunchKernelGGL(rg_zero_kernel, dim3(grid_for(nctr, 256)), dim3(256), 0, st, a.ctr, nctr);
  TM_LAUNCH_CHECK();
  if (schedule == 2) {
    const RgBlk* blks = (const RgBlk*)(a.work - (p.work_off - p.blks_off));
    RgSideStreams* side = rg_side_streams();
    bool used[3] = {false, false, false};  // chain 3 runs on the caller's stream
    for (const RgLaunch& ln : p.launches)
      if (ln.chain < 3) used[ln.chain] = true;
    if (side) {
      if (hipEventRecord(side->fork, st) != hipSuccess) side = nullptr;
      for (int i = 0; i < 3 && side; ++i)
        if (used[i] && hipStreamWaitEvent(side->s[i], side->fork, 0) != hipSuccess) {
          set_error("tm_ragged_dt: stream fork failed");
          return TM_ERR_HIP;
        }
    }
    hipStream_t const st0 = st;
    for (const RgLaunch& ln : p.launches) {
      const RgBlk* b = blks + ln.lo;
      const dim3 g((unsigned)(ln.hi - ln.lo));
      const int d = ln.level;
      hipStream_t st = (side && ln.chain < 3) ? side->s[ln.chain] : st0;
      switch (ln.kind) {
        case RG_EDT0:
          hipLaunchKernelGGL((rg_level_kernel<TIn, RG_EDT0>), g, dim3(kRgThreads), 0, st, in, a,
                             b, d);
          break;
        case RG_EDT_STACK:
          hipLaunchKernelGGL((rg_level_kernel<TIn, RG_EDT_STACK>), g, dim3(kRgThreads), 0, st, in,
                             a, b, d);
          break;
        case RG_CDT0:
          hipLaunchKernelGGL((rg_level_kernel<TIn, RG_CDT0>), g, dim3(kRgThreads), 0, st, in, a,
                             b, d);
          break;
        case RG_CDT_TAXICAB:
          hipLaunchKernelGGL((rg_level_kernel<TIn, RG_CDT_TAXICAB>), g,
                             dim3(kRgThreads), 0, st, in, a, b, d);
          break;
        case RG_CDT_CHEB_WINDOW:
          hipLaunchKernelGGL(rg_cheb_window_kernel, g, dim3(kRgThreads), 0, st, a, b, d);
          break;
        case RG_EDT_ROWS2:
        case RG_EDT_ROWS3:
        case RG_CDT_TAXICAB_ROWS: {
          const int slot = ln.kind == RG_EDT_ROWS2 ? 0 : (ln.kind == RG_EDT_ROWS3 ? 1 : 2);
          const int ch = kRgRowCH[variant][slot];
          const size_t lds = rows_lds(ln.kind, ch);
#define TM_RG_ROWS(K, C) \
  hipLaunchKernelGGL((rg_rows_kernel<K, C>), g, dim3(kCLanes), lds, st, a, b)
#define TM_RG_ROWS_CH(K)             \
  switch (ch) {                      \
    case 8: TM_RG_ROWS(K, 8); break;   \
    case 16: TM_RG_ROWS(K, 16); break; \
    case 32: TM_RG_ROWS(K, 32); break; \
    default: TM_RG_ROWS(K, 64); break; \
  }
          if (ln.kind == RG_EDT_ROWS2) {
            TM_RG_ROWS_CH(RG_EDT_ROWS2);
          } else if (ln.kind == RG_EDT_ROWS3) {
            TM_RG_ROWS_CH(RG_EDT_ROWS3);
          } else {
            TM_RG_ROWS_CH(RG_CDT_TAXICAB_ROWS);
          }
#undef TM_RG_ROWS_CH
#undef TM_RG_ROWS
          break;
        }
        default:
          hipLaunchKernelGGL((rg_raster_kernel<TIn>), g, dim3(kRgThreads), 0, st, in, a, b, d);
          break;
      }
      TM_LAUNCH_CHECK();
    }
    for (int i = 0; i < 3 && side; ++i)
      if (used[i] && (hipEventRecord(side->join[i], side->s[i]) != hipSuccess ||
                      hipStreamWaitEvent(st0, side->join[i], 0) != hipSuccess)) {
        set_error("tm_ragged_dt: stream join failed");
        return TM_ERR_HIP;
      }
    return TM_OK;
  }
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
  TM_CHECK_ARG(schedule >= 0 && schedule <= 2,
               "tm_ragged_dt: schedule must be 0 (persistent), 1 (one launch per level) or 2 "
               "(one launch per level and kind of pass)");
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
    memcpy(h + (p.blks_off - p.jobs_off), p.blks.data(), p.blks.size() * sizeof(RgBlk));
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
  auto run = [&](int v) -> int {
    if (input_dtype == TM_U8) return rg_launch((const uint8_t*)input, p, a, schedule, 0, st, v);
    return rg_launch((const float*)input, p, a, schedule, 1, st, v);
  };
  if (schedule != 2) return run(0);
  // per-pass schedule: the row-staged chunk depth is tuned per shape class
  // of the job list (item count, voxels, longest row, kinds of row pass)
  int64_t vox = 0, maxrow = 0, kinds = 0;
  for (const RgJob& j : p.jobs) {
    vox += j.sh.numel;
    maxrow = std::max<int64_t>(maxrow, j.sh.size[j.sh.ndim - 1]);
  }
  for (const RgLaunch& ln : p.launches)
    if (ln.kind == RG_EDT_ROWS2 || ln.kind == RG_EDT_ROWS3 || ln.kind == RG_CDT_TAXICAB_ROWS)
      kinds |= 1 << (ln.kind - RG_EDT_ROWS2);
  if (!kinds) return run(0);
  auto p2 = [](int64_t v) { int64_t q = 1; while (q < v) q <<= 1; return q; };
  const int64_t key[4] = {kinds, p2((int64_t)p.jobs.size()), p2((vox + 65535) >> 16), p2(maxrow)};
  bool ran = false;
  const int v = tune_pick("ragged_rows", key, 4, kRgRowVariants, 0, run, st, &ran);
  if (ran) return TM_OK;
  return run(v);
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
