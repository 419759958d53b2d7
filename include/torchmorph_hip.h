/*
 * torchmorph_hip.h -- C ABI of libtorchmorph_hip.so (MI355X / gfx950).
 *
 * Plain pointers and sizes only: every buffer is device memory allocated by
 * the caller (PyTorch's caching allocator in the Python layer), every launch
 * goes on the caller's stream (`stream` is a hipStream_t passed as void*,
 * NULL = default stream), nothing synchronises the host unless stated.
 * Every entry point returns TM_OK or a negative status; tm_last_error()
 * returns the message of the last failure on the calling host thread.
 *
 * Layouts: a batch of items is (batch, s0, ..., s{ndim-1}) C-contiguous with
 * batch = B*C of the reference's (B, C, spatial...) tensors.  Index outputs are
 * (ndim, batch, spatial...) planes of int32 or int64 written directly.
 *
 * The reference boundary these replace is the pybind11 module torchmorph._C
 * (reference torchmorph/csrc/torchmorph.cpp:79-155); each declaration names
 * the function it stands in for.
 */
#ifndef TORCHMORPH_HIP_H
#define TORCHMORPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum tm_status {
  TM_OK = 0,
  TM_ERR_ARG = -1,     /* invalid argument (maps to RuntimeError) */
  TM_ERR_HIP = -2,     /* HIP runtime / launch failure */
  TM_ERR_WORKSPACE = -3 /* workspace too small */
};

enum tm_dtype { TM_U8 = 0, TM_F32 = 1, TM_I32 = 2, TM_I64 = 3, TM_F64 = 4 };

/* Copies the last error message of this host thread into buf (NUL
 * terminated, truncated to len).  Returns the full message length. */
int tm_last_error(char* buf, size_t len);

/* Library version string, e.g. "torchmorph_hip 0.1 gfx950". */
const char* tm_version(void);

/* ----------------------------------------------------------- autotune --- */
/* No reference counterpart (the reference hard-codes its launch constants,
 * edt_kernel.cu:18-20, ot_kernel.cu:8-9).  Launch variants (tile, rows per
 * band, columns per thread, lines per wave, LDS depth) are timed per shape
 * class on first use and cached; every variant returns identical results.
 * Tuning is skipped while the stream is capturing a graph. */
void tm_autotune_enable(int enabled);
void tm_autotune_clear(void);
/* Text form of the cache, one "op key... = variant" per line; returns the
 * full length. */
int tm_autotune_export(char* buf, size_t len);
/* Loads entries in the export format; returns the number read. */
int tm_autotune_import(const char* text);

/* ---------------------------------------------------------------- EDT --- */

/* Workspace bytes tm_edt needs for this problem (0 when none). */
size_t tm_edt_workspace_bytes(int64_t batch, int ndim, const int64_t* shape,
                              const double* sampling);

/* Exact Euclidean distance transform with SciPy's feature-transform tie
 * rules.  Replaces torchmorph._C.edt_cuda (reference torchmorph.cpp:35-40,
 * edt_kernel.cu:904-958).
 *   input        mask, nonzero = foreground; dtype TM_U8 or TM_F32
 *   sampling     ndim spacings (NULL = unit)
 *   distances    float32 (batch, spatial...) or NULL
 *   indices      (ndim, batch, spatial...) of index_dtype (TM_I32/TM_I64) or
 *                NULL; an all-foreground item gets index (-1, 0, ..., 0)
 */
int tm_edt(const void* input, int input_dtype, int64_t batch, int ndim, const int64_t* shape,
           const double* sampling, float* distances, void* indices, int index_dtype,
           void* workspace, size_t workspace_bytes, void* stream);

/* No counterpart in the reference (its edt_cuda allocates its scratch with
 * torch on every call).  For callers that cannot allocate -- code under
 * stream capture, tight serving loops, harnesses that forbid hipMalloc --
 * this runs tm_edt with its workspace in a file-scope device array of
 * 32 MiB (each GPU the library runs on holds one until the process ends),
 * looked up once per device.  Same arguments and results as tm_edt without
 * the workspace; a problem whose tm_edt_workspace_bytes exceeds 32 MiB is
 * rejected with TM_ERR_ARG (the 2-D unit-sampling path needs 2 bytes per
 * voxel, the N-D path about 4 (2 ndim - 1) + 8 bytes per voxel plus its
 * stacks).  No hipMalloc, hipFree or other allocating call; the autotuner
 * times nothing here (cached or default launch variant).  Every call on one
 * GPU shares the array: calls that overlap on one GPU race. */
int tm_edt_noalloc(const void* input, int input_dtype, int64_t batch, int ndim,
                   const int64_t* shape, const double* sampling, float* distances,
                   void* indices, int index_dtype, void* stream);

/* ---------------------------------------------------------------- CDT --- */

/* Workspace bytes tm_cdt needs (int32 distance + int32 feature per voxel). */
size_t tm_cdt_workspace_bytes(int64_t batch, int ndim, const int64_t* shape);

/* Chamfer distance transform with SciPy's two-pass raster semantics, values
 * and indices bit-identical to scipy.ndimage.distance_transform_cdt.
 * Replaces torchmorph._C.cdt_cuda (reference torchmorph.cpp:42-47,
 * cdt_kernel.cu:269-463).
 *   metric      0 = chessboard, 1 = taxicab
 *   distances   float32, -1 where no background is reachable, or NULL
 *   indices     (ndim, batch, spatial...) int32/int64 feature coordinates
 */
int tm_cdt(const void* input, int input_dtype, int64_t batch, int ndim, const int64_t* shape,
           int metric, float* distances, void* indices, int index_dtype, void* workspace,
           size_t workspace_bytes, void* stream);

/* ---------------------------------------------------- brute-force DT --- */

/* Brute-force distance transform: every foreground voxel scans every
 * background voxel of its item.  Replaces torchmorph._C.bfdt_cuda (reference
 * torchmorph.cpp:49-55, bfdt_kernel.cu:170-296).  O(voxels^2) per item: a
 * checker for larger fixtures, not a fast path.  Values are computed in FP64
 * and rounded to float32; the nearest background is the first in row-major
 * order at the smallest distance.
 *   metric      0 = euclidean, 1 = taxicab, 2 = chessboard (sampling applies
 *               to all three, as in the reference)
 *   sampling    ndim spacings (NULL = unit)
 *   distances   float32 (batch, spatial...) or NULL; background 0; an item
 *               without background gets +inf (euclidean) or 4294967295
 *   indices     (ndim, batch, spatial...) of index_dtype or NULL; background
 *               voxels index themselves, an item without background gets 0
 * batch <= 65535; no workspace. */
int tm_bfdt(const void* input, int input_dtype, int64_t batch, int ndim, const int64_t* shape,
            const double* sampling, int metric, float* distances, void* indices,
            int index_dtype, void* stream);

/* ------------------------------------------------------------- ragged --- */

/* No reference counterpart: the reference takes dense (B, C, spatial...)
 * tensors only (distance_transform.py:10-138, edt_kernel.cu:904-958,
 * cdt_kernel.cu:269-463), so a list of differently shaped masks costs one
 * call per shape there.  Here every (item, op) pair is a job; all jobs run
 * together (one persistent launch over a device work-list, or one launch per
 * pass level and kind of pass), and each job's result is bit-identical to
 * tm_edt / tm_cdt on that item alone. */
enum tm_ragged_op { TM_RG_EDT = 0, TM_RG_CDT_CHESSBOARD = 1, TM_RG_CDT_TAXICAB = 2 };

typedef struct tm_ragged_job {
  int32_t op;          /* tm_ragged_op */
  int32_t ndim;        /* 1..8 */
  int64_t shape[8];    /* each axis 1..65535, product below 2^31 */
  double sampling[8];  /* EDT spacing per axis, finite > 0; ignored by the chamfer ops */
  int64_t in_offset;   /* element offset of the C-contiguous item in `input` */
  int64_t out_offset;  /* element offset of its float32 distances, or -1 */
  int64_t idx_offset;  /* element offset of its indices (ndim planes of prod(shape)), or -1 */
} tm_ragged_job;

/* Workspace bytes for these jobs (0 on invalid jobs). */
size_t tm_ragged_dt_workspace_bytes(const tm_ragged_job* jobs, int64_t njobs);

/* Runs every job.  input: TM_U8 or TM_F32 masks (nonzero = foreground) shared
 * by all jobs; distances / indices: flat outputs addressed by the jobs'
 * offsets (NULL when no job uses them); index_dtype TM_I32 / TM_I64.
 *   schedule   0 = one persistent launch with dependency counters,
 *              1 = one launch per pass level,
 *              2 = one launch per pass level and kind of pass, one line per
 *                  lane (passes along the contiguous axis staged through LDS,
 *                  chessboard passes per voxel); no dependency waits
 *   upload     1 = copy the job and task descriptors into the workspace
 *              (pinned staging, asynchronous); 0 = the workspace already holds
 *              them from an earlier call with the same jobs, only launch
 *   workspace  256-byte aligned, tm_ragged_dt_workspace_bytes() bytes
 * A chamfer job with indices runs SciPy's raster passes on one workgroup; a
 * chamfer job without indices runs separably (same values).  The persistent
 * schedule bounds every dependency wait; tm_ragged_dt_check() reports a wait
 * that hit its bound.  Not for graph capture when upload = 1. */
int tm_ragged_dt(const tm_ragged_job* jobs, int64_t njobs, const void* input, int input_dtype,
                 float* distances, void* indices, int index_dtype, int schedule, int upload,
                 void* workspace, size_t workspace_bytes, void* stream);

/* Synchronises `stream` and returns TM_OK, or TM_ERR_HIP when the last
 * tm_ragged_dt on this workspace aborted a dependency wait. */
int tm_ragged_dt_check(const void* workspace, void* stream);

/* ---------------------------------------------------- grey morphology --- */

/* Workspace bytes for tm_grey (0 is enough for the fused kernels; the per-axis
 * and general fallbacks need it and report TM_ERR_WORKSPACE before launching
 * anything when it is missing). */
size_t tm_grey_workspace_bytes(int64_t batch, int ndim, const int64_t* shape,
                               const int64_t* se_shape);

/* Grey erosion (dilation = 0) / dilation (dilation = 1) of float32 items with
 * the reference's element offsets, border modes (0 constant, 1 reflect,
 * 2 nearest, 3 mirror, 4 wrap) and cval.  structure (float32) and footprint
 * (uint8) are HOST arrays of prod(se_shape) values, either may be NULL
 * (flat / all active).  Replaces torchmorph._C.grey_erosion_cuda /
 * grey_dilation_cuda (reference torchmorph.cpp:3-19, grey_kernel.cu:172-330). */
int tm_grey(const float* input, int64_t batch, int ndim, const int64_t* shape,
            const float* structure, const uint8_t* footprint, const int64_t* se_shape,
            const int64_t* origin, int mode, float cval, int dilation, float* output,
            void* workspace, size_t workspace_bytes, void* stream);

/* 1 when items of `ndim` axes and this structuring element run on the fused
 * streaming kernels (ndim <= 3; a flat box within the LDS budget, or at most
 * 256 active elements with offsets in -128..127), which tm_grey_ex requires;
 * 0 otherwise. */
int tm_grey_fused_ok(int ndim, const float* structure, const uint8_t* footprint,
                     const int64_t* se_shape, const int64_t* origin, int dilation);

/* 1 when gradient / laplace of a flat box of this shape and origin run in one
 * pass (tm_grey_ex epilogues 4, 5): odd sizes, zero origin, ndim <= 3. */
int tm_grey_dual_ok(int ndim, const int64_t* se_shape, const int64_t* origin);

/* tm_grey with the final arithmetic of a composite fused into the store
 * (replaces the reference's separate tensor expressions of
 * morphological_gradient / morphological_laplace / white_tophat /
 * black_tophat, grey.py:200-289; same float32 operations in the same order):
 *   epilogue 1: out = acc - a      gradient (a = erosion), black tophat (a = input)
 *            2: out = a - acc      white tophat (a = input)
 *            3: out = (acc + a) - 2 b   laplace (a = erosion, b = input)
 * where acc is the erosion / dilation of `input`.  a, b: float32 arrays of
 * the output's layout, not aliasing the output.  Needs tm_grey_fused_ok.
 *   epilogue 4: out = dilation - erosion, 5: out = (dilation + erosion) -
 *            2 input, both in one pass over the input (flat box, no structure
 *            or footprint, tm_grey_dual_ok; `dilation`, a and b unused). */
int tm_grey_ex(const float* input, int64_t batch, int ndim, const int64_t* shape,
               const float* structure, const uint8_t* footprint, const int64_t* se_shape,
               const int64_t* origin, int mode, float cval, int dilation, int epilogue,
               const float* a, const float* b, float* output, void* stream);

/* Workspace bytes for tm_grey_backward's general path (0 is enough for 2-D
 * items in modes other than wrap with elements of at most 33 x 33; otherwise
 * TM_ERR_WORKSPACE is reported before anything is launched). */
size_t tm_grey_backward_workspace_bytes(int64_t batch, int ndim, const int64_t* shape);

/* No reference counterpart (the reference ops are not differentiable).
 * Gradient of tm_grey w.r.t. its input: grad_out of every output is routed to
 * the input voxel read by the first active element (element order) attaining
 * the min/max; deterministic, no atomics.  At most 256 active elements. */
int tm_grey_backward(const float* input, int64_t batch, int ndim, const int64_t* shape,
                     const float* structure, const uint8_t* footprint, const int64_t* se_shape,
                     const int64_t* origin, int mode, float cval, int dilation,
                     const float* grad_out, float* grad_in, void* workspace,
                     size_t workspace_bytes, void* stream);

/* Workspace bytes for tm_ragged_grey on these jobs (0 on invalid jobs). */
size_t tm_ragged_grey_workspace_bytes(const tm_ragged_job* jobs, int64_t njobs);

/* No reference counterpart: the reference's grey_erosion / grey_dilation take
 * one dense (B, C, spatial...) tensor (grey.py:29-105, grey_kernel.cu:172-330).
 * Grey erosion / dilation of a ragged list of float32 items in one launch:
 * every job is one item of ndim (1..3, the same for every job and equal to
 * the structuring element's) spatial axes at in_offset / out_offset (float32
 * elements) of the flat input / output; op, sampling and idx_offset are not
 * read.  Structuring element, origin, mode and cval as tm_grey, and each
 * item's result is bit-identical to tm_grey on that item alone.  The element
 * must be a flat box within the LDS budget or have at most 256 active
 * elements with offsets in -128..127 (TM_ERR_ARG otherwise).
 *   upload     1 = copy the item descriptors and the block -> item table into
 *              the workspace (pinned staging, asynchronous); 0 = the
 *              workspace holds them from an earlier call with the same jobs
 *   workspace  256-byte aligned, tm_ragged_grey_workspace_bytes() bytes */
int tm_ragged_grey(const tm_ragged_job* jobs, int64_t njobs, const float* input,
                   const float* structure, const uint8_t* footprint, const int64_t* se_shape,
                   const int64_t* origin, int mode, float cval, int dilation, float* output,
                   int upload, void* workspace, size_t workspace_bytes, void* stream);

/* -------------------------------------------------- binary morphology --- */

/* Workspace bytes for tm_binary (packed ping/pong/orig/mask + flags). */
size_t tm_binary_workspace_bytes(int64_t batch, int ndim, const int64_t* shape);

/* Binary morphology as one bit-packed pipeline: runs nops operations in
 * order, op k being erosion (op_dilation[k] = 0) or dilation (1) with
 * op_iterations[k] iterations (< 1: until an iteration changes nothing,
 * synchronising the host every 8 iterations).  structure: HOST uint8 array of
 * prod(se_shape); border_value for reads outside the item; mask (device
 * uint8, (batch, spatial...), may be NULL) keeps each step's input where it is
 * 0.  input uint8/bool or float32 (nonzero = set); output uint8 (0/1).
 * Replaces torchmorph._C.binary_erosion_cuda / binary_dilation_cuda and the
 * Python iteration loop around them (reference torchmorph.cpp:21-33,
 * binary_kernel.cu:129-250, binary.py:28-88). */
int tm_binary(const void* input, int input_dtype, int64_t batch, int ndim, const int64_t* shape,
              const uint8_t* structure, const int64_t* se_shape, const int64_t* origin,
              int border_value, int nops, const int* op_dilation, const int* op_iterations,
              const uint8_t* mask, uint8_t* output, void* workspace, size_t workspace_bytes,
              void* stream);

/* Replaces the reference's binary_hit_or_miss composition of two
 * binary_erosion calls and a logical_and (reference morphology/binary.py:
 * 188-228): output = erosion(x, structure1, origin1) AND erosion(NOT x,
 * structure2, origin2), border 0 for both, in one packed pipeline (pack x
 * once, the miss step ANDs into the hit words, unpack once).  An all-zero
 * structure1 hits everywhere; an all-zero structure2 misses nowhere.  Shapes
 * and dtypes as tm_binary (any structure size); workspace as
 * tm_binary_workspace_bytes. */
int tm_binary_hit_or_miss(const void* input, int input_dtype, int64_t batch, int ndim,
                          const int64_t* shape, const uint8_t* structure1,
                          const int64_t* se1_shape, const int64_t* origin1,
                          const uint8_t* structure2, const int64_t* se2_shape,
                          const int64_t* origin2, uint8_t* output, void* workspace,
                          size_t workspace_bytes, void* stream);

/* Replaces the reference's binary_fill_holes composition (reference
 * morphology/binary.py:160-185: a propagation of an empty seed into the
 * background with border_value 1, then a logical_not): the background
 * reachable from outside the item by dilation with `structure` (origin),
 * restricted to the background, is removed; everything else (objects and
 * the holes in them) is set.  One packed pipeline: the background mask is
 * packed and inverted once, the propagation runs until a step changes
 * nothing (the host reads a device change flag after chunks of 8, 16, 32,
 * then 64 steps), the inverted result is unpacked once.  Shapes, dtypes and
 * workspace as tm_binary. */
int tm_binary_fill_holes(const void* input, int input_dtype, int64_t batch, int ndim,
                         const int64_t* shape, const uint8_t* structure, const int64_t* se_shape,
                         const int64_t* origin, uint8_t* output, void* workspace,
                         size_t workspace_bytes, void* stream);

/* ----------------------------------------------------------- Sinkhorn --- */

/* Histogram Sinkhorn, n_iter iterations on (nb, d) rows sharing one (d, d)
 * cost.  Replaces torchmorph._C.sinkhorn_logiter (log_space=1: num = log a /
 * log b, M = cost, u/v = log-scalings) and torchmorph._C.sinkhorn_fastiter
 * (log_space=0: num = a / b, M = K = exp(-C/eps), u/v = scalings)
 * (reference torchmorph.cpp:58-77, ot_kernel.cu:188-283).  u and v are
 * updated in place.  dtype TM_F32 or TM_F64.  n_iter >= 100 replays a
 * captured 25-iteration hipGraph (reference optimal_transport.py:34-68);
 * *used_fallback is set when capture failed and plain launches ran. */
int tm_sinkhorn_hist(int dtype, int log_space, const void* num_a, const void* num_b,
                     const void* M, const void* M_T, void* u, void* v, int64_t nb, int d,
                     int n_iter, double epsilon, int* used_fallback, void* stream);

/* Workspace bytes for tm_sinkhorn_hist_solve (a stop word, plus nb x d
 * elements for the threshold check when threshold > 0 and nb x d for the
 * value shares when want_value). */
size_t tm_sinkhorn_hist_workspace_bytes(int dtype, int64_t nb, int d, double threshold,
                                        int want_value);

/* Replaces the reference's SinkhornSolver solve path for GPU tensors:
 * _log_iterations / _scaling_iterations with their fused kernels and
 * _run_fused (reference optimal_transport.py:34-125), and the value and
 * centred potentials of _SinkhornDistance.forward / potentials()
 * (optimal_transport.py:128-150,235-245) without the n x d x d plan.
 * Inputs as tm_sinkhorn_hist (u, v start as 0 (log) or 1 (scaling) and hold
 * the final state on return), plus C / C_T (cost and its transpose, needed
 * when dist is wanted; in the log form they may equal M / M_T).  threshold
 * > 0 stops after iteration i (i % 10 == 0) once max_b sum_i |u_new - u_old|
 * <= threshold, checked on the device (the reference's torch iterations,
 * optimal_transport.py:86-92,118-124); 0 runs max_iter iterations.  Outputs,
 * each optional (NULL): f, g (nb, d) centred potentials eps log u - mean,
 * eps log v - mean; dist (nb,) values <P, C>; *iters_run the iterations run. */
int tm_sinkhorn_hist_solve(int dtype, int log_space, const void* num_a, const void* num_b,
                           const void* M, const void* M_T, const void* C, const void* C_T,
                           void* u, void* v, int64_t nb, int d, int max_iter, double epsilon,
                           double threshold, void* f, void* g, void* dist, int* iters_run,
                           int* used_fallback, void* workspace, size_t workspace_bytes,
                           void* stream);

/* Replaces the plan reconstruction of SinkhornSolver.plan (reference
 * optimal_transport.py:228-233): plan (nb, d, d) = exp(log u_i - C_ij / eps +
 * log v_j) from a solved state (u, v as left by tm_sinkhorn_hist_solve). */
int tm_sinkhorn_hist_plan(int dtype, int log_space, const void* u, const void* v, const void* C,
                          int64_t nb, int d, double epsilon, void* plan, void* stream);

/* Workspace bytes for tm_sinkhorn_points. */
size_t tm_sinkhorn_points_workspace_bytes(int npairs, const int32_t* n, const int32_t* m);

/* No reference counterpart (the reference solves histograms only).  Entropic
 * OT between npairs point-cloud pairs with cost |x - y|^2: x (sum n, dim),
 * y (sum m, dim), weights a (sum n), b (sum m), per-pair eps (host array),
 * n/m per-pair sizes (host arrays, each <= 4096), dim 1..4.  f/g: initial
 * potentials in (zeros = cold start; a warm start sets f, g is recomputed)
 * and the solution out.  Each pair stops at the first state whose L1
 * row-marginal error after a g-step is <= tol, or after max_iter iterations;
 * iters/err (device, per pair) report it.  The loop runs as a hipGraph of
 * `chunk` iterations with a device-side active-pair count read between
 * replays (the call returns when every pair has stopped). */
int tm_sinkhorn_points(const float* x, const float* y, const float* a, const float* b, int dim,
                       int npairs, const int32_t* n, const int32_t* m, const float* eps,
                       float tol, int max_iter, int chunk, float* f, float* g, int32_t* iters,
                       float* err, void* workspace, size_t workspace_bytes, void* stream);

size_t tm_sinkhorn_points_grad_workspace_bytes(int npairs, const int32_t* n, const int32_t* m);

/* No reference counterpart (the reference's backward only returns the
 * potentials as marginal gradients, optimal_transport.py:146-150).  Gradient
 * of every pair's entropic OT value <f,a> + <g,b> w.r.t. its points from the
 * converged potentials (envelope theorem at the fixed point):
 * gx_i = 2 sum_j P_ij (x_i - y_j), gy_j = 2 sum_i P_ij (y_j - x_i), with the
 * plan P formed on the fly (never materialised).  Layouts as for
 * tm_sinkhorn_points. */
int tm_sinkhorn_points_grad(const float* x, const float* y, const float* a, const float* b,
                            int dim, int npairs, const int32_t* n, const int32_t* m,
                            const float* eps, const float* f, const float* g, float* gx,
                            float* gy, void* workspace, size_t workspace_bytes, void* stream);

/* No reference counterpart (the backward of the implicit Sinkhorn adjoint).
 * Products with the plan of every pair's potentials f, g, formed on the fly
 * (never materialised): transpose = 0: out (sum n, k) = P V for V (sum m, k);
 * transpose = 1: out (sum m, k) = P^T V for V (sum n, k); k in 1..8, V and
 * out row-major.  P_ij = a_i b_j exp((f_i + g_j - |x_i - y_j|^2) / eps).
 * Layouts as tm_sinkhorn_points; workspace as tm_sinkhorn_points_grad. */
int tm_sinkhorn_points_planmat(const float* x, const float* y, const float* a, const float* b,
                               int dim, int npairs, const int32_t* n, const int32_t* m,
                               const float* eps, const float* f, const float* g, int transpose,
                               int k, const float* V, float* out, void* workspace,
                               size_t workspace_bytes, void* stream);

/* ----------------------------------------------------------- warm start --- */
/* No reference counterpart.  DeepSets network predicting the dual potential f
 * of every x point of a point-cloud pair (dim 3) from both clouds and eps:
 * per-point MLP 4 -> 128 -> 128 (shared), mean pooling, head
 * [phi(x_i), mean phi(x), mean phi(y), log(eps/s2)] -> 256 -> 256 -> 1,
 * BF16 MFMA GEMMs with fp32 accumulation.  params: flat fp32 vector of
 * tm_warmstart_param_count() values (layout in tm_warmstart.hip).  Pair
 * sizes n, m and eps are HOST arrays. */
int64_t tm_warmstart_param_count(void);
size_t tm_warmstart_workspace_bytes(int npairs, const int32_t* n, const int32_t* m);
/* f: predicted potentials (sum n), in cost units. */
int tm_warmstart_predict(const float* params, const float* x, const float* y, int npairs,
                         const int32_t* n, const int32_t* m, const float* eps, float* f,
                         void* workspace, size_t workspace_bytes, void* stream);
/* Loss (device scalar) and parameter gradient for targets f_target (sum n,
 * cost units); the per-pair loss is weighted by (s2/eps)^loss_alpha. */
int tm_warmstart_grad(const float* params, const float* x, const float* y, int npairs,
                      const int32_t* n, const int32_t* m, const float* eps,
                      const float* f_target, float loss_alpha, float* grad, float* loss,
                      void* workspace, size_t workspace_bytes, void* stream);
/* Adam (bias-corrected, decoupled weight decay) on n parameters. */
int tm_warmstart_adam(float* params, const float* grad, float* m1, float* m2, int64_t n, int step,
                      float lr, float beta1, float beta2, float eps, float weight_decay,
                      void* stream);
/* BF16 x BF16 -> fp32 GEMM with strided operands (the network's building
 * block, exported for tests): A(m,k) = A[m*sam + k*sak], B(k,n) = B[k*sbk +
 * n*sbn], C dense row-major M x N; splitk_ws (may be NULL) enables split-K. */
int tm_mlp_gemm_bf16(const uint16_t* A, int64_t sam, int64_t sak, const uint16_t* B,
                     int64_t sbk, int64_t sbn, float* C, int M, int N, int K, float* splitk_ws,
                     size_t splitk_ws_elems, void* stream);
/* float32 -> bfloat16 (round to nearest even). */
int tm_cast_bf16(const float* x, uint16_t* y, int64_t n, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* TORCHMORPH_HIP_H */
