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

#ifdef __cplusplus
}
#endif

#endif /* TORCHMORPH_HIP_H */
