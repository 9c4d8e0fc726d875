/*
 * Device-resident k-nearest search (MI355X / HIP build).
 *
 * Drop-in for the reference GPU waist (include/linkcell_gpu.h:28-57,
 * implemented at src/gpu/capi.cpp:19-284): same types, names and
 * signatures. Pointers named xyz, mask, out_nn, out_d2 and the dcell
 * `cell` are HIP device pointers; simbox is a host lc_cell. The queue is a
 * hipStream_t. Returns are 0 on success and 1 on failure with a
 * thread-local message from lc_gpu_last_error (NULL after a success).
 *
 * Differences from the reference GPU path, all toward the reference CPU
 * contract: results are bit-identical to the CPU walk (indices, d2 bits,
 * tie order, -1 / NaN fill), any k >= 1 and up to 16,777,216 cells are
 * accepted, and the mask is length n and shared by all frames
 * (docs/reference/c-api.md:46).
 *
 * Distinct workspaces may run concurrently. Each owns one HIP stream;
 * call lc_gpu_wait before reading results of a search made with wait=0.
 */
#ifndef LINKCELL_GPU_H
#define LINKCELL_GPU_H

#pragma once

#include "linkcell.h"

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lc_gpu_workspace lc_gpu_workspace;

/* 1 when a HIP device is present (src/gpu/capi.cpp:21-27). */
int lc_gpu_available(void);
/* thread-local last error (src/gpu/capi.cpp:29-31). */
const char *lc_gpu_last_error(void);

/* src/gpu/capi.cpp:39-57 */
lc_gpu_workspace *lc_gpu_workspace_new(void);
void lc_gpu_workspace_free(lc_gpu_workspace *ws);

/* one frame, waits (src/gpu/capi.cpp:65-70) */
int lc_gpu_knearest(lc_gpu_workspace *ws, const double *xyz, size_t n,
                    const struct lc_cell *simbox, size_t k, const int *mask,
                    double cell_hint, int *out_nn, double *out_d2);

/* frame-major batch sharing one cell (src/gpu/capi.cpp:72-108) */
int lc_gpu_knearest_many(lc_gpu_workspace *ws, const double *xyz, size_t n,
                         size_t n_frames, const struct lc_cell *simbox,
                         size_t k, const int *mask, double cell_hint,
                         int *out_nn, double *out_d2, int wait);

/* device-packed cell: 3 ortho lengths, 9 rows a b c, or 12 with the
 * origin (src/gpu/capi.cpp:110-148) */
int lc_gpu_knearest_many_dcell(lc_gpu_workspace *ws, const double *xyz,
                               size_t n, size_t n_frames, const double *cell,
                               int cell_n, size_t k, const int *mask,
                               double cell_hint, int *out_nn, double *out_d2,
                               int wait);

/* the workspace stream as a hipStream_t (src/gpu/capi.cpp:150-167) */
void *lc_gpu_queue(lc_gpu_workspace *ws);
void lc_gpu_wait(lc_gpu_workspace *ws);

/* device memory helpers (src/gpu/capi.cpp:169-282); kind follows
 * hipMemcpyKind, numbered like cudaMemcpyKind (2 = device to host) */
int lc_gpu_alloc(void **ptr, size_t bytes);
void lc_gpu_free(void *ptr);
int lc_gpu_fill_i32(void *ptr, int value, size_t n);
int lc_gpu_memcpy(void *dst, const void *src, size_t bytes, int kind);

#ifdef __cplusplus
}
#endif

#endif
