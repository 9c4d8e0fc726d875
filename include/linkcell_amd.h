/*
 * MI355X-build extensions beside the drop-in waist. None of these has a
 * counterpart in the reference; existing callers need none of them.
 */
#ifndef LINKCELL_AMD_H
#define LINKCELL_AMD_H

#pragma once

#include "linkcell_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Walk strategy of a workspace. */
enum {
  LC_GPU_MODE_AUTO = 0,       /* canonical fast walk + exact-order rest */
  LC_GPU_MODE_EXACT_ONLY = 1  /* every source through the exact-order walk */
};

/* lc_gpu_knearest_many plus out_reach: a device buffer of n_frames * n
 * ints (or NULL) that receives each source's final Chebyshev shell, 0 for
 * masked sources. */
int lc_gpu_knearest_ex(lc_gpu_workspace *ws, const double *xyz, size_t n,
                       size_t n_frames, const struct lc_cell *simbox, size_t k,
                       const int *mask, double cell_hint, int *out_nn,
                       double *out_d2, int *out_reach, int wait);

int lc_gpu_set_mode(lc_gpu_workspace *ws, int mode);

/* Per-kernel hipEvent timing on the workspace stream. When on, each
 * search records events around its kernels; lc_gpu_last_timing waits for
 * the last search and returns milliseconds for fold, scan, scatter,
 * cell sort, fast walk, exact walk and the whole pipeline (7 doubles),
 * plus the number of sources the exact-order walk finished. */
int lc_gpu_set_timing(lc_gpu_workspace *ws, int on);
int lc_gpu_last_timing(lc_gpu_workspace *ws, double *ms7, int *deferred);

/* Home-cell occupancy from which the wave-per-cell walk handles a cell
 * (the rest take the lane-per-source walk). Default 32, or the
 * LINKCELL_DENSE_MIN environment variable. */
int lc_gpu_set_dense_min(lc_gpu_workspace *ws, int dense_min);

/* Device work counters of the walk kernels (diagnostics). When on, each
 * search accumulates them; lc_gpu_last_stats waits for the last search
 * and copies up to n counters: dense waves, dense tile tests, dense tiles
 * walked, dense candidate steps, dense heap merge steps, dense appends,
 * sparse sources, sparse tile tests, sparse tiles walked, sparse
 * candidate evaluations, sparse batch merges. */
int lc_gpu_set_stats(lc_gpu_workspace *ws, int on);
int lc_gpu_last_stats(lc_gpu_workspace *ws, unsigned long long *out, int n);

#ifdef __cplusplus
}
#endif

#endif
