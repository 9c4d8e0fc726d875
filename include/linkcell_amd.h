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

/* One system of a batched search: its own points, cell, k and outputs
 * (device pointers as in lc_gpu_knearest; out_d2, out_reach and mask may
 * be NULL; cell_hint <= 0 selects the default). */
typedef struct lc_gpu_frame {
  const double *xyz;
  size_t n;
  struct lc_cell cell;
  size_t k;
  const int *mask;
  double cell_hint;
  int *out_nn;
  double *out_d2;
  int *out_reach;
} lc_gpu_frame;

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

/* Many uneven systems (different n, cell, mesh and k) in one pipeline:
 * one launch per kernel for the whole batch. Errors name the frame. */
int lc_gpu_knearest_batch(lc_gpu_workspace *ws, const lc_gpu_frame *frames,
                          size_t n_frames, int wait);

/* Capture the batch into a hipGraph held by the workspace (after one
 * ordinary run that sizes every buffer); lc_gpu_graph_launch replays it on
 * the workspace stream with whatever the frames' buffers hold then. */
int lc_gpu_graph_capture(lc_gpu_workspace *ws, const lc_gpu_frame *frames,
                         size_t n_frames);
int lc_gpu_graph_launch(lc_gpu_workspace *ws, int wait);
int lc_gpu_graph_reset(lc_gpu_workspace *ws);

/* Time the batch under each walk split and keep the fastest for batches of
 * its size bucket (largest k bucket, log2 point count, log2 mean cell
 * occupancy). Results do not depend on the split. */
int lc_gpu_autotune(lc_gpu_workspace *ws, const lc_gpu_frame *frames,
                    size_t n_frames, int *best_dense_min, double *best_ms);

/* Walk split. LC_GPU_ROW_WALK (-1): one 16-lane row per 16-source tile
 * of a home cell. A positive value: home cells with at least that many
 * sources take the wave-per-cell walk, the rest the lane-per-source walk.
 * 0 (the default) uses the autotuned split for the batch's size bucket,
 * else the row walk. LINKCELL_DENSE_MIN sets a positive split at
 * workspace creation; LINKCELL_TPP=1, 16 or 64 selects the all-lane, row
 * or all-wave walk. */
#define LC_GPU_ROW_WALK (-1)
int lc_gpu_set_dense_min(lc_gpu_workspace *ws, int dense_min);

/* Walk schedule. Results are identical under every schedule.
 *   LC_GPU_SCHED_CELL     plain walk (the default)
 *   LC_GPU_SCHED_LEARNED  the workspace's trained predictor gives each home
 *                         cell a k-th-neighbour distance (plain walk until a
 *                         predictor is trained or loaded)
 *   LC_GPU_SCHED_GIVEN    the per-cell distances set with
 *                         lc_gpu_sched_set_prediction
 * The distance caps the walk's pruning threshold from the first tile on; at
 * every shell end a source whose heap is full is exact under the cap, and
 * one that is not is finished by the exact-order walk when the cap could
 * have hidden the reference's stop. The lane-per-source walk also groups
 * its sources by the predicted depth (ceil(distance / cell_min)). */
#define LC_GPU_SCHED_CELL 0
#define LC_GPU_SCHED_LEARNED 1
#define LC_GPU_SCHED_GIVEN 2
int lc_gpu_set_schedule(lc_gpu_workspace *ws, int mode);

/* Predicted k-th-neighbour distance per home cell for LC_GPU_SCHED_GIVEN:
 * a device array of n_cells floats, one per cell of the batch (frames in
 * order, cells in the reference's (cz*ny + cy)*nx + cx numbering). Kept by
 * pointer; it must stay valid while searches use it. */
int lc_gpu_sched_set_prediction(lc_gpu_workspace *ws, const float *kth_dist,
                                size_t n_cells);

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
