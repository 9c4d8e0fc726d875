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

/* The same timings split further, up to n values: fold, scan, scatter,
 * cell sort, walk (all walk kernels), exact walk, total, then the parts of
 * walk -- schedule set-up (learned caps), first pass, retry pass of rows
 * whose cap was too small -- then the deferred-source and retried-row
 * counts. */
int lc_gpu_last_timing_ex(lc_gpu_workspace *ws, double *ms, int n);

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

/* Time the batch under each walk split (the row walk launched per group
 * or persistent, and the wave-per-cell splits) and keep the fastest for
 * batches of its size bucket (largest k bucket, log2 point count, log2
 * mean cell occupancy). Results do not depend on the split. */
int lc_gpu_autotune(lc_gpu_workspace *ws, const lc_gpu_frame *frames,
                    size_t n_frames, int *best_dense_min, double *best_ms);

/* Walk split. LC_GPU_ROW_WALK (-1): one 16-lane row per 16-source tile
 * of a home cell. LC_GPU_ROW_WALK_PERSISTENT (-2): the same walk as one
 * persistent launch per heap bucket, its grid as large as the GPU holds at
 * once and its waves taking groups of four tiles from a device counter.
 * A positive value: home cells with at least that many sources take the
 * wave-per-cell walk, the rest the lane-per-source walk. 0 (the default)
 * uses the autotuned split for the batch's size bucket, else the row walk.
 * LINKCELL_DENSE_MIN sets a positive split at workspace creation;
 * LINKCELL_TPP=1, 16 or 64 selects the all-lane, row or all-wave walk. */
#define LC_GPU_ROW_WALK (-1)
#define LC_GPU_ROW_WALK_PERSISTENT (-2)
int lc_gpu_set_dense_min(lc_gpu_workspace *ws, int dense_min);

/* Walk schedule. Results are identical under every schedule.
 *   LC_GPU_SCHED_CELL     plain walk (the default)
 *   LC_GPU_SCHED_LEARNED  the workspace's trained predictor gives each source
 *                         tile a k-th-neighbour distance (row walk only;
 *                         plain walk until a predictor is trained or set)
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
/* The learned caps plus the depth schedule: each tile's predicted depth
 * R = ceil(cap / cell_min) (at most the wrap-free reach); the row lists the
 * cells of shells 1 .. R in one pass and the exact stop test decides from
 * shell R on; the tiles of every 16 are packed by R so that rows of one
 * wave need about the same number of shells. Where a per-source reach is
 * asked for (lc_gpu_knearest_ex), rows walk shell by shell. */
#define LC_GPU_SCHED_LEARNED_DEPTH 3
int lc_gpu_set_schedule(lc_gpu_workspace *ws, int mode);

/* Predicted k-th-neighbour distance per home cell for LC_GPU_SCHED_GIVEN:
 * a device array of n_cells floats, one per cell of the batch (frames in
 * order, cells in the reference's (cz*ny + cy)*nx + cx numbering). Kept by
 * pointer; it must stay valid while searches use it. */
int lc_gpu_sched_set_prediction(lc_gpu_workspace *ws, const float *kth_dist,
                                size_t n_cells);

/* Per source tile instead of per cell, in the workspace's tile order (16
 * consecutive sources of one home cell in the walk's cell-sorted order, as
 * numbered by lc_gpu_sched_tile_labels); takes precedence over a per-cell
 * prediction until lc_gpu_sched_set_prediction is called again. Tiles at or
 * beyond n_tiles are not capped. A row whose cap was too small is walked
 * again without it, so results never change. */
int lc_gpu_sched_set_tile_prediction(lc_gpu_workspace *ws, const float *kth_dist,
                                     size_t n_tiles);

/* Per source slot instead (the walk's cell-sorted order, as numbered by
 * lc_gpu_sched_source_labels: n_sources floats, one per active source):
 * each source's own threshold cap; takes precedence over the per-tile and
 * per-cell predictions; NULL clears it. Results never change. */
int lc_gpu_sched_set_source_prediction(lc_gpu_workspace *ws, const float *kth_dist,
                                       size_t n_sources);

/* For every active source of the last search in the walk's slot order, its
 * final k-th-neighbour distance rounded up to float (+inf without one): a
 * device array of at least as many floats as active sources. Waits. */
int lc_gpu_sched_source_labels(lc_gpu_workspace *ws, float *out, size_t n);

/* Number of source tiles of the last search (waits for it). */
int lc_gpu_sched_tile_count(lc_gpu_workspace *ws, size_t *n_tiles);

/* For every source tile of the last search, the largest final
 * k-th-neighbour distance among its sources, read back from that search's
 * out_d2 and rounded up to float (+inf for a source with fewer than k
 * neighbours, or without out_d2): a device array of at least
 * lc_gpu_sched_tile_count floats. The outputs of the last search must still
 * be valid. Waits for completion. */
int lc_gpu_sched_tile_labels(lc_gpu_workspace *ws, float *out, size_t n);

/* Device-cell searches (lc_gpu_knearest_many_dcell) invert the cell and plan
 * the mesh on the device, so nothing is read back and the search can be
 * captured into a graph; their buffers are sized for this many cells per
 * frame (default 262144, the reference's device limit), and a plan above it
 * fails with "too many cells". The failure of a call made with wait = 0 is
 * reported by the next lc_gpu_wait (through lc_gpu_last_error); its output
 * rows are -1 / NaN. */
int lc_gpu_set_dcell_capacity(lc_gpu_workspace *ws, size_t cells_per_frame);

/* Diagnostics: the device's plan of a device cell as 36 doubles: H by
 * columns (9), Hinv (9), origin (3), widths (3), ortho, cell_min, nx, ny,
 * nz, cells, max_reach, status (0 planned, 1 bad cell, 2 too many cells),
 * then 4 zeros. Waits. */
int lc_gpu_dcell_plan(lc_gpu_workspace *ws, const double *cell, int cell_n,
                      double cell_hint, double *out36);

/* A search that allocates nothing, for callers that cannot: code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation.
 * No counterpart in the reference. Arguments as lc_gpu_knearest_many, with
 * no workspace: the work goes onto `stream` (a hipStream_t; NULL is the
 * null stream) and is not waited for. Every buffer of the search lives in
 * the library's fixed device scratch of lc_gpu_noalloc_scratch_bytes()
 * (64 MiB) per GPU, held until the process ends; a search that needs more
 * fails with a message that says how much (roughly n_frames * n above
 * 400,000 points). Calls on one GPU share the scratch: calls that overlap
 * on that GPU race. Besides the frames' own inputs and outputs the call
 * makes no device or host allocation on success. */
int lc_gpu_knearest_noalloc(const double *xyz, size_t n, size_t n_frames,
                            const struct lc_cell *simbox, size_t k, const int *mask,
                            double cell_hint, int *out_nn, double *out_d2,
                            void *stream);
size_t lc_gpu_noalloc_scratch_bytes(void);

/* Stage timings (milliseconds) of up to n_runs of the most recent timed
 * searches of this workspace, oldest first, 10 per run: fold, scan,
 * scatter, sort, walk, exact, total, schedule set-up, first walk pass,
 * retry; *got receives the number of runs written (at most 64 are kept).
 * Waits for the latest. */
int lc_gpu_timing_history(lc_gpu_workspace *ws, double *ms, int n_runs, int *got);

/* Diagnostics of the learned scheduler on the last search (host arrays):
 * x_collect (n_tiles x 160 BF16 bit patterns) are the inputs the training
 * set takes for every tile; x_predict (same shape) and caps (n_tiles
 * floats) are the inputs and caps of inference, written only when the
 * predictor is trained (either may be NULL to skip them); tile_cell (home
 * cell per tile), tile_box (8 doubles per tile: lx, hx, ly, hy, lz, hz,
 * pad, pad) and count (occupancy per cell, n_cells ints) are the layout
 * they come from. n_tiles and n_cells must cover the search. Waits. */
int lc_gpu_sched_dump(lc_gpu_workspace *ws, size_t n_tiles, size_t n_cells,
                      unsigned short *x_collect, unsigned short *x_predict,
                      float *caps, int *tile_cell, double *tile_box, int *count);

/* Rows of the last timed search that were walked again without their
 * predicted cap (needs timing on, like lc_gpu_last_timing). */
int lc_gpu_last_retried(lc_gpu_workspace *ws, int *retried);

/* Learned scheduler (LC_GPU_SCHED_LEARNED). Every workspace holds one small
 * MLP (inputs: the home cell's 5x5x5 occupancies and the tile's shape and
 * place; 160 -> 64 -> 64 -> 1, BF16 weights on MFMA, FP32 accumulation) that
 * predicts each source tile's largest final k-th-neighbour distance; the
 * row walk uses prediction * 2^margin as its pruning cap. Training runs on
 * the device with the library's own forward, backward and Adam kernels.
 *
 * lc_gpu_sched_collect appends the tiles of the last search (which must
 * have had out_d2) to the training set, labelled by that search's own
 * k-th distances; lc_gpu_sched_train initialises the network from seed and
 * runs epochs of minibatch Adam (batch a multiple of 64), then sets the
 * margin to the given quantile of (label - prediction) in log2 units. */
#define LC_GPU_SCHED_PARAMS 14529
int lc_gpu_sched_collect(lc_gpu_workspace *ws, size_t *n_added);
int lc_gpu_sched_samples(lc_gpu_workspace *ws, size_t *n);
int lc_gpu_sched_clear(lc_gpu_workspace *ws);
int lc_gpu_sched_train(lc_gpu_workspace *ws, int epochs, int batch, double lr,
                       unsigned seed, double quantile, double *last_loss,
                       double *margin);

/* Network parameters as LC_GPU_SCHED_PARAMS host floats: W1 (160 x 64,
 * input-major), b1 (64), W2 (64 x 64, input-major), b2 (64), w3 (64), b3;
 * set_params also sets the margin and marks the predictor trained. */
int lc_gpu_sched_get_params(lc_gpu_workspace *ws, float *params, size_t n,
                            double *margin);
int lc_gpu_sched_set_params(lc_gpu_workspace *ws, const float *params,
                            size_t n, double margin);

/* The network on host data (inputs rounded to BF16 on the device):
 * forward: y for n samples of 160 inputs; gradient: of the loss
 * sum_i w_i (y_i - t_i)^2 / (2 sum_i w_i) at the current parameters, in the
 * parameter layout above, plus the loss; adam: `steps` full-batch Adam steps
 * (beta 0.9 / 0.999, eps 1e-8, moments from zero), returning the loss
 * before the last step. */
int lc_gpu_sched_forward(lc_gpu_workspace *ws, const float *x, size_t n,
                         float *y);
int lc_gpu_sched_gradient(lc_gpu_workspace *ws, const float *x, const float *t,
                          const float *w, size_t n, float *grad, double *loss);
int lc_gpu_sched_adam(lc_gpu_workspace *ws, const float *x, const float *t,
                      const float *w, size_t n, int steps, double lr,
                      double *loss);

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
