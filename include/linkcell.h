/*
 * linkcell C API (MI355X build).
 *
 * Same declarations and contract as the reference header
 * (include/linkcell.h:37-230 of linkcell 0.3.3, generated there from
 * src/capi.rs:49-297). Memory ownership and thread safety are unchanged:
 *   - lc_knearest writes caller-owned host out_nn (n * k ints), unused -1;
 *   - lc_knearest_d2 also writes out_d2 (n * k doubles), unused NaN;
 *   - lc_knearest_many is the frame-major batch (n_frames * n * k);
 *   - lc_last_error is a thread-local slot, NULL after a success;
 *   - lc_version is process-static and leaves the error slot alone.
 *
 * In this library the search itself runs on the GPU (the host buffers are
 * staged through device memory); results are bit-identical to the
 * reference CPU walk. A process without a usable HIP device gets a
 * nonzero return and the message "HIP device not available".
 */
#ifndef LINKCELL_H
#define LINKCELL_H

#pragma once

#include <stddef.h>
#include <stdint.h>

/* Periodic parallelepiped: lattice vectors a, b, c (rows, as vesin) and
 * the dump-cell origin. Reference: src/capi.rs:49-77. */
typedef struct lc_cell {
  double ax, ay, az;
  double bx, by, bz;
  double cx, cy, cz;
  double ox, oy, oz;
} lc_cell;

#ifdef __cplusplus
extern "C" {
#endif

/* k nearest neighbours of n points (src/capi.rs:132-186). */
int lc_knearest(const double *xyz, size_t n, const struct lc_cell *simbox,
                size_t k, const int *mask, double cell_hint, int *out_nn);

/* lc_knearest plus squared distances (src/capi.rs:188-209). */
int lc_knearest_d2(const double *xyz, size_t n, const struct lc_cell *simbox,
                   size_t k, const int *mask, double cell_hint, int *out_nn,
                   double *out_d2);

/* Frame-major batch sharing one cell; mask is length n
 * (src/capi.rs:211-297). */
int lc_knearest_many(const double *xyz, size_t n, size_t n_frames,
                     const struct lc_cell *simbox, size_t k, const int *mask,
                     double cell_hint, int *out_nn, double *out_d2);

/* Thread-local last error (src/capi.rs:108-124). */
const char *lc_last_error(void);

/* Version string (src/capi.rs:126-130). */
const char *lc_version(void);

#ifdef __cplusplus
}
#endif

#endif /* LINKCELL_H */

#ifndef LINKCELL_LC_CELL_ORTHO_H
#define LINKCELL_LC_CELL_ORTHO_H

/* Orthorhombic cell with origin at zero (include/linkcell.h:206-224). */
static inline lc_cell lc_cell_ortho(double lx, double ly, double lz) {
  lc_cell c = {lx, 0.0, 0.0, 0.0, ly, 0.0, 0.0, 0.0, lz, 0.0, 0.0, 0.0};
  return c;
}

#endif /* LINKCELL_LC_CELL_ORTHO_H */
