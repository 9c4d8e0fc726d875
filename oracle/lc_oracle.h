/*
 * CPU oracle for the periodic linked-cell k-nearest search.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP
 * product path is compared against; it is also the CPU baseline timed
 * beside the GPU. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it. The product library never links it.
 *
 * It restates linkcell 0.3.3's Rust CPU path (src/cell.rs,
 * src/knearest.rs, src/capi.rs, src/error.rs) in plain C with the same
 * IEEE-754 operation order and no FMA contraction (built with
 * -ffp-contract=off). Each function cites the file:line it follows.
 *
 * Parity pins (see tests/test_oracle.py):
 *   cell construction  - closed forms on power-of-two boxes, BadBox
 *                        cases (tests/agree.rs:18-27), ortho flag
 *                        (src/cell.rs:395-421), widths doc test.
 *   fold               - doc-test known answers (src/cell.rs:161-174),
 *                        0 <= s < 1 and primary-cell invariants.
 *   shift / dist2      - lattice_shift(-1,0,0) == [-10,0,0] exactly,
 *                        MIC 0.64 known answers.
 *   walk + heap        - independent 27-image brute force (numpy,
 *                        bitwise d2) and the restated knearest_brute on
 *                        every reference known-answer and LCG-seeded case
 *                        (tests/agree.rs:30-558, tests/abi.rs:83-367).
 *   boundary ties      - hand-derived answers for exact d2 ties across
 *                        cells, inside a cell, across evictions and
 *                        shells (tests/golden/tie_answers.json), which
 *                        builds with a planted mistake (LC_ORACLE_PLANT)
 *                        fail.
 * Host sanitizers: tests/sanitize/oracle_check.c runs it under ASan,
 * UBSan and TSan.
 */
#ifndef LC_ORACLE_H
#define LC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes follow src/error.rs:7-22; messages src/error.rs:27-33. */
enum {
  ORC_OK = 0,
  ORC_ZERO_K = 1,
  ORC_BAD_BOX = 2,
  ORC_EMPTY = 3,
  ORC_BUFFER_SIZE = 4,
  ORC_OVERFLOW = 5,
  ORC_MASK_LEN = 6,
  ORC_TOO_MANY_CELLS = 7,
  ORC_NULL = 8
};

/* Cell (src/cell.rs:35-46). h[col][row]: h[0] is lattice vector a. */
typedef struct orc_cell {
  double h[3][3];
  double hinv[3][3];
  double origin[3];
  double widths[3];
  int ortho;
} orc_cell;

/* Mesh plan (src/knearest.rs:377-397, :430). */
typedef struct orc_plan {
  int nx, ny, nz;
  int64_t ncell;
  double cell_min;
  int max_reach;
} orc_plan;

const char *orc_error_message(int code);
const char *orc_version(void);

int orc_cell_from_vectors(const double a[3], const double b[3],
                          const double c[3], const double origin[3],
                          orc_cell *out);
void orc_fractional(const orc_cell *cell, const double r[3], double s[3]);
void orc_cartesian(const orc_cell *cell, const double s[3], double r[3]);
void orc_lattice_shift(const orc_cell *cell, int na, int nb, int nc,
                       double out[3]);
double orc_dist2_shifted(const double p[3], const double q[3],
                         const double shift[3]);
double orc_dist2(const orc_cell *cell, const double p[3], const double q[3]);

/* hint: the value handed to search() (NaN or <= 0 means default 3.0). */
int orc_plan_mesh(const orc_cell *cell, double hint, orc_plan *out);

/* Fold + bin of one point: folded Cartesian image and bin triple. */
void orc_fold_bin(const orc_cell *cell, const orc_plan *plan,
                  const double r[3], double folded[3], int bin[3]);

/*
 * lc_knearest_many restated (src/capi.rs:222-297 validation order, then
 * src/knearest.rs:225-307 + :351-486). cell12 is the lc_cell field order
 * (a, b, c, origin). out_d2 may be NULL. out_reach / out_cand, when not
 * NULL, receive per-source final reach and candidates visited (0 for
 * inactive sources); they are diagnostics with no reference counterpart.
 * nthreads <= 1 runs serially; otherwise sources are split across
 * pthreads (the rayon par_iter equivalent, src/knearest.rs:477-481).
 */
int orc_knearest_many(const double *xyz, size_t n, size_t n_frames,
                      const double *cell12, size_t k, const int *mask,
                      double cell_hint, int *out_nn, double *out_d2,
                      int *out_reach, int *out_cand, int nthreads);

/*
 * Walk only the listed sources of one frame (the CPU baseline's bounded
 * sample). Rows of unlisted sources are left untouched. Times in seconds
 * for the preparation (fold/bin/lists over all points) and the walk are
 * written to times[0], times[1] when times is not NULL.
 */
int orc_knearest_sample(const double *xyz, size_t n, const double *cell12,
                        size_t k, const int *mask, double cell_hint,
                        const int *sources, size_t n_sources, int *out_nn,
                        double *out_d2, int *out_reach, int *out_cand,
                        int nthreads, double *times);

/* knearest_brute restated (src/knearest.rs:24-40, :508-541). */
int orc_knearest_brute(const double *xyz, size_t n, const double *cell12,
                       size_t k, const int *mask, int *out_nn,
                       double *out_d2);

#ifdef __cplusplus
}
#endif

#endif
