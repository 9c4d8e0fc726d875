/*
 * CPU oracle: C restatement of linkcell 0.3.3's Rust CPU walk.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline); the
 * product library never links this file. Build with -O2
 * -ffp-contract=off and no fast-math: every floating-point expression
 * below is evaluated in the reference's source order with correctly
 * rounded IEEE-754 double operations, which is what rustc emits for the
 * reference (no target-cpu, so no FMA; Cargo.toml:35-37).
 */
#define _GNU_SOURCE
#include "lc_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#define MAX_CELLS 16777216LL /* src/knearest.rs:20 */

const char *orc_version(void) { return "0.3.3"; }

/* src/error.rs:27-33 plus the C ABI's "null pointer" (src/capi.rs:240). */
const char *orc_error_message(int code) {
  switch (code) {
  case ORC_ZERO_K: return "k must be at least 1";
  case ORC_BAD_BOX: return "singular or non-positive cell";
  case ORC_EMPTY: return "no points";
  case ORC_BUFFER_SIZE: return "out buffer length must be n * k";
  case ORC_OVERFLOW: return "n * k overflows";
  case ORC_MASK_LEN: return "mask length must be n";
  case ORC_TOO_MANY_CELLS: return "linked-cell mesh is too fine";
  case ORC_NULL: return "null pointer";
  default: return NULL;
  }
}

/* Rust f64::min / f64::max: a NaN operand yields the other operand. */
static double rmin(double a, double b) { return fmin(a, b); }
static double rmax(double a, double b) { return fmax(a, b); }

/* Rust `x as i32`: saturating, NaN -> 0. */
static int32_t as_i32(double x) {
  if (x != x) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return (int32_t)-2147483647 - 1;
  return (int32_t)x;
}

static int32_t rem_euclid(int32_t a, int32_t n) {
  int32_t r = a % n;
  return r < 0 ? r + n : r;
}

static int32_t div_euclid(int32_t a, int32_t n) {
  int32_t q = a / n;
  if (a % n < 0) q -= 1;
  return q;
}

/* src/cell.rs:335-341: H stored by columns, mul(h, v) = H v. */
static void mul(const double h[3][3], const double v[3], double out[3]) {
  double r0 = h[0][0] * v[0] + h[1][0] * v[1] + h[2][0] * v[2];
  double r1 = h[0][1] * v[0] + h[1][1] * v[1] + h[2][1] * v[2];
  double r2 = h[0][2] * v[0] + h[1][2] * v[1] + h[2][2] * v[2];
  out[0] = r0;
  out[1] = r1;
  out[2] = r2;
}

/* src/cell.rs:343-349 */
static void cross(const double u[3], const double v[3], double out[3]) {
  out[0] = u[1] * v[2] - u[2] * v[1];
  out[1] = u[2] * v[0] - u[0] * v[2];
  out[2] = u[0] * v[1] - u[1] * v[0];
}

/* src/cell.rs:351-353 */
static double norm(const double v[3]) {
  return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
}

/* src/cell.rs:322-331 */
static int is_axis_aligned(const double h[3][3]) {
  double scale = rmax(norm(h[0]) + norm(h[1]) + norm(h[2]), 1.0);
  double tol = 1e-12 * scale;
  return fabs(h[0][1]) <= tol && fabs(h[0][2]) <= tol &&
         fabs(h[1][0]) <= tol && fabs(h[1][2]) <= tol &&
         fabs(h[2][0]) <= tol && fabs(h[2][1]) <= tol;
}

/* src/cell.rs:355-389. Returns 0 when H is singular or non-finite. */
static int invert_columns(const double h[3][3], double inv[3][3],
                          double *det_out) {
  const double *a = h[0], *b = h[1], *c = h[2];
  double det = a[0] * (b[1] * c[2] - b[2] * c[1]) -
               a[1] * (b[0] * c[2] - b[2] * c[0]) +
               a[2] * (b[0] * c[1] - b[1] * c[0]);
  if (!isfinite(det) || fabs(det) < 1e-18) return 0;
  double invdet = 1.0 / det;
  inv[0][0] = (b[1] * c[2] - b[2] * c[1]) * invdet;
  inv[0][1] = (a[2] * c[1] - a[1] * c[2]) * invdet;
  inv[0][2] = (a[1] * b[2] - a[2] * b[1]) * invdet;
  inv[1][0] = (b[2] * c[0] - b[0] * c[2]) * invdet;
  inv[1][1] = (a[0] * c[2] - a[2] * c[0]) * invdet;
  inv[1][2] = (a[2] * b[0] - a[0] * b[2]) * invdet;
  inv[2][0] = (b[0] * c[1] - b[1] * c[0]) * invdet;
  inv[2][1] = (a[1] * c[0] - a[0] * c[1]) * invdet;
  inv[2][2] = (a[0] * b[1] - a[1] * b[0]) * invdet;
  *det_out = det;
  return 1;
}

/* Cell::from_vectors, src/cell.rs:97-123 */
int orc_cell_from_vectors(const double a[3], const double b[3],
                          const double c[3], const double origin[3],
                          orc_cell *out) {
  orc_cell cell;
  memset(&cell, 0, sizeof cell);
  for (int i = 0; i < 3; ++i) {
    cell.h[0][i] = a[i];
    cell.h[1][i] = b[i];
    cell.h[2][i] = c[i];
    cell.origin[i] = origin[i];
  }
  double det;
  if (!invert_columns(cell.h, cell.hinv, &det)) return ORC_BAD_BOX;
  if (!isfinite(det) || fabs(det) < 1e-18) return ORC_BAD_BOX;
  double bc[3], ca[3], ab[3];
  cross(b, c, bc);
  cross(c, a, ca);
  cross(a, b, ab);
  double wa = fabs(det) / norm(bc);
  double wb = fabs(det) / norm(ca);
  double wc = fabs(det) / norm(ab);
  if (!(wa > 0.0 && wb > 0.0 && wc > 0.0)) return ORC_BAD_BOX;
  cell.widths[0] = wa;
  cell.widths[1] = wb;
  cell.widths[2] = wc;
  cell.ortho = is_axis_aligned(cell.h);
  *out = cell;
  return ORC_OK;
}

/* src/cell.rs:291-298 */
static double wrap01(double s) {
  s -= floor(s);
  return s >= 1.0 ? 0.0 : s;
}

/* Cell::fractional, src/cell.rs:176-194 */
void orc_fractional(const orc_cell *cell, const double r[3], double s[3]) {
  if (cell->ortho) {
    s[0] = wrap01((r[0] - cell->origin[0]) / cell->widths[0]);
    s[1] = wrap01((r[1] - cell->origin[1]) / cell->widths[1]);
    s[2] = wrap01((r[2] - cell->origin[2]) / cell->widths[2]);
    return;
  }
  double d[3] = {r[0] - cell->origin[0], r[1] - cell->origin[1],
                 r[2] - cell->origin[2]};
  mul(cell->hinv, d, s);
  for (int i = 0; i < 3; ++i) s[i] = wrap01(s[i]);
}

/* Cell::cartesian, src/cell.rs:200-207 */
void orc_cartesian(const orc_cell *cell, const double s[3], double r[3]) {
  double t[3];
  mul(cell->h, s, t);
  r[0] = t[0] + cell->origin[0];
  r[1] = t[1] + cell->origin[1];
  r[2] = t[2] + cell->origin[2];
}

/* Cell::lattice_shift, src/cell.rs:227-247 */
void orc_lattice_shift(const orc_cell *cell, int na, int nb, int nc,
                       double out[3]) {
  if (cell->ortho) {
    out[0] = (double)na * cell->widths[0];
    out[1] = (double)nb * cell->widths[1];
    out[2] = (double)nc * cell->widths[2];
    return;
  }
  const double *a = cell->h[0], *b = cell->h[1], *c = cell->h[2];
  double fa = (double)na, fb = (double)nb, fc = (double)nc;
  out[0] = fa * a[0] + fb * b[0] + fc * c[0];
  out[1] = fa * a[1] + fb * b[1] + fc * c[1];
  out[2] = fa * a[2] + fb * b[2] + fc * c[2];
}

/* Cell::dist2_shifted, src/cell.rs:269-274 */
double orc_dist2_shifted(const double p[3], const double q[3],
                         const double shift[3]) {
  double dx = q[0] + shift[0] - p[0];
  double dy = q[1] + shift[1] - p[1];
  double dz = q[2] + shift[2] - p[2];
  return dx * dx + dy * dy + dz * dz;
}

/* dist2_ortho, src/cell.rs:300-309 */
static double dist2_ortho(const double l[3], const double p[3],
                          const double q[3]) {
  double r2 = 0.0;
  for (int k = 0; k < 3; ++k) {
    double d = fabs(q[k] - p[k]);
    d -= l[k] * round(d / l[k]);
    r2 += d * d;
  }
  return r2;
}

/* dist2_general, src/cell.rs:311-320 */
static double dist2_general(const double h[3][3], const double hinv[3][3],
                            const double p[3], const double q[3]) {
  double dp[3] = {q[0] - p[0], q[1] - p[1], q[2] - p[2]};
  double ds[3], dr[3];
  mul(hinv, dp, ds);
  for (int i = 0; i < 3; ++i) ds[i] -= round(ds[i]);
  mul(h, ds, dr);
  return dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
}

/* Cell::dist2, src/cell.rs:282-287 */
double orc_dist2(const orc_cell *cell, const double p[3], const double q[3]) {
  if (cell->ortho) return dist2_ortho(cell->widths, p, q);
  return dist2_general(cell->h, cell->hinv, p, q);
}

/* bins_1d, src/knearest.rs:42-48 */
static int bins_1d(double width, double edge, int *out) {
  double n = rmax(floor(width / edge), 1.0);
  if (!isfinite(n) || n > 1000000.0) return ORC_TOO_MANY_CELLS;
  *out = as_i32(n);
  return ORC_OK;
}

/* mesh plan, src/knearest.rs:377-397 and :430 */
int orc_plan_mesh(const orc_cell *cell, double hint, orc_plan *out) {
  double edge = hint;
  if (!isfinite(edge) || edge <= 0.0) edge = 3.0;
  const double *w = cell->widths;
  edge = rmin(rmin(rmin(edge, w[0]), w[1]), w[2]);
  int nx, ny, nz, rc;
  if ((rc = bins_1d(w[0], edge, &nx))) return rc;
  if ((rc = bins_1d(w[1], edge, &ny))) return rc;
  if ((rc = bins_1d(w[2], edge, &nz))) return rc;
  int64_t ncell = (int64_t)nx * (int64_t)ny * (int64_t)nz;
  if (!(ncell > 0 && ncell <= MAX_CELLS)) return ORC_TOO_MANY_CELLS;
  out->nx = nx;
  out->ny = ny;
  out->nz = nz;
  out->ncell = ncell;
  out->cell_min = rmin(rmin(w[0] / (double)nx, w[1] / (double)ny),
                       w[2] / (double)nz);
  int m = nx > ny ? nx : ny;
  m = m > nz ? m : nz;
  out->max_reach = m / 2 + 1;
  return ORC_OK;
}

/* fold + bin of one point, src/knearest.rs:405-412 */
void orc_fold_bin(const orc_cell *cell, const orc_plan *plan,
                  const double r[3], double folded[3], int bin[3]) {
  double s[3];
  orc_fractional(cell, r, s);
  orc_cartesian(cell, s, folded);
  int n[3] = {plan->nx, plan->ny, plan->nz};
  double inv[3] = {(double)plan->nx, (double)plan->ny, (double)plan->nz};
  for (int a = 0; a < 3; ++a) {
    int v = as_i32(s[a] * inv[a]);
    if (v < 0) v = 0;
    if (v > n[a] - 1) v = n[a] - 1;
    bin[a] = v;
  }
}

/* ---------------- KHeap, src/knearest.rs:53-159 ---------------- */

typedef struct kheap {
  double *d2;
  size_t *idx;
  size_t n, k;
} kheap;

/* Planted mistakes, for tests/test_oracle.py only: each variant is a
 * plausible misreading of KHeap's tie rules, compiled into a separate
 * library by the test to show that the hand-derived tie fixtures
 * (tests/golden/tie_answers.json) reject it. The oracle itself is built
 * with LC_ORACLE_PLANT 0 (none). */
#ifndef LC_ORACLE_PLANT
#define LC_ORACLE_PLANT 0
#endif
#define LC_PLANT_REPLACE_ON_EQUAL 1   /* d2 <= worst replaces */
#define LC_PLANT_LAST_ARGMAX 2        /* the last of equal maxima is the worst */
#define LC_PLANT_CANONICAL_INDEX 3    /* ties retained by smallest index */
#define LC_PLANT_INCREASING_LISTS 4   /* cell lists in increasing index */
#define LC_PLANT_FINISH_DESCENDING 5  /* finish orders equal d2 by larger index */

/* KHeap::push, src/knearest.rs:106-129 */
static void kheap_push(kheap *h, double d2, size_t j) {
  for (size_t t = 0; t < h->n; ++t) {
    if (h->idx[t] == j) {
      if (d2 < h->d2[t]) {
        h->d2[t] = d2;
        h->idx[t] = j;
      }
      return;
    }
  }
  if (h->n < h->k) {
    h->d2[h->n] = d2;
    h->idx[h->n] = j;
    h->n += 1;
    return;
  }
  size_t worst = 0;
  for (size_t t = 1; t < h->n; ++t) {
#if LC_ORACLE_PLANT == LC_PLANT_LAST_ARGMAX
    if (h->d2[t] >= h->d2[worst]) worst = t;
#elif LC_ORACLE_PLANT == LC_PLANT_CANONICAL_INDEX
    if (h->d2[t] > h->d2[worst] ||
        (h->d2[t] == h->d2[worst] && h->idx[t] > h->idx[worst]))
      worst = t;
#else
    if (h->d2[t] > h->d2[worst]) worst = t;
#endif
  }
#if LC_ORACLE_PLANT == LC_PLANT_REPLACE_ON_EQUAL
  if (d2 <= h->d2[worst]) {
#elif LC_ORACLE_PLANT == LC_PLANT_CANONICAL_INDEX
  if (d2 < h->d2[worst] || (d2 == h->d2[worst] && j < h->idx[worst])) {
#else
  if (d2 < h->d2[worst]) {
#endif
    h->d2[worst] = d2;
    h->idx[worst] = j;
  }
}

/* KHeap::worst, src/knearest.rs:135-144 */
static double kheap_worst(const kheap *h) {
  double w = h->d2[0];
  for (size_t t = 1; t < h->n; ++t) {
    double v = h->d2[t];
    if (v > w) w = v;
  }
  return w;
}

/* f64::total_cmp key */
static int64_t total_key(double x) {
  int64_t b;
  memcpy(&b, &x, sizeof b);
  return b ^ (int64_t)((uint64_t)(b >> 63) >> 1);
}

/* KHeap::finish sort order, src/knearest.rs:146-158 */
static void kheap_finish(kheap *h) {
  /* insertion sort: n <= k and k is small in practice */
  for (size_t a = 1; a < h->n; ++a) {
    double dv = h->d2[a];
    size_t jv = h->idx[a];
    int64_t kv = total_key(dv);
    size_t p = a;
    while (p > 0) {
      int64_t kp = total_key(h->d2[p - 1]);
#if LC_ORACLE_PLANT == LC_PLANT_FINISH_DESCENDING
      if (kp > kv || (kp == kv && h->idx[p - 1] < jv)) {
#else
      if (kp > kv || (kp == kv && h->idx[p - 1] > jv)) {
#endif
        h->d2[p] = h->d2[p - 1];
        h->idx[p] = h->idx[p - 1];
        --p;
      } else {
        break;
      }
    }
    h->d2[p] = dv;
    h->idx[p] = jv;
  }
}

/* ---------------- search, src/knearest.rs:351-486 ---------------- */

typedef struct frame_ctx {
  const orc_cell *cell;
  orc_plan plan;
  size_t n, k;
  double (*folded)[3];
  int (*bin)[3];
  int64_t *head; /* per cell, -1 terminated */
  int64_t *next; /* per point */
  int *out_nn;
  double *out_d2;
  int *out_reach;
  int *out_cand;
} frame_ctx;

static int64_t cell_index(const frame_ctx *f, int ix, int iy, int iz) {
  int cx = rem_euclid(ix, f->plan.nx);
  int cy = rem_euclid(iy, f->plan.ny);
  int cz = rem_euclid(iz, f->plan.nz);
  return ((int64_t)cz * f->plan.ny + cy) * f->plan.nx + cx;
}

/* the closure `one(i)`, src/knearest.rs:431-475, then write_one's
 * packing (src/knearest.rs:298-305). */
static void walk_one(const frame_ctx *f, size_t i, double *hd2, size_t *hidx) {
  kheap heap = {hd2, hidx, 0, f->k};
  const int ix = f->bin[i][0], iy = f->bin[i][1], iz = f->bin[i][2];
  const int max_reach = f->plan.max_reach;
  int reach = 1;
  int64_t cand = 0;
  int final_reach = max_reach;
  while (reach <= max_reach) {
    for (int dx = -reach; dx <= reach; ++dx) {
      for (int dy = -reach; dy <= reach; ++dy) {
        for (int dz = -reach; dz <= reach; ++dz) {
          int shell = reach == 1 || abs(dx) == reach || abs(dy) == reach ||
                      abs(dz) == reach;
          if (!shell && reach > 1) continue;
          int jx = ix + dx, jy = iy + dy, jz = iz + dz;
          int64_t c = cell_index(f, jx, jy, jz);
          double shift[3];
          orc_lattice_shift(f->cell, div_euclid(jx, f->plan.nx),
                            div_euclid(jy, f->plan.ny),
                            div_euclid(jz, f->plan.nz), shift);
          int64_t j = f->head[c];
          while (j >= 0) {
            if ((size_t)j != i) {
              kheap_push(&heap, orc_dist2_shifted(f->folded[i], f->folded[j],
                                                  shift),
                         (size_t)j);
              cand += 1;
            }
            j = f->next[j];
          }
        }
      }
    }
    if (heap.n >= heap.k) {
      double bound = (double)reach * f->plan.cell_min;
      if (kheap_worst(&heap) <= bound * bound) {
        final_reach = reach;
        break;
      }
    }
    reach += 1;
  }
  kheap_finish(&heap);
  for (size_t t = 0; t < heap.n; ++t) {
    f->out_nn[i * f->k + t] = (int)heap.idx[t];
    if (f->out_d2) f->out_d2[i * f->k + t] = heap.d2[t];
  }
  if (f->out_reach) f->out_reach[i] = final_reach;
  if (f->out_cand) f->out_cand[i] = (int)(cand > 2147483647 ? 2147483647 : cand);
}

typedef struct worker_arg {
  const frame_ctx *f;
  const size_t *list;
  size_t count;
  atomic_size_t *cursor;
} worker_arg;

#define CHUNK 32

static void *worker(void *argp) {
  worker_arg *arg = (worker_arg *)argp;
  const frame_ctx *f = arg->f;
  double *hd2 = (double *)malloc(f->k * sizeof(double));
  size_t *hidx = (size_t *)malloc(f->k * sizeof(size_t));
  for (;;) {
    size_t lo = atomic_fetch_add(arg->cursor, CHUNK);
    if (lo >= arg->count) break;
    size_t hi = lo + CHUNK < arg->count ? lo + CHUNK : arg->count;
    for (size_t t = lo; t < hi; ++t) walk_one(f, arg->list[t], hd2, hidx);
  }
  free(hd2);
  free(hidx);
  return NULL;
}

static void run_sources(const frame_ctx *f, const size_t *list, size_t count,
                        int nthreads) {
  if (nthreads <= 1 || count < 2 * CHUNK) {
    double *hd2 = (double *)malloc(f->k * sizeof(double));
    size_t *hidx = (size_t *)malloc(f->k * sizeof(size_t));
    for (size_t t = 0; t < count; ++t) walk_one(f, list[t], hd2, hidx);
    free(hd2);
    free(hidx);
    return;
  }
  atomic_size_t cursor;
  atomic_init(&cursor, 0);
  pthread_t *tids = (pthread_t *)malloc(sizeof(pthread_t) * (size_t)nthreads);
  worker_arg arg = {f, list, count, &cursor};
  int started = 0;
  for (int t = 0; t < nthreads; ++t) {
    if (pthread_create(&tids[t], NULL, worker, &arg) == 0) started++;
    else break;
  }
  if (started == 0) worker(&arg);
  for (int t = 0; t < started; ++t) pthread_join(tids[t], NULL);
  free(tids);
}

/* Fold, bin and chain the active points of one frame
 * (src/knearest.rs:370-428). Returns the active list (caller frees). */
static size_t *prepare_frame(frame_ctx *f, const double *xyz,
                             const int *mask, size_t *n_active) {
  size_t n = f->n;
  size_t *active = (size_t *)malloc((n ? n : 1) * sizeof(size_t));
  size_t na = 0;
  for (size_t i = 0; i < n; ++i)
    if (!mask || mask[i] != 0) active[na++] = i;
  f->folded = calloc(n ? n : 1, sizeof(double[3]));
  f->bin = calloc(n ? n : 1, sizeof(int[3]));
  f->head = (int64_t *)malloc((size_t)f->plan.ncell * sizeof(int64_t));
  f->next = (int64_t *)malloc((n ? n : 1) * sizeof(int64_t));
  for (int64_t c = 0; c < f->plan.ncell; ++c) f->head[c] = -1;
  for (size_t i = 0; i < n; ++i) f->next[i] = -1;
  for (size_t t = 0; t < na; ++t) {
    size_t i = active[t];
    orc_fold_bin(f->cell, &f->plan, xyz + 3 * i, f->folded[i], f->bin[i]);
  }
  for (size_t t = 0; t < na; ++t) {
#if LC_ORACLE_PLANT == LC_PLANT_INCREASING_LISTS
    size_t i = active[na - 1 - t];
#else
    size_t i = active[t];
#endif
    int64_t c = cell_index(f, f->bin[i][0], f->bin[i][1], f->bin[i][2]);
    f->next[i] = f->head[c];
    f->head[c] = (int64_t)i;
  }
  *n_active = na;
  return active;
}

static void free_frame(frame_ctx *f) {
  free(f->folded);
  free(f->bin);
  free(f->head);
  free(f->next);
}

/* write_one, src/knearest.rs:275-307 (search inlined). */
static int write_one(const orc_cell *cell, const double *xyz, size_t n,
                     size_t k, const int *mask, double hint, int *out_nn,
                     double *out_d2, int *out_reach, int *out_cand,
                     int nthreads) {
  for (size_t t = 0; t < n * k; ++t) out_nn[t] = -1;
  if (out_d2) {
    const uint64_t nanbits = 0x7FF8000000000000ULL; /* f64::NAN */
    double nanv;
    memcpy(&nanv, &nanbits, sizeof nanv);
    for (size_t t = 0; t < n * k; ++t) out_d2[t] = nanv;
  }
  if (out_reach) memset(out_reach, 0, n * sizeof(int));
  if (out_cand) memset(out_cand, 0, n * sizeof(int));
  if (k == 0) return ORC_ZERO_K;
  if (n == 0) return ORC_EMPTY;
  size_t any = 0;
  for (size_t i = 0; i < n; ++i)
    if (!mask || mask[i] != 0) { any = 1; break; }
  if (!any) return ORC_OK;
  frame_ctx f;
  memset(&f, 0, sizeof f);
  f.cell = cell;
  int rc = orc_plan_mesh(cell, hint, &f.plan);
  if (rc) return rc;
  f.n = n;
  f.k = k;
  f.out_nn = out_nn;
  f.out_d2 = out_d2;
  f.out_reach = out_reach;
  f.out_cand = out_cand;
  size_t na;
  size_t *active = prepare_frame(&f, xyz, mask, &na);
  run_sources(&f, active, na, nthreads);
  free(active);
  free_frame(&f);
  return ORC_OK;
}

static int cell_from12(const double *c12, orc_cell *cell) {
  return orc_cell_from_vectors(c12, c12 + 3, c12 + 6, c12 + 9, cell);
}

/* lc_knearest_many, src/capi.rs:222-297, then knearest_into_d2 /
 * knearest_into_many (src/knearest.rs:210-273). */
int orc_knearest_many(const double *xyz, size_t n, size_t n_frames,
                      const double *cell12, size_t k, const int *mask,
                      double cell_hint, int *out_nn, double *out_d2,
                      int *out_reach, int *out_cand, int nthreads) {
  if (n == 0 || n_frames == 0) return ORC_EMPTY;
  if (k == 0) return ORC_ZERO_K;
  if (!xyz || !cell12 || !out_nn) return ORC_NULL;
  if (n > SIZE_MAX / n_frames) return ORC_OVERFLOW;
  size_t n_pts = n * n_frames;
  if (n_pts > SIZE_MAX / k) return ORC_OVERFLOW;
  size_t need = n_pts * k;
  size_t max_xyz = ((size_t)PTRDIFF_MAX) / 3;
  size_t max_out = ((size_t)PTRDIFF_MAX) / sizeof(int);
  if (n_pts > max_xyz || need > max_out) return ORC_OVERFLOW;
  orc_cell cell;
  int rc = cell_from12(cell12, &cell);
  if (rc) return rc;
  /* the C ABI maps cell_hint <= 0 (and NaN) to None (src/capi.rs:273-277) */
  double hint = cell_hint > 0.0 ? cell_hint : NAN;
  for (size_t fr = 0; fr < n_frames; ++fr) {
    rc = write_one(&cell, xyz + fr * n * 3, n, k, mask, hint,
                   out_nn + fr * n * k, out_d2 ? out_d2 + fr * n * k : NULL,
                   out_reach ? out_reach + fr * n : NULL,
                   out_cand ? out_cand + fr * n : NULL, nthreads);
    if (rc) return rc;
  }
  return ORC_OK;
}

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

int orc_knearest_sample(const double *xyz, size_t n, const double *cell12,
                        size_t k, const int *mask, double cell_hint,
                        const int *sources, size_t n_sources, int *out_nn,
                        double *out_d2, int *out_reach, int *out_cand,
                        int nthreads, double *times) {
  if (n == 0) return ORC_EMPTY;
  if (k == 0) return ORC_ZERO_K;
  if (!xyz || !cell12 || !out_nn || (!sources && n_sources)) return ORC_NULL;
  orc_cell cell;
  int rc = cell_from12(cell12, &cell);
  if (rc) return rc;
  double hint = cell_hint > 0.0 ? cell_hint : NAN;
  double t0 = now_s();
  frame_ctx f;
  memset(&f, 0, sizeof f);
  f.cell = &cell;
  rc = orc_plan_mesh(&cell, hint, &f.plan);
  if (rc) return rc;
  f.n = n;
  f.k = k;
  f.out_nn = out_nn;
  f.out_d2 = out_d2;
  f.out_reach = out_reach;
  f.out_cand = out_cand;
  size_t na;
  size_t *active = prepare_frame(&f, xyz, mask, &na);
  size_t *list = (size_t *)malloc((n_sources ? n_sources : 1) * sizeof(size_t));
  size_t m = 0;
  for (size_t t = 0; t < n_sources; ++t) {
    int s = sources[t];
    if (s < 0 || (size_t)s >= n) continue;
    if (mask && mask[s] == 0) continue;
    for (size_t q = 0; q < k; ++q) {
      out_nn[(size_t)s * k + q] = -1;
      if (out_d2) {
        const uint64_t nanbits = 0x7FF8000000000000ULL;
        memcpy(&out_d2[(size_t)s * k + q], &nanbits, sizeof(double));
      }
    }
    list[m++] = (size_t)s;
  }
  double t1 = now_s();
  run_sources(&f, list, m, nthreads);
  double t2 = now_s();
  free(list);
  free(active);
  free_frame(&f);
  if (times) {
    times[0] = t1 - t0;
    times[1] = t2 - t1;
  }
  return ORC_OK;
}

/* pair_dist2, src/knearest.rs:24-40 */
static double pair_dist2(const orc_cell *cell, const double p[3],
                         const double q[3]) {
  if (cell->ortho) return orc_dist2(cell, p, q);
  double best = INFINITY;
  for (int na = -1; na <= 1; ++na)
    for (int nb = -1; nb <= 1; ++nb)
      for (int nc = -1; nc <= 1; ++nc) {
        double s[3];
        orc_lattice_shift(cell, na, nb, nc, s);
        double d2 = orc_dist2_shifted(p, q, s);
        if (d2 < best) best = d2;
      }
  return best;
}

typedef struct pair_t {
  double d2;
  size_t j;
} pair_t;

static int pair_cmp(const void *a, const void *b) {
  const pair_t *x = (const pair_t *)a, *y = (const pair_t *)b;
  int64_t kx = total_key(x->d2), ky = total_key(y->d2);
  if (kx != ky) return kx < ky ? -1 : 1;
  return x->j < y->j ? -1 : (x->j > y->j ? 1 : 0);
}

/* knearest_brute, src/knearest.rs:508-541 (packed like write_one). */
int orc_knearest_brute(const double *xyz, size_t n, const double *cell12,
                       size_t k, const int *mask, int *out_nn,
                       double *out_d2) {
  if (k == 0) return ORC_ZERO_K;
  if (n == 0) return ORC_EMPTY;
  orc_cell cell;
  int rc = cell_from12(cell12, &cell);
  if (rc) return rc;
  for (size_t t = 0; t < n * k; ++t) out_nn[t] = -1;
  if (out_d2) {
    const uint64_t nanbits = 0x7FF8000000000000ULL;
    for (size_t t = 0; t < n * k; ++t)
      memcpy(&out_d2[t], &nanbits, sizeof(double));
  }
  pair_t *pairs = (pair_t *)malloc(n * sizeof(pair_t));
  for (size_t i = 0; i < n; ++i) {
    if (mask && mask[i] == 0) continue;
    size_t m = 0;
    for (size_t j = 0; j < n; ++j) {
      if (j == i || (mask && mask[j] == 0)) continue;
      pairs[m].d2 = pair_dist2(&cell, xyz + 3 * i, xyz + 3 * j);
      pairs[m].j = j;
      m++;
    }
    qsort(pairs, m, sizeof(pair_t), pair_cmp);
    size_t lim = m < k ? m : k;
    for (size_t t = 0; t < lim; ++t) {
      out_nn[i * k + t] = (int)pairs[t].j;
      if (out_d2) out_d2[i * k + t] = pairs[t].d2;
    }
  }
  free(pairs);
  return ORC_OK;
}
