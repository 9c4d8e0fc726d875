/*
 * Host sanitizer driver for the oracle (test infrastructure): built by
 * tests/test_host_sanitizers.py with ASan + UBSan and, separately, TSan,
 * it runs the C restatement on seeded cases that reach every branch of the
 * walk (orthorhombic and triclinic cells, points far outside the primary
 * cell, masks, several frames, k above the point count, small boxes where
 * the stencil wraps, coincident points, 1 and 4 threads) and checks the
 * linked-cell search against the restated brute force: indices equal
 * (except between exact ties at a row's k-th distance, where the walk keeps
 * the index it met first), d2 within the reference's 1e-12 (the brute force
 * measures unfolded points).
 * Points far outside the primary cell, beyond the brute force's 27 images,
 * are searched and checked for well-formed rows only. With the argument
 * "threads" only the multi-threaded cases run (the TSan build). Exit
 * status 0 when every case agrees.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../oracle/lc_oracle.h"

static uint64_t lcg_state;
static double lcg01(void) {
  lcg_state = lcg_state * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(lcg_state >> 11) * (1.0 / 9007199254740992.0);
}

static int almost(double a, double b) {
  const double m = fabs(b) > 1.0 ? fabs(b) : 1.0;
  return fabs(a - b) <= 1e-12 * m;
}

/* brute: 1 compares with the brute force, 0 checks that every row holds
 * distinct in-range indices with non-decreasing d2, then -1 / NaN */
static int check_case(const char *name, const double cell12[12], size_t n, size_t frames,
                      size_t k, int use_mask, double hint, double spread, int coincident,
                      int nthreads, int brute) {
  const size_t np = n * frames;
  double *xyz = malloc(sizeof(double) * 3 * np);
  int *mask = use_mask ? malloc(sizeof(int) * n) : NULL;
  int *nn = malloc(sizeof(int) * np * k), *nb = malloc(sizeof(int) * n * k);
  double *d2 = malloc(sizeof(double) * np * k), *db = malloc(sizeof(double) * n * k);
  int *reach = malloc(sizeof(int) * np), *cand = malloc(sizeof(int) * np);
  if (!xyz || (use_mask && !mask) || !nn || !nb || !d2 || !db || !reach || !cand) return 2;
  for (size_t i = 0; i < np; ++i) {
    /* fractional coordinates in [-spread, 1 + spread), mapped by the cell */
    double s[3];
    for (int a = 0; a < 3; ++a) s[a] = -spread + (1.0 + 2.0 * spread) * lcg01();
    for (int a = 0; a < 3; ++a) {
      xyz[3 * i + a] = s[0] * cell12[a] + s[1] * cell12[3 + a] + s[2] * cell12[6 + a] +
                       cell12[9 + a];
    }
    if (coincident && i % 7 == 3) memcpy(&xyz[3 * i], &xyz[3 * (i - 1)], sizeof(double) * 3);
  }
  for (size_t i = 0; use_mask && i < n; ++i) mask[i] = lcg01() < 0.9;
  int rc = orc_knearest_many(xyz, n, frames, cell12, k, mask, hint, nn, d2, reach, cand,
                             nthreads);
  if (rc) {
    fprintf(stderr, "%s: knearest_many failed: %s\n", name, orc_error_message(rc));
    return 1;
  }
  int bad = 0;
  for (size_t i = 0; !brute && i < np && !bad; ++i) {
    int ended = 0;
    for (size_t t = 0; t < k; ++t) {
      const int j = nn[i * k + t];
      const double v = d2[i * k + t];
      if (j < 0) {
        ended = 1;
        bad |= !isnan(v);
        continue;
      }
      bad |= ended || j >= (int)n || !(v >= 0.0);
      if (t > 0 && nn[i * k + t - 1] >= 0) bad |= v < d2[i * k + t - 1];
      for (size_t u = 0; u < t; ++u) bad |= nn[i * k + u] == j;
    }
    if (bad) fprintf(stderr, "%s: row %zu is malformed\n", name, i);
  }
  for (size_t f = 0; brute && f < frames && !bad; ++f) {
    rc = orc_knearest_brute(xyz + 3 * n * f, n, cell12, k, mask, nb, db);
    if (rc) {
      fprintf(stderr, "%s: brute failed: %s\n", name, orc_error_message(rc));
      return 1;
    }
    for (size_t e = 0; e < n * k; ++e) {
      const int a = nn[f * n * k + e], b = nb[e];
      const double x = d2[f * n * k + e], y = db[e];
      /* an exact tie at the row's k-th distance may keep either index:
         the walk keeps the one it met first, the brute force the lower */
      const double last = d2[f * n * k + (e / k) * k + k - 1];
      const int tie = !isnan(x) && x == last;
      const int same = (a == b || tie) && ((isnan(x) && isnan(y)) || almost(x, y));
      if (!same) {
        fprintf(stderr, "%s: frame %zu slot %zu: %d/%.17g against brute %d/%.17g\n", name,
                f, e, a, x, b, y);
        bad = 1;
        break;
      }
    }
  }
  /* the sampled driver on a few sources of frame 0 */
  size_t ns = n < 5 ? n : 5;
  int src[5];
  for (size_t i = 0; i < ns; ++i) src[i] = (int)((i * 7919u) % n);
  double times[2];
  rc = orc_knearest_sample(xyz, n, cell12, k, mask, hint, src, ns, nn, d2, reach, cand,
                           nthreads, times);
  if (rc) {
    fprintf(stderr, "%s: knearest_sample failed: %s\n", name, orc_error_message(rc));
    bad = 1;
  }
  free(xyz), free(mask), free(nn), free(nb), free(d2), free(db), free(reach), free(cand);
  if (!bad) printf("%s: ok\n", name);
  return bad;
}

int main(int argc, char **argv) {
  const int threads_only = argc > 1 && strcmp(argv[1], "threads") == 0;
  lcg_state = 12345;
  const double ortho16[12] = {16, 0, 0, 0, 16, 0, 0, 0, 16, 0, 0, 0};
  const double tri[12] = {40, 0, 0, 20, 34.64, 0, 8, 6, 38, 1, -2, 3};
  const double small[12] = {4, 0, 0, 0, 5, 0, 0, 0, 6, 0, 0, 0};
  const double slab[12] = {20, 0, 0, 0, 20, 0, 0, 0, 60, 0, 0, 0};
  int bad = 0;
  bad |= check_case("ortho16_k12_threads", ortho16, 1200, 1, 12, 0, 3.0, 0.0, 0, 4, 1);
  bad |= check_case("triclinic_masked_threads", tri, 900, 1, 16, 1, 0.0, 0.2, 0, 4, 1);
  bad |= check_case("coincident_points_threads", slab, 900, 1, 8, 0, 2.0, 0.0, 1, 3, 1);
  bad |= check_case("far_outside_threads", tri, 900, 2, 12, 1, 0.0, 2.0, 0, 4, 0);
  if (!threads_only) {
    bad |= check_case("ortho16_k12", ortho16, 1200, 1, 12, 0, 3.0, 0.0, 0, 1, 1);
    bad |= check_case("small_box_k100_frames", small, 300, 3, 100, 0, 0.0, 0.05, 0, 1, 1);
    bad |= check_case("k_above_n", small, 40, 2, 64, 1, 1.0, 0.0, 0, 1, 1);
    bad |= check_case("one_point", ortho16, 1, 1, 3, 0, 0.0, 0.0, 0, 1, 1);
  }
  return bad;
}
