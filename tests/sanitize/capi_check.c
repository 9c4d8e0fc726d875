/*
 * Host sanitizer driver for the library's host code (lc_api.hip and the
 * engine's host side), built by tests/test_host_sanitizers.py against a
 * host-only ASan + UBSan build of the library. It runs, with no GPU, the
 * contract the reference's ABI tests check (tests/abi.rs, src/capi.rs,
 * src/gpu/capi.cpp): validation order and messages, the fill before a
 * failed search, the all-masked success, thread-local error slots, and
 * the GPU waist failing cleanly without a device. Exit status 0 when every
 * check holds.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "linkcell.h"
#include "linkcell_amd.h"
#include "linkcell_gpu.h"

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, \
              #cond);                                                 \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static int err_is(const char *want) {
  const char *e = lc_last_error();
  return e != NULL && strcmp(e, want) == 0;
}
static int gpu_err_is(const char *want) {
  const char *e = lc_gpu_last_error();
  return e != NULL && strcmp(e, want) == 0;
}

static struct lc_cell ortho(double a, double b, double c) {
  struct lc_cell x;
  memset(&x, 0, sizeof x);
  x.ax = a;
  x.by = b;
  x.cz = c;
  return x;
}

struct slot_arg {
  size_t n, k;
  char seen[64];
};
static void *slot_thread(void *p) {
  struct slot_arg *a = (struct slot_arg *)p;
  struct lc_cell box = ortho(10, 10, 10);
  int out[1] = {-1};
  double pt[3] = {0, 0, 0};
  (void)lc_knearest(pt, a->n, &box, a->k, NULL, 0.0, out);
  const char *e = lc_last_error();
  snprintf(a->seen, sizeof a->seen, "%s", e ? e : "(none)");
  return NULL;
}

int main(void) {
  CHECK(strcmp(lc_version(), "0.3.3") == 0);
  struct lc_cell box = ortho(10, 10, 10);
  int out1[1] = {-1};
  double pt[3] = {0, 0, 0};
  CHECK(lc_knearest(pt, 0, &box, 1, NULL, 0.0, out1) != 0 && err_is("no points"));
  (void)lc_version();
  CHECK(err_is("no points"));
  CHECK(lc_knearest(pt, 1, &box, 0, NULL, 0.0, out1) != 0 && err_is("k must be at least 1"));
  CHECK(lc_knearest(NULL, 1, &box, 1, NULL, 0.0, out1) != 0 && err_is("null pointer"));
  CHECK(lc_knearest(pt, (SIZE_MAX >> 1) / 3 + 1, &box, 1, NULL, 0.0, out1) != 0 &&
        err_is("n * k overflows"));

  int out2[2] = {-7, -7};
  double d2[4] = {0, 0, 0, 0};
  double pts[6] = {0, 0, 0, 1, 0, 0};
  struct lc_cell bad = ortho(0, 10, 10);
  CHECK(lc_knearest_d2(pts, 2, &bad, 1, NULL, 0.0, out2, d2) != 0 &&
        err_is("singular or non-positive cell"));
  struct lc_cell big = ortho(1e4, 1e4, 1e4);
  CHECK(lc_knearest_d2(pts, 2, &big, 1, NULL, 1e-3, out2, d2) != 0 &&
        err_is("linked-cell mesh is too fine"));
  CHECK(out2[0] == -1 && out2[1] == -1);

  int out4[4] = {-7, -7, -7, -7};
  int mask[2] = {0, 0};
  CHECK(lc_knearest_d2(pts, 2, &big, 2, mask, 1e-3, out4, d2) == 0 && lc_last_error() == NULL);
  for (int i = 0; i < 4; ++i) CHECK(out4[i] == -1 && isnan(d2[i]));

  struct slot_arg a = {0, 1, ""}, b = {1, 0, ""};
  pthread_t ta, tb;
  pthread_create(&ta, NULL, slot_thread, &a);
  pthread_create(&tb, NULL, slot_thread, &b);
  pthread_join(ta, NULL);
  pthread_join(tb, NULL);
  CHECK(strcmp(a.seen, "no points") == 0);
  CHECK(strcmp(b.seen, "k must be at least 1") == 0);

  if (lc_gpu_available() == 0) {
    CHECK(lc_gpu_workspace_new() == NULL && gpu_err_is("HIP device not available"));
    void *p = NULL;
    CHECK(lc_gpu_alloc(&p, 64) != 0);
    CHECK(lc_gpu_knearest(NULL, NULL, 1, &box, 1, NULL, 0.0, NULL, NULL) != 0 &&
          gpu_err_is("null pointer"));
    CHECK(lc_gpu_fill_i32(NULL, 5, 0) != 0);
    int o[3];
    double dd[3];
    CHECK(lc_gpu_knearest_noalloc(pts, 2, 1, &box, 1, NULL, 0.0, o, dd, NULL) != 0);
  }
  if (failures == 0) printf("capi host checks: ok\n");
  return failures != 0;
}
