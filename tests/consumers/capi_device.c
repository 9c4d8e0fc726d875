/* C consumer of the lc_gpu_* waist in the order of the reference's Python
 * binding (python/src/gpu.rs:252-359): workspace, queue, device alloc,
 * host-to-device copy (kind 1), lc_gpu_fill_i32(-1) of the index buffer,
 * lc_gpu_knearest_many with wait = 0, lc_gpu_wait, device-to-host copy
 * (kind 2). Then lc_gpu_knearest (which waits) into a zero-filled buffer,
 * and the argument errors of the helpers.
 *
 *   capi_device <input.bin> <output.bin>
 * Same input format as gpu_agree.cpp; output: the lc_gpu_knearest_many
 * indices and d2, then (single frame only) the lc_gpu_knearest ones. */
#include "linkcell_gpu.h"

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static int fail(const char *what) {
  const char *e = lc_gpu_last_error();
  fprintf(stderr, "capi_device: %s (%s)\n", what, e ? e : "no message");
  return 1;
}

int main(int argc, char **argv) {
  if (argc != 3) return fail("usage");
  if (lc_gpu_available() != 1) return fail("no device");
  FILE *in = fopen(argv[1], "rb");
  if (!in) return fail("open input");
  int64_t hdr[4];
  double hint, c12[12];
  if (fread(hdr, 8, 4, in) != 4 || fread(&hint, 8, 1, in) != 1 ||
      fread(c12, 8, 12, in) != 12)
    return fail("read header");
  size_t n = (size_t)hdr[0], frames = (size_t)hdr[1], k = (size_t)hdr[2];
  int has_mask = hdr[3] != 0;
  size_t nx = frames * n * 3, m = frames * n * k;
  double *xyz = malloc(nx * sizeof(double));
  int *mask = malloc((has_mask ? n : 1) * sizeof(int));
  int *nn = malloc(m * sizeof(int));
  double *d2 = malloc(m * sizeof(double));
  if (fread(xyz, 8, nx, in) != nx) return fail("read xyz");
  if (has_mask && fread(mask, 4, n, in) != n) return fail("read mask");
  fclose(in);
  lc_cell box;
  memcpy(&box, c12, sizeof box);

  lc_gpu_workspace *ws = lc_gpu_workspace_new();
  if (!ws) return fail("workspace");
  if (lc_gpu_queue(ws) == NULL) return fail("queue");
  void *dx = NULL, *dm = NULL, *dn = NULL, *dd = NULL;
  if (lc_gpu_alloc(&dx, nx * sizeof(double)) != 0) return fail("alloc xyz");
  if (lc_gpu_memcpy(dx, xyz, nx * sizeof(double), 1) != 0) return fail("HtoD xyz");
  if (has_mask) {
    if (lc_gpu_alloc(&dm, n * sizeof(int)) != 0) return fail("alloc mask");
    if (lc_gpu_memcpy(dm, mask, n * sizeof(int), 1) != 0) return fail("HtoD mask");
  }
  if (lc_gpu_alloc(&dn, m * sizeof(int)) != 0) return fail("alloc nn");
  if (lc_gpu_fill_i32(dn, -1, m) != 0) return fail("fill -1");
  if (lc_gpu_alloc(&dd, m * sizeof(double)) != 0) return fail("alloc d2");
  if (lc_gpu_knearest_many(ws, dx, n, frames, &box, k, dm, hint, dn, dd, 0) != 0)
    return fail("knearest_many");
  lc_gpu_wait(ws);
  if (lc_gpu_memcpy(nn, dn, m * sizeof(int), 2) != 0) return fail("DtoH nn");
  if (lc_gpu_memcpy(d2, dd, m * sizeof(double), 2) != 0) return fail("DtoH d2");
  FILE *out = fopen(argv[2], "wb");
  if (!out) return fail("open output");
  fwrite(nn, 4, m, out);
  fwrite(d2, 8, m, out);
  if (frames == 1) {
    /* zero fill, read back, then the waiting single-frame entry point */
    if (lc_gpu_fill_i32(dn, 0, m) != 0) return fail("fill 0");
    if (lc_gpu_memcpy(nn, dn, m * sizeof(int), 2) != 0) return fail("DtoH zeros");
    for (size_t i = 0; i < m; ++i)
      if (nn[i] != 0) return fail("fill 0 left a nonzero");
    if (lc_gpu_knearest(ws, dx, n, &box, k, dm, hint, dn, dd) != 0)
      return fail("knearest");
    if (lc_gpu_memcpy(nn, dn, m * sizeof(int), 2) != 0) return fail("DtoH nn");
    if (lc_gpu_memcpy(d2, dd, m * sizeof(double), 2) != 0) return fail("DtoH d2");
    fwrite(nn, 4, m, out);
    fwrite(d2, 8, m, out);
  }
  fclose(out);
  /* argument contract of the helpers (src/gpu/capi.cpp:247-282) */
  if (lc_gpu_fill_i32(dn, 5, m) == 0) return fail("fill 5 accepted");
  if (strstr(lc_gpu_last_error(), "0 and -1") == NULL) return fail("fill message");
  if (lc_gpu_memcpy(nn, NULL, 4, 2) == 0) return fail("null memcpy accepted");
  lc_gpu_free(dx);
  lc_gpu_free(dm);
  lc_gpu_free(dn);
  lc_gpu_free(dd);
  lc_gpu_workspace_free(ws);
  free(xyz);
  free(mask);
  free(nn);
  free(d2);
  printf("capi_device: n=%zu frames=%zu k=%zu done\n", n, frames, k);
  return 0;
}
