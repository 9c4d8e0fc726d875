// C++ consumer of linkcell::gpu::Workspace on device buffers, after the
// reference's tests/gpu_agree.cpp:21-59: the host knearest_into and the
// device Workspace::knearest_into must agree element for element. Device
// memory comes from the library's own C helpers (lc_gpu_alloc /
// lc_gpu_memcpy), so the program needs no HIP headers.
//
//   gpu_agree <input.bin> <output.bin>
// input: int64 n, frames, k, has_mask; double hint; double cell[12]
// (rows a, b, c, origin); frames*n*3 doubles; n int32 mask if has_mask.
// output: frames*n*k int32 indices, then frames*n*k doubles (d2), from the
// device search; the host search is compared here.
#include "linkcell.hpp"
#include "linkcell_gpu.hpp"
#include "linkcell_gpu.h"

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {

int fail(const char* what) {
  std::fprintf(stderr, "gpu_agree: %s\n", what);
  return 1;
}

template <typename T>
bool read_all(std::FILE* f, T* p, std::size_t n) {
  return std::fread(p, sizeof(T), n, f) == n;
}

struct DevBuf {
  void* p = nullptr;
  explicit DevBuf(std::size_t bytes) {
    if (lc_gpu_alloc(&p, bytes) != 0) p = nullptr;
  }
  ~DevBuf() { lc_gpu_free(p); }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc != 3) return fail("usage: gpu_agree in out");
  if (!linkcell::gpu::available()) return fail("no HIP device");
  std::FILE* in = std::fopen(argv[1], "rb");
  if (!in) return fail("open input");
  std::int64_t hdr[4];
  double hint = 0.0, c12[12];
  if (!read_all(in, hdr, 4) || !read_all(in, &hint, 1) || !read_all(in, c12, 12)) {
    return fail("read header");
  }
  const std::size_t n = static_cast<std::size_t>(hdr[0]);
  const std::size_t frames = static_cast<std::size_t>(hdr[1]);
  const std::size_t k = static_cast<std::size_t>(hdr[2]);
  const bool has_mask = hdr[3] != 0;
  std::vector<double> xyz(frames * n * 3);
  std::vector<int> mask(has_mask ? n : 0);
  if (!read_all(in, xyz.data(), xyz.size())) return fail("read xyz");
  if (has_mask && !read_all(in, mask.data(), n)) return fail("read mask");
  std::fclose(in);
  const linkcell::Cell cell = linkcell::Cell::from_vectors(
      {c12[0], c12[1], c12[2]}, {c12[3], c12[4], c12[5]}, {c12[6], c12[7], c12[8]},
      {c12[9], c12[10], c12[11]});
  const std::size_t m = frames * n * k;
  std::vector<int> host(m, -2);
  std::vector<double> host_d2(m, 0.0);
  linkcell::knearest_into_many(xyz.data(), n, frames, cell, k, host.data(), m,
                               has_mask ? mask.data() : nullptr, hint, host_d2.data());

  DevBuf dx(xyz.size() * sizeof(double)), dm(has_mask ? n * sizeof(int) : 4),
      dn(m * sizeof(int)), dd(m * sizeof(double));
  if (!dx.p || !dm.p || !dn.p || !dd.p) return fail("lc_gpu_alloc");
  if (lc_gpu_memcpy(dx.p, xyz.data(), xyz.size() * sizeof(double), 1) != 0) {
    return fail("memcpy xyz");
  }
  if (has_mask && lc_gpu_memcpy(dm.p, mask.data(), n * sizeof(int), 1) != 0) {
    return fail("memcpy mask");
  }
  const int* dmask = has_mask ? static_cast<const int*>(dm.p) : nullptr;
  std::vector<int> dev(m, -3);
  std::vector<double> dev_d2(m, 0.0);
  try {
    linkcell::gpu::Workspace ws;
    if (frames == 1) {
      ws.knearest_into(static_cast<const double*>(dx.p), n, cell, k,
                       static_cast<int*>(dn.p), m, dmask, hint,
                       static_cast<double*>(dd.p));
      if (lc_gpu_memcpy(dev.data(), dn.p, m * sizeof(int), 2) != 0) return fail("DtoH");
      for (std::size_t i = 0; i < m; ++i) {
        if (dev[i] != host[i]) {
          std::fprintf(stderr, "knearest_into mismatch at %zu: host %d device %d\n", i,
                       host[i], dev[i]);
          return 1;
        }
      }
    }
    // frame batch without a wait, then wait(): the results are only read
    // after it; the same workspace is reused (grow-only buffers)
    if (lc_gpu_fill_i32(dn.p, -1, m) != 0) return fail("fill");
    ws.knearest_into_many(static_cast<const double*>(dx.p), n, frames, cell, k,
                          static_cast<int*>(dn.p), m, dmask, hint, false, nullptr,
                          static_cast<double*>(dd.p));
    if (ws.queue() == nullptr) return fail("queue");
    ws.wait();
    if (lc_gpu_memcpy(dev.data(), dn.p, m * sizeof(int), 2) != 0 ||
        lc_gpu_memcpy(dev_d2.data(), dd.p, m * sizeof(double), 2) != 0) {
      return fail("DtoH");
    }
  } catch (const linkcell::Error& e) {
    std::fprintf(stderr, "gpu_agree: %s\n", e.what());
    return 1;
  }
  for (std::size_t i = 0; i < m; ++i) {
    if (dev[i] != host[i] || std::memcmp(&dev_d2[i], &host_d2[i], sizeof(double)) != 0) {
      std::fprintf(stderr, "knearest_into_many mismatch at %zu: host %d device %d\n", i,
                   host[i], dev[i]);
      return 1;
    }
  }
  std::FILE* out = std::fopen(argv[2], "wb");
  if (!out) return fail("open output");
  std::fwrite(dev.data(), sizeof(int), m, out);
  std::fwrite(dev_d2.data(), sizeof(double), m, out);
  std::fclose(out);
  std::printf("gpu_agree: n=%zu frames=%zu k=%zu agree\n", n, frames, k);
  return 0;
}
