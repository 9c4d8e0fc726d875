// Drop-in surfaces over the engine:
//   linkcell::gpu::Workspace      (reference include/linkcell_gpu.hpp:29-78,
//                                  src/gpu/knearest.cpp:676-878)
//   lc_gpu_* C waist              (reference src/gpu/capi.cpp:19-284)
//   lc_knearest* host-pointer API (reference src/capi.rs:79-297), staged
//                                  through device memory
//   lc_gpu_knearest_ex and the timing / mode extensions
#include "../../include/linkcell_amd.h"
#include "lc_engine.hpp"
#include "lc_sched.hpp"
#include "../../include/linkcell_gpu.h"
#include "../../include/linkcell_gpu.hpp"
#include "lc_cell.hpp"
#include "lc_engine.hpp"

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <string>
#include <vector>

namespace {

thread_local std::string g_gpu_err;
thread_local std::string g_cpu_err;
thread_local bool g_cpu_err_set = false;

void set_gpu_err(const char* m) { g_gpu_err = m ? m : ""; }

bool device_present() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return false;
  return n > 0;
}

const char* kNoDevice = "HIP device not available";

double hint_option(double cell_hint) {
  // The C waist maps cell_hint <= 0 (and NaN) to None (src/capi.rs:273-277);
  // the search then maps None and non-finite hints to 3.0
  // (src/knearest.rs:377-380).
  return cell_hint > 0.0 ? cell_hint : std::numeric_limits<double>::quiet_NaN();
}

struct Geometry {
  lcamd::CellGeom geom;
  lcamd::MeshPlan plan;
};

}  // namespace

namespace linkcell {
namespace gpu {

struct Workspace::Impl {
  lcamd::Engine engine;
};

bool available() { return device_present(); }

Workspace::Workspace() : impl_(nullptr) {
  if (!device_present()) throw Error(kNoDevice);
  impl_ = new Impl();
}

Workspace::~Workspace() { delete impl_; }

Workspace::Workspace(Workspace&& o) noexcept : impl_(o.impl_) { o.impl_ = nullptr; }

Workspace& Workspace::operator=(Workspace&& o) noexcept {
  if (this != &o) {
    delete impl_;
    impl_ = o.impl_;
    o.impl_ = nullptr;
  }
  return *this;
}

void* Workspace::queue() { return impl_->engine.stream(); }

void Workspace::wait() {
  try {
    impl_->engine.sync();
  } catch (const std::runtime_error& e) {
    throw Error(e.what());
  }
}

}  // namespace gpu
}  // namespace linkcell

namespace lcamd {

// Shared validation and frame construction for every device entry point
// (order follows src/gpu/knearest.cpp:711-762).
void search_device(Engine& eng, const double* xyz, size_t n, size_t nframes,
                   const double* a, const double* b, const double* c,
                   const double* o, size_t k, int* out, size_t out_len,
                   const int* mask, double cell_hint, bool wait,
                   const double* frame_box, double* out_d2, int* out_reach) {
  using linkcell::Error;
  if (!device_present()) throw Error(kNoDevice);
  if (n == 0 || nframes == 0) throw Error("no points");
  if (k == 0) throw Error("k must be at least 1");
  if (n > SIZE_MAX / nframes || n * nframes > SIZE_MAX / k) {
    throw Error("n * k overflows");
  }
  if (out_len != n * k * nframes) {
    throw Error("out buffer length must be nFrames * n * k");
  }
  if (xyz == nullptr || out == nullptr) throw Error("null pointer");
  if (n * nframes >= static_cast<size_t>(INT_MAX) || k >= static_cast<size_t>(INT_MAX)) {
    throw Error("n * n_frames and k must be below 2^31 - 1");
  }
  CellGeom base{};
  if (cell_from_vectors(a, b, c, o, base) != kOk) throw Error("bad cell");
  const double hint = hint_option(cell_hint);
  std::vector<FrameSpec> frames(nframes);
  if (frame_box != nullptr && !base.ortho) {
    throw Error("frame boxes on a sheared cell need a shared H");
  }
  MeshPlan plan{};
  if (frame_box == nullptr && plan_mesh(base, hint, plan) != kOk) {
    throw Error("too many cells");
  }
  for (size_t f = 0; f < nframes; ++f) {
    FrameSpec& s = frames[f];
    s.xyz = xyz + f * n * 3;
    s.mask = mask;
    s.out_nn = out + f * n * k;
    s.out_d2 = out_d2 ? out_d2 + f * n * k : nullptr;
    s.out_reach = out_reach ? out_reach + f * n : nullptr;
    s.n = static_cast<int>(n);
    s.k = static_cast<int>(k);
    if (frame_box != nullptr) {
      const double* L = frame_box + 3 * f;
      if (!(L[0] > 0.0 && L[1] > 0.0 && L[2] > 0.0)) throw Error("bad frame box");
      const double fa[3] = {L[0], 0.0, 0.0}, fb[3] = {0.0, L[1], 0.0},
                   fc[3] = {0.0, 0.0, L[2]};
      if (cell_from_vectors(fa, fb, fc, o, s.geom) != kOk) throw Error("bad frame box");
      if (plan_mesh(s.geom, hint, s.plan) != kOk) throw Error("too many cells");
    } else {
      s.geom = base;
      s.plan = plan;
    }
  }
  try {
    eng.run(frames, wait);
  } catch (const std::runtime_error& e) {
    throw Error(e.what());
  }
}

}  // namespace lcamd

namespace linkcell {
namespace gpu {

void Workspace::knearest_into(const double* xyz, std::size_t n, const Cell& cell,
                              std::size_t k, int* out, std::size_t out_len,
                              const int* mask, double cell_hint, double* out_d2) {
  knearest_into_many(xyz, n, 1, cell, k, out, out_len, mask, cell_hint, true,
                     nullptr, out_d2);
}

void Workspace::knearest_into_many(const double* xyz, std::size_t n,
                                   std::size_t nFrames, const Cell& cell,
                                   std::size_t k, int* out, std::size_t out_len,
                                   const int* mask, double cell_hint, bool wait,
                                   const double* frameBox, double* out_d2) {
  lcamd::search_device(impl_->engine, xyz, n, nFrames, cell.a.data(),
                       cell.b.data(), cell.c.data(), cell.origin.data(), k, out,
                       out_len, mask, cell_hint, wait, frameBox, out_d2, nullptr);
}

void Workspace::knearest_into_many_dcell(const double* xyz, std::size_t n,
                                         std::size_t nFrames, const double* cell,
                                         int cell_n, std::size_t k, int* out,
                                         std::size_t out_len, const int* mask,
                                         double cell_hint, bool wait,
                                         double* out_d2) {
  if (!device_present()) throw Error(kNoDevice);
  if (n == 0 || nFrames == 0) throw Error("no points");
  if (k == 0) throw Error("k must be at least 1");
  if (xyz == nullptr || out == nullptr || cell == nullptr) throw Error("null pointer");
  if (cell_n != 3 && cell_n != 9 && cell_n != 12) {
    throw Error("device cell must have 3, 9, or 12 doubles");
  }
  if (n > SIZE_MAX / nFrames || n * nFrames > SIZE_MAX / k) throw Error("n * k overflows");
  if (out_len != n * k * nFrames) {
    throw Error("out buffer length must be nFrames * n * k");
  }
  if (n * nFrames >= static_cast<size_t>(INT_MAX) || k >= static_cast<size_t>(INT_MAX)) {
    throw Error("n * n_frames and k must be below 2^31 - 1");
  }
  // The cell stays on the device: it is inverted and the mesh planned
  // there (no read-back, no host copy), so the search can be captured.
  // A bad cell or too fine a mesh is reported after the wait.
  lcamd::FrameSpec spec;
  spec.xyz = xyz;
  spec.mask = mask;
  spec.out_nn = out;
  spec.out_d2 = out_d2;
  spec.n = static_cast<int>(n);
  spec.k = static_cast<int>(k);
  try {
    impl_->engine.run_dcell(spec, static_cast<int>(nFrames), cell, cell_n,
                            hint_option(cell_hint), wait);
  } catch (const std::runtime_error& e) {
    throw Error(e.what());
  }
}

}  // namespace gpu
}  // namespace linkcell

struct lc_gpu_workspace {
  linkcell::gpu::Workspace ws;
};

namespace {

template <typename F>
int guarded(F&& fn) {
  try {
    fn();
    set_gpu_err(nullptr);
    return 0;
  } catch (const std::exception& e) {
    set_gpu_err(e.what());
    return 1;
  }
}

}  // namespace

extern "C" {

int lc_gpu_available(void) { return device_present() ? 1 : 0; }

const char* lc_gpu_last_error(void) {
  return g_gpu_err.empty() ? nullptr : g_gpu_err.c_str();
}

lc_gpu_workspace* lc_gpu_workspace_new(void) {
  if (!device_present()) {
    set_gpu_err(kNoDevice);
    return nullptr;
  }
  try {
    return new lc_gpu_workspace();
  } catch (const std::exception& e) {
    set_gpu_err(e.what());
    return nullptr;
  }
}

void lc_gpu_workspace_free(lc_gpu_workspace* ws) { delete ws; }

int lc_gpu_knearest(lc_gpu_workspace* ws, const double* xyz, size_t n,
                    const struct lc_cell* simbox, size_t k, const int* mask,
                    double cell_hint, int* out_nn, double* out_d2) {
  return lc_gpu_knearest_many(ws, xyz, n, 1, simbox, k, mask, cell_hint, out_nn,
                              out_d2, 1);
}

int lc_gpu_knearest_ex(lc_gpu_workspace* ws, const double* xyz, size_t n,
                       size_t n_frames, const struct lc_cell* simbox, size_t k,
                       const int* mask, double cell_hint, int* out_nn,
                       double* out_d2, int* out_reach, int wait) {
  if (ws == nullptr || xyz == nullptr || simbox == nullptr || out_nn == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    const lc_cell& cb = *simbox;
    const double a[3] = {cb.ax, cb.ay, cb.az}, b[3] = {cb.bx, cb.by, cb.bz},
                 c[3] = {cb.cx, cb.cy, cb.cz}, o[3] = {cb.ox, cb.oy, cb.oz};
    const size_t out_len = n * k * n_frames;
    lcamd::search_device(ws->ws.impl()->engine, xyz, n, n_frames, a, b, c, o, k,
                         out_nn, out_len, mask, cell_hint, wait != 0, nullptr,
                         out_d2, out_reach);
  });
}

int lc_gpu_knearest_many(lc_gpu_workspace* ws, const double* xyz, size_t n,
                         size_t n_frames, const struct lc_cell* simbox,
                         size_t k, const int* mask, double cell_hint,
                         int* out_nn, double* out_d2, int wait) {
  return lc_gpu_knearest_ex(ws, xyz, n, n_frames, simbox, k, mask, cell_hint,
                            out_nn, out_d2, nullptr, wait);
}

int lc_gpu_knearest_many_dcell(lc_gpu_workspace* ws, const double* xyz,
                               size_t n, size_t n_frames, const double* cell,
                               int cell_n, size_t k, const int* mask,
                               double cell_hint, int* out_nn, double* out_d2,
                               int wait) {
  if (ws == nullptr || xyz == nullptr || cell == nullptr || out_nn == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    const size_t out_len = n * k * n_frames;
    ws->ws.knearest_into_many_dcell(xyz, n, n_frames, cell, cell_n, k, out_nn,
                                    out_len, mask, cell_hint, wait != 0, out_d2);
  });
}

void* lc_gpu_queue(lc_gpu_workspace* ws) {
  return ws == nullptr ? nullptr : ws->ws.queue();
}

void lc_gpu_wait(lc_gpu_workspace* ws) {
  if (ws != nullptr) {
    try {
      ws->ws.wait();
      set_gpu_err(nullptr);
    } catch (const std::exception& e) {
      // a failure the stream reports (a device-cell plan of a call made
      // with wait = 0) goes to the error slot
      set_gpu_err(e.what());
    }
  }
}

int lc_gpu_set_mode(lc_gpu_workspace* ws, int mode) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  if (mode != LC_GPU_MODE_AUTO && mode != LC_GPU_MODE_EXACT_ONLY) {
    set_gpu_err("unknown mode");
    return 1;
  }
  ws->ws.impl()->engine.set_mode(static_cast<lcamd::WalkMode>(mode));
  set_gpu_err(nullptr);
  return 0;
}

int lc_gpu_set_timing(lc_gpu_workspace* ws, int on) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  ws->ws.impl()->engine.set_timing(on != 0);
  set_gpu_err(nullptr);
  return 0;
}

int lc_gpu_last_timing(lc_gpu_workspace* ws, double* ms7, int* deferred) {
  if (ws == nullptr || ms7 == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    const lcamd::Timings& t = ws->ws.impl()->engine.last_timings();
    ms7[0] = t.fold;
    ms7[1] = t.scan;
    ms7[2] = t.scatter;
    ms7[3] = t.sort;
    ms7[4] = t.walk;
    ms7[5] = t.exact;
    ms7[6] = t.total;
    if (deferred) *deferred = t.deferred;
  });
}

int lc_gpu_last_timing_ex(lc_gpu_workspace* ws, double* ms, int n) {
  if (ws == nullptr || ms == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    const lcamd::Timings& t = ws->ws.impl()->engine.last_timings();
    const double v[12] = {t.fold,  t.scan,  t.scatter, t.sort,
                          t.walk,  t.exact, t.total,   t.setup,
                          t.pass1, t.retry, static_cast<double>(t.deferred),
                          static_cast<double>(t.retried)};
    for (int i = 0; i < n; ++i) ms[i] = i < 12 ? v[i] : 0.0;
  });
}

}  // extern "C"

namespace {

// Validate and convert a batch (src/gpu/knearest.cpp:711-762 per frame).
std::vector<lcamd::FrameSpec> batch_specs(const lc_gpu_frame* frames,
                                          size_t n_frames) {
  using linkcell::Error;
  if (!device_present()) throw Error(kNoDevice);
  if (frames == nullptr) throw Error("null pointer");
  if (n_frames == 0) throw Error("no points");
  std::vector<lcamd::FrameSpec> out(n_frames);
  size_t total = 0;
  for (size_t f = 0; f < n_frames; ++f) {
    const lc_gpu_frame& fr = frames[f];
    const std::string tag = "frame " + std::to_string(f) + ": ";
    if (fr.n == 0) throw Error(tag + "no points");
    if (fr.k == 0) throw Error(tag + "k must be at least 1");
    if (fr.xyz == nullptr || fr.out_nn == nullptr) throw Error(tag + "null pointer");
    if (fr.n >= static_cast<size_t>(INT_MAX) || fr.k >= static_cast<size_t>(INT_MAX) ||
        fr.n > SIZE_MAX / fr.k) {
      throw Error(tag + "n * k overflows");
    }
    total += fr.n;
    if (total >= static_cast<size_t>(INT_MAX) / 2) {
      throw Error("batch exceeds 2^30 points");
    }
    const lc_cell& cb = fr.cell;
    const double a[3] = {cb.ax, cb.ay, cb.az}, b[3] = {cb.bx, cb.by, cb.bz},
                 c[3] = {cb.cx, cb.cy, cb.cz}, o[3] = {cb.ox, cb.oy, cb.oz};
    lcamd::FrameSpec& s = out[f];
    if (lcamd::cell_from_vectors(a, b, c, o, s.geom) != lcamd::kOk) {
      throw Error(tag + "bad cell");
    }
    if (lcamd::plan_mesh(s.geom, hint_option(fr.cell_hint), s.plan) != lcamd::kOk) {
      throw Error(tag + "too many cells");
    }
    s.xyz = fr.xyz;
    s.mask = fr.mask;
    s.out_nn = fr.out_nn;
    s.out_d2 = fr.out_d2;
    s.out_reach = fr.out_reach;
    s.n = static_cast<int>(fr.n);
    s.k = static_cast<int>(fr.k);
  }
  return out;
}

}  // namespace

extern "C" {

int lc_gpu_knearest_batch(lc_gpu_workspace* ws, const lc_gpu_frame* frames,
                          size_t n_frames, int wait) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    const auto specs = batch_specs(frames, n_frames);
    try {
      ws->ws.impl()->engine.run(specs, wait != 0);
    } catch (const std::runtime_error& e) {
      throw linkcell::Error(e.what());
    }
  });
}

int lc_gpu_graph_capture(lc_gpu_workspace* ws, const lc_gpu_frame* frames,
                         size_t n_frames) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    const auto specs = batch_specs(frames, n_frames);
    ws->ws.impl()->engine.capture(specs);
  });
}

int lc_gpu_graph_launch(lc_gpu_workspace* ws, int wait) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] { ws->ws.impl()->engine.replay(wait != 0); });
}

int lc_gpu_graph_reset(lc_gpu_workspace* ws) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] { ws->ws.impl()->engine.reset_graph(); });
}

int lc_gpu_autotune(lc_gpu_workspace* ws, const lc_gpu_frame* frames,
                    size_t n_frames, int* best_dense_min, double* best_ms) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    const auto specs = batch_specs(frames, n_frames);
    float ms = 0.0f;
    const int best = ws->ws.impl()->engine.autotune(specs, &ms);
    if (best_dense_min) *best_dense_min = best;
    if (best_ms) *best_ms = ms;
  });
}

int lc_gpu_set_dense_min(lc_gpu_workspace* ws, int dense_min) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  ws->ws.impl()->engine.set_dense_min(dense_min);
  set_gpu_err(nullptr);
  return 0;
}

int lc_gpu_set_stats(lc_gpu_workspace* ws, int on) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  ws->ws.impl()->engine.set_stats(on != 0);
  set_gpu_err(nullptr);
  return 0;
}

int lc_gpu_set_schedule(lc_gpu_workspace* ws, int mode) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  if (mode != LC_GPU_SCHED_CELL && mode != LC_GPU_SCHED_LEARNED &&
      mode != LC_GPU_SCHED_GIVEN && mode != LC_GPU_SCHED_LEARNED_DEPTH) {
    set_gpu_err("unknown schedule");
    return 1;
  }
  ws->ws.impl()->engine.set_schedule(mode);
  set_gpu_err(nullptr);
  return 0;
}

int lc_gpu_sched_set_prediction(lc_gpu_workspace* ws, const float* kth_dist,
                                size_t n_cells) {
  if (ws == nullptr || (kth_dist == nullptr && n_cells != 0)) {
    set_gpu_err("null pointer");
    return 1;
  }
  ws->ws.impl()->engine.set_given_prediction(kth_dist,
                                             static_cast<long long>(n_cells));
  set_gpu_err(nullptr);
  return 0;
}

int lc_gpu_sched_set_tile_prediction(lc_gpu_workspace* ws,
                                     const float* kth_dist, size_t n_tiles) {
  if (ws == nullptr || (kth_dist == nullptr && n_tiles != 0)) {
    set_gpu_err("null pointer");
    return 1;
  }
  ws->ws.impl()->engine.set_given_tile_prediction(
      kth_dist, static_cast<long long>(n_tiles));
  set_gpu_err(nullptr);
  return 0;
}

int lc_gpu_sched_tile_count(lc_gpu_workspace* ws, size_t* n_tiles) {
  if (ws == nullptr || n_tiles == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    *n_tiles = static_cast<size_t>(ws->ws.impl()->engine.tile_count());
  });
}

int lc_gpu_sched_tile_labels(lc_gpu_workspace* ws, float* out, size_t n) {
  if (ws == nullptr || out == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    try {
      ws->ws.impl()->engine.tile_labels(out, static_cast<long long>(n));
    } catch (const std::runtime_error& e) {
      throw linkcell::Error(e.what());
    }
  });
}

int lc_gpu_knearest_noalloc(const double* xyz, size_t n, size_t n_frames,
                            const struct lc_cell* simbox, size_t k, const int* mask,
                            double cell_hint, int* out_nn, double* out_d2,
                            void* stream) {
  // validation in the order of the device entry points
  // (src/gpu/knearest.cpp:711-762), with no frame vector
  if (!device_present()) {
    set_gpu_err(kNoDevice);
    return 1;
  }
  if (n == 0 || n_frames == 0) {
    set_gpu_err("no points");
    return 1;
  }
  if (k == 0) {
    set_gpu_err("k must be at least 1");
    return 1;
  }
  if (n > SIZE_MAX / n_frames || n * n_frames > SIZE_MAX / k) {
    set_gpu_err("n * k overflows");
    return 1;
  }
  if (xyz == nullptr || simbox == nullptr || out_nn == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  if (n * n_frames >= static_cast<size_t>(INT_MAX) || k >= static_cast<size_t>(INT_MAX)) {
    set_gpu_err("n * n_frames and k must be below 2^31 - 1");
    return 1;
  }
  const lc_cell& cb = *simbox;
  const double a[3] = {cb.ax, cb.ay, cb.az}, b[3] = {cb.bx, cb.by, cb.bz},
               c[3] = {cb.cx, cb.cy, cb.cz}, o[3] = {cb.ox, cb.oy, cb.oz};
  lcamd::FrameSpec s;
  if (lcamd::cell_from_vectors(a, b, c, o, s.geom) != lcamd::kOk) {
    set_gpu_err("bad cell");
    return 1;
  }
  if (lcamd::plan_mesh(s.geom, hint_option(cell_hint), s.plan) != lcamd::kOk) {
    set_gpu_err("too many cells");
    return 1;
  }
  s.xyz = xyz;
  s.mask = mask;
  s.out_nn = out_nn;
  s.out_d2 = out_d2;
  s.n = static_cast<int>(n);
  s.k = static_cast<int>(k);
  try {
    lcamd::search_noalloc(s, static_cast<int>(n_frames), static_cast<hipStream_t>(stream));
  } catch (const std::exception& e) {
    set_gpu_err(e.what());
    return 1;
  }
  set_gpu_err(nullptr);
  return 0;
}

size_t lc_gpu_noalloc_scratch_bytes(void) { return lcamd::kNoAllocScratchBytes; }

int lc_gpu_sched_set_source_prediction(lc_gpu_workspace* ws, const float* kth_dist,
                                       size_t n_sources) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  ws->ws.impl()->engine.set_given_source_prediction(kth_dist,
                                                    static_cast<long long>(n_sources));
  set_gpu_err(nullptr);
  return 0;
}

int lc_gpu_sched_source_labels(lc_gpu_workspace* ws, float* out, size_t n) {
  if (ws == nullptr || out == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    try {
      ws->ws.impl()->engine.source_labels(out, static_cast<long long>(n));
    } catch (const std::runtime_error& e) {
      throw linkcell::Error(e.what());
    }
  });
}

int lc_gpu_set_dcell_capacity(lc_gpu_workspace* ws, size_t cells_per_frame) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  if (cells_per_frame == 0 || cells_per_frame > static_cast<size_t>(lcamd::kMaxCells)) {
    set_gpu_err("cell capacity must be between 1 and 16777216");
    return 1;
  }
  ws->ws.impl()->engine.set_dcell_capacity(static_cast<long long>(cells_per_frame));
  set_gpu_err(nullptr);
  return 0;
}

int lc_gpu_dcell_plan(lc_gpu_workspace* ws, const double* cell, int cell_n, double cell_hint,
                      double* out36) {
  if (ws == nullptr || cell == nullptr || out36 == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  if (cell_n != 3 && cell_n != 9 && cell_n != 12) {
    set_gpu_err("device cell must have 3, 9, or 12 doubles");
    return 1;
  }
  return guarded([&] {
    try {
      ws->ws.impl()->engine.dcell_plan(cell, cell_n, hint_option(cell_hint), out36);
    } catch (const std::runtime_error& e) {
      throw linkcell::Error(e.what());
    }
  });
}

int lc_gpu_timing_history(lc_gpu_workspace* ws, double* ms, int n_runs, int* got) {
  if (ws == nullptr || ms == nullptr || got == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    const std::vector<lcamd::Timings> h = ws->ws.impl()->engine.timing_history(n_runs);
    for (size_t i = 0; i < h.size(); ++i) {
      const lcamd::Timings& t = h[i];
      const double v[10] = {t.fold, t.scan,  t.scatter, t.sort,  t.walk,
                            t.exact, t.total, t.setup,   t.pass1, t.retry};
      for (int j = 0; j < 10; ++j) ms[10 * i + j] = v[j];
    }
    *got = static_cast<int>(h.size());
  });
}

int lc_gpu_sched_dump(lc_gpu_workspace* ws, size_t n_tiles, size_t n_cells,
                      unsigned short* x_collect, unsigned short* x_predict,
                      float* caps, int* tile_cell, double* tile_box, int* count) {
  if (ws == nullptr || x_collect == nullptr || tile_cell == nullptr ||
      tile_box == nullptr || count == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    try {
      ws->ws.impl()->engine.sched_dump(x_collect, x_predict, caps, tile_cell, tile_box,
                                       count, static_cast<long long>(n_tiles),
                                       static_cast<long long>(n_cells));
    } catch (const std::runtime_error& e) {
      throw linkcell::Error(e.what());
    }
  });
}

int lc_gpu_last_retried(lc_gpu_workspace* ws, int* retried) {
  if (ws == nullptr || retried == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    *retried = ws->ws.impl()->engine.last_timings().retried;
  });
}

int lc_gpu_sched_collect(lc_gpu_workspace* ws, size_t* n_added) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    const long long a = ws->ws.impl()->engine.sched_collect();
    if (n_added) *n_added = static_cast<size_t>(a);
  });
}

int lc_gpu_sched_samples(lc_gpu_workspace* ws, size_t* n) {
  if (ws == nullptr || n == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    *n = static_cast<size_t>(ws->ws.impl()->engine.predictor().samples());
  });
}

int lc_gpu_sched_clear(lc_gpu_workspace* ws) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] { ws->ws.impl()->engine.predictor().clear_samples(); });
}

int lc_gpu_sched_train(lc_gpu_workspace* ws, int epochs, int batch, double lr,
                       unsigned seed, double quantile, double* last_loss,
                       double* margin) {
  if (ws == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    lcamd::Engine& e = ws->ws.impl()->engine;
    try {
      const double l = e.predictor().train(epochs, batch, lr, seed, quantile, e.stream());
      if (last_loss) *last_loss = l;
      if (margin) *margin = e.predictor().margin();
    } catch (const std::runtime_error& err) {
      throw linkcell::Error(err.what());
    }
  });
}

int lc_gpu_sched_get_params(lc_gpu_workspace* ws, float* params, size_t n,
                            double* margin) {
  if (ws == nullptr || params == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  if (n != static_cast<size_t>(lcamd::kParams)) {
    set_gpu_err("parameter count must be LC_GPU_SCHED_PARAMS");
    return 1;
  }
  return guarded([&] {
    lcamd::Predictor& p = ws->ws.impl()->engine.predictor();
    try {
      p.get_params(params);
    } catch (const std::runtime_error& err) {
      throw linkcell::Error(err.what());
    }
    if (margin) *margin = p.margin();
  });
}

int lc_gpu_sched_set_params(lc_gpu_workspace* ws, const float* params, size_t n,
                            double margin) {
  if (ws == nullptr || params == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  if (n != static_cast<size_t>(lcamd::kParams)) {
    set_gpu_err("parameter count must be LC_GPU_SCHED_PARAMS");
    return 1;
  }
  return guarded([&] {
    lcamd::Engine& e = ws->ws.impl()->engine;
    e.predictor().set_params(params, margin, e.stream());
  });
}

int lc_gpu_sched_forward(lc_gpu_workspace* ws, const float* x, size_t n,
                         float* y) {
  if (ws == nullptr || x == nullptr || y == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    lcamd::Engine& e = ws->ws.impl()->engine;
    try {
      e.predictor().forward_host(x, static_cast<long long>(n), y, e.stream());
    } catch (const std::runtime_error& err) {
      throw linkcell::Error(err.what());
    }
  });
}

int lc_gpu_sched_gradient(lc_gpu_workspace* ws, const float* x, const float* t,
                          const float* w, size_t n, float* grad, double* loss) {
  if (ws == nullptr || x == nullptr || t == nullptr || w == nullptr ||
      grad == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    lcamd::Engine& e = ws->ws.impl()->engine;
    try {
      const double l = e.predictor().grad_host(x, t, w, static_cast<long long>(n),
                                               grad, e.stream());
      if (loss) *loss = l;
    } catch (const std::runtime_error& err) {
      throw linkcell::Error(err.what());
    }
  });
}

int lc_gpu_sched_adam(lc_gpu_workspace* ws, const float* x, const float* t,
                      const float* w, size_t n, int steps, double lr,
                      double* loss) {
  if (ws == nullptr || x == nullptr || t == nullptr || w == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] {
    lcamd::Engine& e = ws->ws.impl()->engine;
    try {
      const double l = e.predictor().adam_host(x, t, w, static_cast<long long>(n),
                                               steps, lr, e.stream());
      if (loss) *loss = l;
    } catch (const std::runtime_error& err) {
      throw linkcell::Error(err.what());
    }
  });
}

int lc_gpu_last_stats(lc_gpu_workspace* ws, unsigned long long* out, int n) {
  if (ws == nullptr || out == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  return guarded([&] { ws->ws.impl()->engine.read_stats(out, n); });
}

int lc_gpu_alloc(void** ptr, size_t bytes) {
  if (ptr == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  if (!device_present()) {
    set_gpu_err(kNoDevice);
    return 1;
  }
  const hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    set_gpu_err(hipGetErrorString(e));
    return 1;
  }
  set_gpu_err(nullptr);
  return 0;
}

void lc_gpu_free(void* ptr) {
  if (ptr != nullptr && device_present()) (void)hipFree(ptr);
}

int lc_gpu_memcpy(void* dst, const void* src, size_t bytes, int kind) {
  if (bytes == 0) {
    set_gpu_err(nullptr);
    return 0;
  }
  if (dst == nullptr || src == nullptr) {
    set_gpu_err("null pointer");
    return 1;
  }
  if (!device_present()) {
    set_gpu_err(kNoDevice);
    return 1;
  }
  if (kind < 0 || kind > 4) {
    set_gpu_err("unknown memcpy kind");
    return 1;
  }
  const hipError_t e = hipMemcpy(dst, src, bytes, static_cast<hipMemcpyKind>(kind));
  if (e != hipSuccess) {
    set_gpu_err(hipGetErrorString(e));
    return 1;
  }
  set_gpu_err(nullptr);
  return 0;
}

int lc_gpu_fill_i32(void* ptr, int value, size_t n) {
  if (ptr == nullptr && n != 0) {
    set_gpu_err("null pointer");
    return 1;
  }
  if (!device_present()) {
    set_gpu_err(kNoDevice);
    return 1;
  }
  if (value != 0 && value != -1) {
    set_gpu_err("lc_gpu_fill_i32 only supports 0 and -1");
    return 1;
  }
  const hipError_t e = hipMemset(ptr, value == 0 ? 0 : 0xFF, n * sizeof(int));
  if (e != hipSuccess) {
    set_gpu_err(hipGetErrorString(e));
    return 1;
  }
  set_gpu_err(nullptr);
  return 0;
}

}  // extern "C"

// ------------------------------------------------------------------
// Host-pointer C API (src/capi.rs:79-297), staged through the device.

namespace {

enum CpuErr { kCpuOk = 0 };

const char* kMsgZeroK = "k must be at least 1";
const char* kMsgBadBox = "singular or non-positive cell";
const char* kMsgEmpty = "no points";
const char* kMsgOverflow = "n * k overflows";
const char* kMsgTooMany = "linked-cell mesh is too fine";

int cpu_fail(const char* m) {
  g_cpu_err = m;
  g_cpu_err_set = true;
  return 1;
}

struct HostStage {
  lcamd::Engine engine;
  void* dxyz = nullptr;
  size_t dxyz_bytes = 0;
  void* dmask = nullptr;
  size_t dmask_bytes = 0;
  void* dnn = nullptr;
  size_t dnn_bytes = 0;
  void* dd2 = nullptr;
  size_t dd2_bytes = 0;
};

void* grow_dev(void*& p, size_t& cap, size_t bytes) {
  if (bytes > cap || p == nullptr) {
    if (p != nullptr) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc(&p, bytes < 256 ? 256 : bytes) != hipSuccess) {
      throw std::runtime_error("hipMalloc failed");
    }
    cap = bytes < 256 ? 256 : bytes;
  }
  return p;
}

HostStage& host_stage() {
  // one per thread, kept for the life of the process
  thread_local HostStage* st = new HostStage();
  return *st;
}

void fill_rows(int* nn, double* d2, size_t count) {
  const uint64_t nanbits = 0x7FF8000000000000ull;
  double nanv;
  std::memcpy(&nanv, &nanbits, sizeof nanv);
  for (size_t t = 0; t < count; ++t) {
    nn[t] = -1;
    if (d2) d2[t] = nanv;
  }
}

}  // namespace

extern "C" {

const char* lc_last_error(void) {
  return g_cpu_err_set ? g_cpu_err.c_str() : nullptr;
}

const char* lc_version(void) { return "0.3.3"; }

int lc_knearest_many(const double* xyz, size_t n, size_t n_frames,
                     const struct lc_cell* simbox, size_t k, const int* mask,
                     double cell_hint, int* out_nn, double* out_d2) {
  // validation order of src/capi.rs:233-265
  if (n == 0 || n_frames == 0) return cpu_fail(kMsgEmpty);
  if (k == 0) return cpu_fail(kMsgZeroK);
  if (xyz == nullptr || simbox == nullptr || out_nn == nullptr) {
    return cpu_fail("null pointer");
  }
  if (n > SIZE_MAX / n_frames) return cpu_fail(kMsgOverflow);
  const size_t n_pts = n * n_frames;
  if (n_pts > SIZE_MAX / k) return cpu_fail(kMsgOverflow);
  const size_t need = n_pts * k;
  const size_t max_xyz = static_cast<size_t>(PTRDIFF_MAX) / 3;
  const size_t max_out = static_cast<size_t>(PTRDIFF_MAX) / sizeof(int);
  if (n_pts > max_xyz || need > max_out) return cpu_fail(kMsgOverflow);
  const lc_cell& cb = *simbox;
  const double a[3] = {cb.ax, cb.ay, cb.az}, b[3] = {cb.bx, cb.by, cb.bz},
               c[3] = {cb.cx, cb.cy, cb.cz}, o[3] = {cb.ox, cb.oy, cb.oz};
  lcamd::CellGeom geom{};
  if (lcamd::cell_from_vectors(a, b, c, o, geom) != lcamd::kOk) {
    return cpu_fail(kMsgBadBox);
  }
  // write_one fills before searching (src/knearest.rs:293-296); a search
  // with no active point returns before planning the mesh (:370-375).
  bool any = mask == nullptr;
  for (size_t i = 0; !any && i < n; ++i) any = mask[i] != 0;
  if (!any) {
    fill_rows(out_nn, out_d2, need);
    g_cpu_err_set = false;
    return 0;
  }
  lcamd::MeshPlan plan{};
  if (lcamd::plan_mesh(geom, hint_option(cell_hint), plan) != lcamd::kOk) {
    fill_rows(out_nn, out_d2, n * k);  // frame 0 was filled, then failed
    return cpu_fail(kMsgTooMany);
  }
  if (!device_present()) return cpu_fail(kNoDevice);
  try {
    HostStage& hs = host_stage();
    hipStream_t st = hs.engine.stream();
    double* dxyz = static_cast<double*>(
        grow_dev(hs.dxyz, hs.dxyz_bytes, n_pts * 3 * sizeof(double)));
    int* dmask = nullptr;
    if (mask != nullptr) {
      dmask = static_cast<int*>(grow_dev(hs.dmask, hs.dmask_bytes, n * sizeof(int)));
    }
    int* dnn = static_cast<int*>(grow_dev(hs.dnn, hs.dnn_bytes, need * sizeof(int)));
    double* dd2 = nullptr;
    if (out_d2 != nullptr) {
      dd2 = static_cast<double*>(grow_dev(hs.dd2, hs.dd2_bytes, need * sizeof(double)));
    }
    if (hipMemcpyAsync(dxyz, xyz, n_pts * 3 * sizeof(double),
                       hipMemcpyHostToDevice, st) != hipSuccess ||
        (dmask && hipMemcpyAsync(dmask, mask, n * sizeof(int),
                                 hipMemcpyHostToDevice, st) != hipSuccess)) {
      return cpu_fail("HtoD copy failed");
    }
    std::vector<lcamd::FrameSpec> frames(n_frames);
    for (size_t f = 0; f < n_frames; ++f) {
      lcamd::FrameSpec& s = frames[f];
      s.xyz = dxyz + f * n * 3;
      s.mask = dmask;
      s.out_nn = dnn + f * n * k;
      s.out_d2 = dd2 ? dd2 + f * n * k : nullptr;
      s.n = static_cast<int>(n);
      s.k = static_cast<int>(k);
      s.geom = geom;
      s.plan = plan;
    }
    if (n_pts >= static_cast<size_t>(INT_MAX) || k >= static_cast<size_t>(INT_MAX)) {
      return cpu_fail("n * n_frames and k must be below 2^31 - 1");
    }
    hs.engine.run(frames, false);
    if (hipMemcpyAsync(out_nn, dnn, need * sizeof(int), hipMemcpyDeviceToHost,
                       st) != hipSuccess ||
        (dd2 && hipMemcpyAsync(out_d2, dd2, need * sizeof(double),
                               hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess) {
      return cpu_fail("DtoH copy failed");
    }
  } catch (const std::exception& e) {
    return cpu_fail(e.what());
  }
  g_cpu_err_set = false;
  return 0;
}

int lc_knearest(const double* xyz, size_t n, const struct lc_cell* simbox,
                size_t k, const int* mask, double cell_hint, int* out_nn) {
  return lc_knearest_many(xyz, n, 1, simbox, k, mask, cell_hint, out_nn, nullptr);
}

int lc_knearest_d2(const double* xyz, size_t n, const struct lc_cell* simbox,
                   size_t k, const int* mask, double cell_hint, int* out_nn,
                   double* out_d2) {
  return lc_knearest_many(xyz, n, 1, simbox, k, mask, cell_hint, out_nn, out_d2);
}

}  // extern "C"
