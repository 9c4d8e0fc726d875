// Host-side engine behind linkcell::gpu::Workspace and the lc_gpu_* waist.
//
// One engine owns one HIP stream and grow-only device buffers (the
// reference Workspace::Impl, src/gpu/knearest.cpp:571-674). A call runs a
// batch of frames; each frame has its own points, cell, mesh and k, so
// the shared-cell entry points (src/gpu/knearest.cpp:705-816) and the
// batched uneven-frame path use the same kernels.
#pragma once

#include "lc_cell.hpp"

#include <hip/hip_runtime.h>

#include <cstddef>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace lcamd {

// Device-side frame descriptor.
struct FrameDesc {
  const double* xyz;  // n x 3, AoS, device
  const int* mask;    // n ints or null, device
  int* out_nn;        // n x k, device
  double* out_d2;     // n x k or null, device
  int* out_reach;     // n or null (final shell per source; diagnostics)
  double h[9];
  double hinv[9];
  double origin[3];
  double widths[3];
  double cell_min;
  long long pt_base;    // first global point id of the frame
  long long cell_base;  // first global cell id of the frame
  int n;
  int k;
  int nx, ny, nz;
  int max_reach;
  int r_safe;  // largest reach free of stencil wrap (see safe_reach)
  int ortho;
  int kc;  // register-heap bucket for the fast walk, 0 = exact walk only
  int skip;  // 1: the frame's device-side plan failed; every point counts as masked
};

// Where the source tiles of a search live (device pointers of the engine);
// read by the learned scheduler's feature map.
struct TileSrc {
  const FrameDesc* fr;
  const long long* fcb;
  int F;
  const int* count;      // occupancy per cell of the batch
  const int* off;        // cell offsets (ncells + 1)
  const int* tile_off;   // first tile per cell (ncells + 1)
  const int* tile_cell;  // home cell per tile
  const double* tbox;    // per tile: lx, hx, ly, hy, lz, hz, pad, pad
  long long ncells;
};

class Predictor;

// Host description of one frame of a call.
struct FrameSpec {
  const double* xyz = nullptr;
  const int* mask = nullptr;
  int* out_nn = nullptr;
  double* out_d2 = nullptr;
  int* out_reach = nullptr;
  int n = 0;
  int k = 0;
  CellGeom geom{};
  MeshPlan plan{};
};

// Per-call kernel timings in milliseconds (hipEvent pairs on the engine
// stream), valid after a call made with timing on and a wait.
struct Timings {
  float fold = 0, scan = 0, scatter = 0, sort = 0, walk = 0, exact = 0,
        total = 0;
  // parts of walk: schedule set-up (learned caps, depth order), the walk
  // kernels, and what follows them before the exact walk
  float setup = 0, pass1 = 0, retry = 0;
  int deferred = 0;  // sources finished by the exact-order walk
  int retried = 0;   // rows walked a second time without their predicted cap
};

enum class WalkMode : int {
  kAuto = 0,       // canonical fast walk, exact-order walk for the rest
  kExactOnly = 1,  // every source through the exact-order walk
};

class Engine {
 public:
  Engine();
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  hipStream_t stream();
  // Waits for the stream; reports a pending device-cell plan failure.
  void sync();
  // Runs a batch. Throws std::runtime_error on HIP failures.
  void run(const std::vector<FrameSpec>& frames, bool wait);
  // n_frames frames sharing a device-resident cell of cell_n doubles (3
  // ortho lengths, 9 rows a b c, 12 with the origin): the cell is inverted
  // and the mesh planned on the device, with the same arithmetic as the
  // host plan, so nothing is read back and the sequence has no host copy
  // (a graph can capture it once buffers are sized). Buffers are sized for
  // dcell_capacity() cells per frame. A bad cell or a plan above that
  // capacity makes every point of the search count as masked (-1 / NaN
  // rows) and is reported as "bad cell" / "too many cells" after a wait
  // (at once when wait is set, else by the next sync()). spec gives the
  // points, outputs, mask, n and k of frame 0; frame f follows at strides
  // n points and n * k outputs.
  void run_dcell(const FrameSpec& spec, int n_frames, const double* cell, int cell_n,
                 double hint, bool wait);
  void set_dcell_capacity(long long cells) { dcell_cap_ = cells < 1 ? 1 : cells; }
  long long dcell_capacity() const { return dcell_cap_; }
  // Diagnostics: the device's plan of a device cell as 36 doubles (H by
  // columns, Hinv, origin, widths, ortho, cell_min, nx, ny, nz, ncell,
  // max_reach, status); waits.
  void dcell_plan(const double* cell, int cell_n, double hint, double* out36);

  // hipGraph: run the batch once (sizing every buffer), then capture the
  // same batch into a graph that replay() relaunches. The frames' device
  // pointers and sizes are baked in; callers refill the same buffers.
  void capture(const std::vector<FrameSpec>& frames);
  void replay(bool wait);
  bool has_graph() const { return graph_exec_ != nullptr; }
  void reset_graph();

  // Times the batch under each walk split (see set_dense_min) and keeps
  // the fastest for the batch's size bucket; later runs of that bucket use
  // it. Returns the best split and its pipeline time in milliseconds.
  int autotune(const std::vector<FrameSpec>& frames, float* best_ms);

  void set_timing(bool on) { timing_ = on; }
  const Timings& last_timings();
  // Stage timings of up to n of the most recent timed runs, oldest first
  // (at most kTimingRing); waits for the latest. Deferred and retried
  // counts are those of the last run only.
  std::vector<Timings> timing_history(int n);
  void set_mode(WalkMode m) { mode_ = m; }
  // Home cells with at least this many sources take the wave-per-cell
  // walk, the rest the lane-per-source walk.
  // An explicit value overrides the tuned table; 0 returns to it; -1
  // selects the row walk (one 16-lane row per 16-source tile).
  void set_dense_min(int v) { dense_min_override_ = v < -2 ? 0 : v; }
  int dense_min_for(const std::vector<FrameSpec>& frames) const;
  // Walk schedule: 0 = plain; 1 = the learned predictor's k-th-neighbour
  // distance per source tile as the row walk's threshold cap; 3 = the
  // same caps plus the depth schedule (each row lists the shells below its
  // predicted depth in one pass, rows packed by predicted depth); 2 = a
  // caller-given distance per cell (device
  // array of floats, one per cell of the batch in frame order). The
  // distance seeds the walk's pruning threshold (checked exactly at every
  // shell end; a source it does not cover is finished by the exact walk)
  // and orders the lane-per-source walk by predicted depth.
  void set_schedule(int mode) { sched_mode_ = mode; }
  int schedule() const { return sched_mode_; }
  void set_given_prediction(const float* kth, long long n) {
    given_pred_ = kth;
    given_pred_n_ = n;
    given_tile_ = nullptr;
    given_tile_n_ = 0;
  }
  // Per source tile instead of per cell (the engine's tile order, see
  // tile_labels); takes precedence over a per-cell prediction.
  void set_given_tile_prediction(const float* kth, long long n) {
    given_tile_ = kth;
    given_tile_n_ = n;
  }
  // Per source slot (the walk's cell-sorted order, see source_labels);
  // takes precedence over both; null clears it.
  void set_given_source_prediction(const float* kth, long long n) {
    given_src_ = kth;
    given_src_n_ = n;
  }
  // For every source slot of the last search, its final k-th-neighbour
  // distance rounded up to float (+inf without one), into the device
  // array out (at least as many floats as the search's active sources).
  void source_labels(float* out, long long n);
  // Source tiles (16 consecutive slots of one home cell in the walk's
  // cell-sorted order) of the last search; waits for it.
  long long tile_count();
  // For every tile of the last search, the largest final k-th-neighbour
  // distance among its sources (from the frames' out_d2), rounded up to
  // float, into the device array out (at least tile_count() floats).
  void tile_labels(float* out, long long n);
  // The workspace's learned scheduler (created on first use).
  Predictor& predictor();
  // Append the source tiles of the last search (run with out_d2) to the
  // predictor's training set; returns the samples added.
  long long sched_collect();
  // Diagnostics of the learned scheduler on the last search (host
  // arrays, nt >= tile_count(), nc >= the search's cells): the inputs the
  // training set takes for every tile, the inputs and caps of inference
  // (when trained), and the tile layout they come from: home cell, box
  // (lx, hx, ly, hy, lz, hz, pad, pad) and cell occupancy.
  void sched_dump(unsigned short* x_collect, unsigned short* x_predict, float* caps,
                  int* tile_cell, double* tile_box, int* count, long long nt,
                  long long nc);
  // Device work counters of the walks (diagnostics).
  void set_stats(bool on) { stats_on_ = on; }
  void read_stats(unsigned long long* out, int n);

 private:
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  void grow(Buf& b, size_t bytes);
  void free_all();

  hipStream_t stream_ = nullptr;
  int device_ = -1;
  Buf frames_, fpt_, fcb_, count_, status_, counters_, off_, tile_off_,
      item_off_, items_, cellslot_, tmp_key_, cand_, scell_,
      byidx_, tbox_, cbox_, defer_, heap_d2_, heap_idx_, flags_, stats_,
      sched_, sched_cnt_, tile_cell_;
  int sched_mode_ = 0;
  // cells per frame the device-cell path sizes its buffers for (the
  // reference's device limit, src/gpu/knearest.cpp:23)
  long long dcell_cap_ = 262144;
  bool dcell_pending_ = false;  // a device-cell search's status is unread
  void check_dcell_status();
  const float* given_pred_ = nullptr;
  long long given_pred_n_ = 0;
  const float* given_tile_ = nullptr;
  long long given_tile_n_ = 0;
  const float* given_src_ = nullptr;
  long long given_src_n_ = 0;
  int last_F_ = 0;
  long long last_ntot_ = 0, last_ctot_ = 0;
  std::unique_ptr<Predictor> pred_;
  Buf pred_tile_, labels_, tile_perm_;
  Buf cap_list_;  // tiles the learned caps are inferred for, then their count
  TileSrc tile_src(long long ctot, int F);
  void* staging_ = nullptr;
  size_t staging_bytes_ = 0;
  hipEvent_t staging_ev_ = nullptr;
  bool staging_pending_ = false;
  bool timing_ = false;
  bool timings_pending_ = false;
  WalkMode mode_ = WalkMode::kAuto;
  int dense_min_override_ = 0;
  bool capturing_ = false;
  hipGraphExec_t graph_exec_ = nullptr;
  void* graph_staging_ = nullptr;
  size_t graph_staging_bytes_ = 0;
  std::vector<std::pair<std::string, int>> tuned_;
  bool stats_on_ = false;
  long long stats_n_ = 0;
  // timing events of the last kTimingRing timed runs (slot ring_pos_ - 1
  // is the latest), so a timed loop of many calls can be read afterwards
  static constexpr int kTimingRing = 64;
  hipEvent_t ring_[kTimingRing][10] = {};
  long long ring_pos_ = 0;
  hipEvent_t* ev_ = nullptr;  // the events of the run being recorded
  Timings timings_{};
};

// Heap-capacity bucket of the register-resident fast walk for k.
int fast_bucket(int k);

// Search without allocating: n_frames frames sharing frame0's cell, k and
// mask, frame f at strides n (points) and n * k (outputs), every buffer in
// the library's fixed device scratch (kNoAllocScratchBytes per GPU), on
// stream st, not waited for. Throws std::runtime_error when the search
// needs more scratch (noalloc_bytes) or on a HIP failure. Calls that
// overlap on one GPU share the scratch and race.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
size_t noalloc_bytes(int n_frames, const FrameSpec& frame0);
void search_noalloc(const FrameSpec& frame0, int n_frames, hipStream_t st);

}  // namespace lcamd
