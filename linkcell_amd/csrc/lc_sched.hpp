// Learned shell scheduler: a small MLP that predicts, for every source tile
// of a search (16 consecutive sources of one home cell), the largest final
// k-th-neighbour distance among its sources. The row walk starts from that
// distance as its pruning cap; a row whose cap proves too small is walked
// again without it, so the prediction changes the work done, never the
// results.
//
// Inputs (160 per tile, the MFMA K dimension): log2(1 + occupancy) / 8 of
// the home cell's 5x5x5 periodic neighbourhood in (dz, dy, dx) order (125
// values), then the tile: log2(box edge / cell_min + 1/256) / 8 per axis,
// the box centre's fractional position inside the home cell per axis, the
// fill (sources / 16) and log2(k) / 8; the last 27 inputs are zero. Network
// 160 -> 64 -> 64 -> 1 with ReLU, BF16 weights and activations into MFMA
// (v_mfma_f32_16x16x32_bf16), FP32 accumulation, FP32 master weights and
// Adam moments. Target: log2(label / cell_min). Loss: weighted mean of
// 0.5 (y - target)^2.
//
// The predictor has no counterpart in the reference; its labels come from
// the plain walk (src/knearest.rs:431-475 restated on the GPU).
#pragma once

#include "lc_engine.hpp"

#include <hip/hip_runtime.h>

#include <vector>

namespace lcamd {

constexpr int kFeat = 160;  // inputs per tile (125 occupancies + 8 + 27 zero)
constexpr int kOcc = 125;   // 5x5x5 occupancies
constexpr int kHid = 64;    // hidden width
// FP32 parameter vector: W1 [kFeat][kHid] (input-major), b1, W2 [kHid][kHid]
// (input-major), b2, w3, b3.
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kFeat * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
constexpr int kParams = kOffB3 + 1;
constexpr int kMlpRows = 64;  // samples per 256-thread block

class Predictor {
 public:
  Predictor() = default;
  ~Predictor();
  Predictor(const Predictor&) = delete;
  Predictor& operator=(const Predictor&) = delete;

  bool trained() const { return trained_; }
  double margin() const { return margin_; }

  // Learned caps of every tile of a search: out[T] = cell_min *
  // 2^(y_T + margin) for T below the search's tile count (read on the
  // device), tcap the caller's upper bound of that count.
  // xdump (optional, device): the BF16 inputs of every tile as the
  // forward pass reads them (diagnostics). list / list_n (optional,
  // device): only the tiles list[0 .. *list_n) (at most tcap of them).
  void predict_caps(const TileSrc& src, long long tcap, float* out,
                    hipStream_t st, unsigned short* xdump = nullptr,
                    const int* list = nullptr, const int* list_n = nullptr);

  // Append the tiles of a finished search with a finite label to the
  // training set (labels: device floats per tile, from the engine's
  // tile_labels). Returns the number of samples added.
  long long collect(const TileSrc& src, long long ntiles, const float* labels,
                    hipStream_t st);
  long long samples() const { return n_samples_; }
  // The BF16 inputs the training set would take for every tile of a
  // search, into the device array x (ntiles x kFeat; diagnostics).
  void collect_features(const TileSrc& src, long long ntiles, unsigned short* x,
                        hipStream_t st);
  void clear_samples();

  // Fresh initialisation from seed, then epochs of minibatch Adam over the
  // training set; the margin becomes the given quantile of (target -
  // prediction) over the set. Returns the last epoch's mean loss.
  double train(int epochs, int batch, double lr, unsigned seed,
               double quantile, hipStream_t st);

  // Host-staged access for tests and persistence.
  void get_params(float* out) const;
  void set_params(const float* in, double margin, hipStream_t st);
  // y = network(x) for n host samples of kFeat inputs (rounded to BF16).
  void forward_host(const float* x, long long n, float* y, hipStream_t st);
  // Gradient of the loss at the current parameters over n host samples
  // (targets t, weights w) and the loss itself; nothing is updated.
  double grad_host(const float* x, const float* t, const float* w, long long n,
                   float* grad, hipStream_t st);
  // steps full-batch Adam steps on n host samples from the current
  // parameters (moments restart at zero). Returns the last loss.
  double adam_host(const float* x, const float* t, const float* w, long long n,
                   int steps, double lr, hipStream_t st);

 private:
  struct Buf {
    void* p = nullptr;
    size_t bytes = 0;
  };
  void grow(Buf& b, size_t bytes, bool keep = false);
  void ensure_params();
  void upload_params(const float* host, hipStream_t st);
  // one Adam step over the rows idx(i), i < nrows, of the sample arrays
  void step(const unsigned short* X, const float* t, const float* w,
            long long n, long long nrows, long long stride, long long start,
            double lr, int t_step, float inv_wsum, hipStream_t st,
            bool update, float* grad_out);

  Buf master_, m_, v_, w1t_, w2t_, w2_, partial_, loss_, gout_;
  Buf X_, T_, W_;  // training set: features (BF16), targets, weights
  Buf tmpX_, tmpT_, tmpW_, tmpY_;
  long long n_samples_ = 0;
  bool trained_ = false;
  bool params_ready_ = false;
  double margin_ = 0.0;
};

}  // namespace lcamd
