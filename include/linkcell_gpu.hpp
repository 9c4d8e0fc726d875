// Device-resident periodic k-nearest search: linkcell::gpu::Workspace.
//
// Same class, methods and defaults as the reference
// (include/linkcell_gpu.hpp:29-89 of linkcell 0.3.3). Pointers are HIP
// device pointers, queue() is a hipStream_t, methods throw
// linkcell::Error. Any k >= 1; results are bit-identical to the CPU walk.
#ifndef LINKCELL_GPU_HPP
#define LINKCELL_GPU_HPP

#include "linkcell.hpp"

#include <cstddef>

// This build always has the device path. The reference gates the class on
// LINKCELL_HAS_GPULITE (include/linkcell_gpu.hpp:88); define it too so
// sources written against that macro compile unchanged.
#ifndef LINKCELL_HAS_HIP
#define LINKCELL_HAS_HIP 1
#endif
#ifndef LINKCELL_HAS_GPULITE
#define LINKCELL_HAS_GPULITE 1
#endif

namespace linkcell {
namespace gpu {

bool available();

class Workspace {
 public:
  Workspace();
  ~Workspace();
  Workspace(Workspace&&) noexcept;
  Workspace& operator=(Workspace&&) noexcept;
  Workspace(const Workspace&) = delete;
  Workspace& operator=(const Workspace&) = delete;

  void knearest_into(const double* xyz, std::size_t n, const Cell& cell,
                     std::size_t k, int* out, std::size_t out_len,
                     const int* mask = nullptr, double cell_hint = 0.0,
                     double* out_d2 = nullptr);

  /// frameBox: optional host array of nFrames * 3 orthorhombic lengths,
  /// one box per frame (the cell must then be orthorhombic); each frame
  /// is searched in its own box with the cell's origin.
  void knearest_into_many(const double* xyz, std::size_t n,
                          std::size_t nFrames, const Cell& cell, std::size_t k,
                          int* out, std::size_t out_len,
                          const int* mask = nullptr, double cell_hint = 0.0,
                          bool wait = true, const double* frameBox = nullptr,
                          double* out_d2 = nullptr);

  void knearest_into_many_dcell(const double* xyz, std::size_t n,
                                std::size_t nFrames, const double* cell,
                                int cell_n, std::size_t k, int* out,
                                std::size_t out_len, const int* mask = nullptr,
                                double cell_hint = 0.0, bool wait = true,
                                double* out_d2 = nullptr);

  void* queue();
  void wait();

  struct Impl;
  /// Build-specific access for the C waist's extensions.
  Impl* impl() noexcept { return impl_; }

 private:
  Impl* impl_;
};

}  // namespace gpu
}  // namespace linkcell

#endif
