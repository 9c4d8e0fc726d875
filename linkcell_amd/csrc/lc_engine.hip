// Linked-cell k-nearest pipeline on gfx950 (MI355X).
//
//   fold_count  fold every point into its primary cell, bin it, count
//               occupants with one atomic per point (the slot it returns
//               is the point's arrival rank in the cell)
//   scan        single-pass decoupled look-back exclusive scan over cell
//               counts; one packed 64-bit word per tile carries the
//               occupant prefix and the 64-source chunk prefix together
//   scatter     arrival-order scatter of point ids into cell segments
//   cell_sort   per cell chunk: rank by original index, re-fold, write
//               SoA FP64 coordinates, ids and home cells in cell order
//               with ascending ids inside each cell
//   walk_fast   one lane per source; Chebyshev shells with a canonical
//               (d2, index) register heap and the exact stop test
//   walk_exact  exact-order emulation of the reference KHeap for the
//               sources the fast walk defers (boundary ties, stencil
//               wrap, k above the register buckets, non-finite input)
//
// Reference: src/knearest.rs:351-486 (CPU walk, the parity target) and
// src/gpu/knearest.cpp:192-441 (the CUDA kernels this file replaces).
#include "lc_engine.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <stdexcept>
#include <string>

#pragma clang fp contract(off)

namespace lcamd {

namespace {

constexpr int kWave = 64;
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr unsigned long long kFlagAgg = 1ull << 62;
constexpr unsigned long long kFlagPre = 2ull << 62;
constexpr unsigned long long kValMask = (1ull << 62) - 1;
constexpr int kFrameNonFinite = 1;
constexpr long long kInfBits = 0x7ff0000000000000LL;
constexpr long long kUnusedBits = LLONG_MIN;
constexpr unsigned long long kNanBits = 0x7ff8000000000000ull;

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }
}

__device__ __forceinline__ int find_frame(const long long* base, int F,
                                          long long x) {
  int lo = 0, hi = F - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (base[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__device__ __forceinline__ CellGeom geom_of(const FrameDesc& d) {
  CellGeom g;
  for (int t = 0; t < 9; ++t) {
    g.h[t] = d.h[t];
    g.hinv[t] = d.hinv[t];
  }
  for (int t = 0; t < 3; ++t) {
    g.origin[t] = d.origin[t];
    g.widths[t] = d.widths[t];
  }
  g.ortho = d.ortho;
  return g;
}

__device__ __forceinline__ double nan_value() {
  return __longlong_as_double(static_cast<long long>(kNanBits));
}

// ---------------------------------------------------------------- fill

__global__ void k_zero_u32(unsigned* p, long long n) {
  const long long t = blockIdx.x * static_cast<long long>(blockDim.x) +
                      threadIdx.x;
  if (t < n) p[t] = 0u;
}

// ---------------------------------------------------------- fold/count

// Fold + bin (src/knearest.rs:405-412; Cell::fractional/cartesian
// src/cell.rs:176-207). Masked points get their -1 / NaN row here
// (src/knearest.rs:293-296), because no walk touches them.
__global__ __launch_bounds__(256) void k_fold_count(
    const FrameDesc* __restrict__ fr, const long long* __restrict__ fpt,
    int F, long long ntot, int* __restrict__ count,
    int2* __restrict__ cellslot, int* __restrict__ fflags) {
  const long long p = blockIdx.x * static_cast<long long>(blockDim.x) +
                      threadIdx.x;
  if (p >= ntot) return;
  const int f = find_frame(fpt, F, p);
  const FrameDesc& d = fr[f];
  const int i = static_cast<int>(p - d.pt_base);
  if (d.mask != nullptr && d.mask[i] == 0) {
    cellslot[p] = make_int2(-1, 0);
    const size_t row = static_cast<size_t>(i) * d.k;
    for (int t = 0; t < d.k; ++t) {
      d.out_nn[row + t] = -1;
      if (d.out_d2) d.out_d2[row + t] = nan_value();
    }
    if (d.out_reach) d.out_reach[i] = 0;
    return;
  }
  const CellGeom g = geom_of(d);
  const double* r = d.xyz + 3 * static_cast<size_t>(i);
  double s0, s1, s2, fx, fy, fz;
  fold_point(g, r[0], r[1], r[2], s0, s1, s2, fx, fy, fz);
  if (!(isfinite(fx) && isfinite(fy) && isfinite(fz))) {
    atomicOr(&fflags[f], kFrameNonFinite);
  }
  const int bx = bin_axis(s0, d.nx);
  const int by = bin_axis(s1, d.ny);
  const int bz = bin_axis(s2, d.nz);
  const long long gcell =
      d.cell_base + (static_cast<long long>(bz) * d.ny + by) * d.nx + bx;
  const int slot = atomicAdd(&count[gcell], 1);
  cellslot[p] = make_int2(static_cast<int>(gcell), slot);
}

// ---------------------------------------------------------------- scan

__device__ __forceinline__ unsigned long long pack_cell(int cnt) {
  const unsigned long long ch = static_cast<unsigned long long>((cnt + 63) >> 6);
  return static_cast<unsigned long long>(cnt) | (ch << 31);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(
    unsigned long long v) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

// Single-pass exclusive scan with decoupled look-back. Tiles take ids
// from an atomic counter, so a tile only ever waits on tiles that are
// already resident. Each tile publishes its aggregate at once and its
// inclusive prefix after the look-back; the whole first wave inspects 64
// predecessors per step with one ballot.
__global__ __launch_bounds__(kScanThreads) void k_scan(
    const int* __restrict__ count, long long ncells, int* __restrict__ off,
    int* __restrict__ item_off, unsigned long long* status,
    int* tile_counter) {
  __shared__ int s_tile;
  __shared__ unsigned long long s_wave[kScanThreads / kWave];
  __shared__ unsigned long long s_prefix;
  if (threadIdx.x == 0) s_tile = atomicAdd(tile_counter, 1);
  __syncthreads();
  const int tile = s_tile;
  const int lane = threadIdx.x & (kWave - 1);
  const int wave = threadIdx.x / kWave;
  const long long base = static_cast<long long>(tile) * kScanTile +
                         static_cast<long long>(threadIdx.x) * kScanItems;
  unsigned long long ex[kScanItems];
  unsigned long long sum = 0;
#pragma unroll
  for (int t = 0; t < kScanItems; ++t) {
    const long long idx = base + t;
    const int c = idx < ncells ? count[idx] : 0;
    ex[t] = sum;
    sum += pack_cell(c);
  }
  unsigned long long incl = sum;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const unsigned long long up = __shfl_up(incl, d, kWave);
    if (lane >= d) incl += up;
  }
  if (lane == kWave - 1) s_wave[wave] = incl;
  __syncthreads();
  unsigned long long wave_ex = 0, tile_total = 0;
#pragma unroll
  for (int w = 0; w < kScanThreads / kWave; ++w) {
    if (w < wave) wave_ex += s_wave[w];
    tile_total += s_wave[w];
  }
  if (wave == 0) {
    unsigned long long prefix = 0;
    if (tile == 0) {
      if (lane == 0) {
        __hip_atomic_store(&status[0], kFlagPre | tile_total,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    } else {
      if (lane == 0) {
        __hip_atomic_store(&status[tile], kFlagAgg | tile_total,
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      long long top = tile - 1;
      for (;;) {
        const long long pj = top - lane;
        unsigned long long st = kFlagPre;
        if (pj >= 0) {
          st = __hip_atomic_load(&status[pj], __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_AGENT);
        }
        const unsigned long long flag = st >> 62;
        const unsigned long long pmask = __ballot(flag == 2);
        const unsigned long long zmask = __ballot(flag == 0);
        const int first_p = pmask ? __ffsll(static_cast<long long>(pmask)) - 1
                                  : kWave;
        const unsigned long long need =
            first_p >= kWave - 1 ? ~0ull : ((2ull << first_p) - 1);
        if (zmask & need) continue;  // a predecessor has not published yet
        const unsigned long long v = lane <= first_p ? (st & kValMask) : 0ull;
        prefix += wave_sum_u64(v);
        if (first_p < kWave) break;
        top -= kWave;
      }
      if (lane == 0) {
        __hip_atomic_store(&status[tile], kFlagPre | (prefix + tile_total),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
    }
    if (lane == 0) s_prefix = prefix;
  }
  __syncthreads();
  const unsigned long long thr = s_prefix + wave_ex + incl - sum;
#pragma unroll
  for (int t = 0; t < kScanItems; ++t) {
    const long long idx = base + t;
    if (idx <= ncells) {
      const unsigned long long e = thr + ex[t];
      off[idx] = static_cast<int>(e & 0x7fffffffull);
      item_off[idx] = static_cast<int>((e >> 31) & 0x7fffffffull);
    }
  }
}

// ------------------------------------------------------------- scatter

__global__ __launch_bounds__(256) void k_emit_items(
    const int* __restrict__ count, const int* __restrict__ item_off,
    long long ncells, int* __restrict__ items) {
  const long long g = blockIdx.x * static_cast<long long>(blockDim.x) +
                      threadIdx.x;
  if (g >= ncells) return;
  const int c = count[g];
  if (c == 0) return;
  const int nch = (c + 63) >> 6;
  const int b = item_off[g];
  for (int q = 0; q < nch; ++q) items[b + q] = static_cast<int>(g);
}

__global__ __launch_bounds__(256) void k_scatter(
    const int2* __restrict__ cellslot, const long long* __restrict__ fpt,
    int F, long long ntot, const int* __restrict__ off,
    int* __restrict__ tmp_idx) {
  const long long p = blockIdx.x * static_cast<long long>(blockDim.x) +
                      threadIdx.x;
  if (p >= ntot) return;
  const int2 cs = cellslot[p];
  if (cs.x < 0) return;
  const int f = find_frame(fpt, F, p);
  tmp_idx[off[cs.x] + cs.y] = static_cast<int>(p - fpt[f]);
}

// One wave per (cell, 64-occupant chunk). Each lane ranks its occupant
// among the whole segment by original index, so segments come out in
// ascending index order: the reference linked list walks a cell in
// decreasing index order (src/knearest.rs:423-428), which the exact
// walk replays by reading a segment backwards.
__global__ __launch_bounds__(256) void k_cell_sort(
    const int* __restrict__ items, const int* __restrict__ item_off,
    long long ncells, const int* __restrict__ off,
    const int* __restrict__ tmp_idx, const FrameDesc* __restrict__ fr,
    const long long* __restrict__ fcb, int F, double* __restrict__ sx,
    double* __restrict__ sy, double* __restrict__ sz, int* __restrict__ sidx,
    int* __restrict__ scell) {
  const int w = __builtin_amdgcn_readfirstlane(
      static_cast<int>(blockIdx.x * (blockDim.x / kWave) +
                       threadIdx.x / kWave));
  const int lane = threadIdx.x & (kWave - 1);
  const int total = item_off[ncells];
  if (w >= total) return;
  const int g = __builtin_amdgcn_readfirstlane(items[w]);
  const int q = w - __builtin_amdgcn_readfirstlane(item_off[g]);
  const int o0 = __builtin_amdgcn_readfirstlane(off[g]);
  const int o1 = __builtin_amdgcn_readfirstlane(off[g + 1]);
  const int a = o0 + q * kWave + lane;
  const bool valid = a < o1;
  const int mine = valid ? tmp_idx[a] : INT_MAX;
  int rank = 0;
  for (int b = o0; b < o1; ++b) rank += tmp_idx[b] < mine ? 1 : 0;
  if (!valid) return;
  const int f = find_frame(fcb, F, g);
  const FrameDesc& d = fr[f];
  const CellGeom gm = geom_of(d);
  const double* r = d.xyz + 3 * static_cast<size_t>(mine);
  double s0, s1, s2, fx, fy, fz;
  fold_point(gm, r[0], r[1], r[2], s0, s1, s2, fx, fy, fz);
  const int dst = o0 + rank;
  sx[dst] = fx;
  sy[dst] = fy;
  sz[dst] = fz;
  sidx[dst] = mine;
  scell[dst] = g;
}

// ----------------------------------------------------------- fast walk

// Canonical order on (d2 bits, index). d2 >= 0 or +inf for finite input,
// so the IEEE bit pattern orders like the value.
__device__ __forceinline__ bool key_gt(long long da, int ja, long long db,
                                       int jb) {
  return da > db || (da == db && ja > jb);
}

// Register heap of KC slots kept in DEScending canonical order. Slots
// [0, k) start as (+inf, kEmptyJ); slots [k, KC) hold kUnusedBits and
// never move. Slot 0 is the current k-th best, so the filter threshold
// sits at a fixed register.
template <int KC>
__device__ __forceinline__ void heap_insert(long long (&D)[KC], int (&J)[KC],
                                            long long xd, int xj,
                                            double& min_rej) {
  if (!key_gt(D[0], J[0], xd, xj)) {
    min_rej = fmin(min_rej, __longlong_as_double(xd));
    return;
  }
  if (J[0] != kEmptyJ) min_rej = fmin(min_rej, __longlong_as_double(D[0]));
  bool gt[KC];
#pragma unroll
  for (int t = 0; t < KC; ++t) gt[t] = key_gt(D[t], J[t], xd, xj);
#pragma unroll
  for (int t = 0; t < KC - 1; ++t) {
    const long long nd = gt[t + 1] ? D[t + 1] : (gt[t] ? xd : D[t]);
    const int nj = gt[t + 1] ? J[t + 1] : (gt[t] ? xj : J[t]);
    D[t] = nd;
    J[t] = nj;
  }
  D[KC - 1] = gt[KC - 1] ? xd : D[KC - 1];
  J[KC - 1] = gt[KC - 1] ? xj : J[KC - 1];
}

__device__ __forceinline__ void defer_source(int s, int* defer_list,
                                             int* defer_count) {
  const int pos = atomicAdd(defer_count, 1);
  defer_list[pos] = s;
}

// One lane per active source in cell order (neighbouring lanes share
// home or adjacent cells, so their candidate loads hit the same lines).
// The result equals the reference's whenever no candidate outside the
// final k ties the k-th distance: KHeap::push keeps, per index, the min
// over visited images and the k smallest of those (src/knearest.rs:106-129),
// which is order independent except for that tie. Ties are detected
// exactly (min over rejected and evicted d2 equals the k-th d2) and
// deferred, as are sources still unfinished at r_safe.
template <int KC>
__global__ __launch_bounds__(256) void k_walk_fast(
    const FrameDesc* __restrict__ fr, const long long* __restrict__ fcb,
    int F, const int* __restrict__ fflags, const double* __restrict__ sx,
    const double* __restrict__ sy, const double* __restrict__ sz,
    const int* __restrict__ sidx, const int* __restrict__ scell,
    const int* __restrict__ off, long long ncells, int* __restrict__ defer_list,
    int* __restrict__ defer_count) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int nact = off[ncells];
  if (s >= nact) return;
  const int g = scell[s];
  const int f = find_frame(fcb, F, g);
  const FrameDesc& d = fr[f];
  if (d.kc != KC) return;
  if ((fflags[f] & kFrameNonFinite) != 0 || d.r_safe < 1) {
    defer_source(s, defer_list, defer_count);
    return;
  }
  const int k = d.k;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const long long cb = d.cell_base;
  const int c = static_cast<int>(g - cb);
  const int hx = c % nx;
  const int hy = (c / nx) % ny;
  const int hz = c / (nx * ny);
  const double px = sx[s], py = sy[s], pz = sz[s];
  const int me = sidx[s];
  const int ortho = d.ortho;
  const double w0 = d.widths[0], w1 = d.widths[1], w2 = d.widths[2];
  double hm[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) hm[t] = d.h[t];

  long long D[KC];
  int J[KC];
#pragma unroll
  for (int t = 0; t < KC; ++t) {
    D[t] = t < k ? kInfBits : kUnusedBits;
    J[t] = kEmptyJ;
  }
  double min_rej = __longlong_as_double(kInfBits);
  int done_reach = 0;
  for (int r = 1; r <= d.r_safe; ++r) {
    for (int dx = -r; dx <= r; ++dx) {
      const int jx = hx + dx;
      const int cx = jx < 0 ? jx + nx : (jx >= nx ? jx - nx : jx);
      const int na = jx < 0 ? -1 : (jx >= nx ? 1 : 0);
      for (int dy = -r; dy <= r; ++dy) {
        const int jy = hy + dy;
        const int cy = jy < 0 ? jy + ny : (jy >= ny ? jy - ny : jy);
        const int nb = jy < 0 ? -1 : (jy >= ny ? 1 : 0);
        const bool face = r == 1 || dx == -r || dx == r || dy == -r || dy == r;
        const int step = face ? 1 : 2 * r;
        for (int dz = -r; dz <= r; dz += step) {
          const int jz = hz + dz;
          const int cz = jz < 0 ? jz + nz : (jz >= nz ? jz - nz : jz);
          const int nc = jz < 0 ? -1 : (jz >= nz ? 1 : 0);
          double shx, shy, shz;
          if (ortho) {
            shx = static_cast<double>(na) * w0;
            shy = static_cast<double>(nb) * w1;
            shz = static_cast<double>(nc) * w2;
          } else {
            const double fa = na, fb = nb, fc = nc;
            shx = fa * hm[0] + fb * hm[3] + fc * hm[6];
            shy = fa * hm[1] + fb * hm[4] + fc * hm[7];
            shz = fa * hm[2] + fb * hm[5] + fc * hm[8];
          }
          const long long gq = cb + (static_cast<long long>(cz) * ny + cy) * nx + cx;
          const int a0 = off[gq];
          const int a1 = off[gq + 1];
          for (int t = a0; t < a1; ++t) {
            const int j = sidx[t];
            const double d2 = dist2_shifted(px, py, pz, sx[t], sy[t], sz[t],
                                            shx, shy, shz);
            if (j != me && d2 <= __longlong_as_double(D[0])) {
              heap_insert<KC>(D, J, __double_as_longlong(d2), j, min_rej);
            }
          }
        }
      }
    }
    // stop test, src/knearest.rs:466-471
    if (J[0] != kEmptyJ) {
      const double bound = static_cast<double>(r) * d.cell_min;
      if (__longlong_as_double(D[0]) <= bound * bound) {
        done_reach = r;
        break;
      }
    }
  }
  if (done_reach == 0 || min_rej == __longlong_as_double(D[0])) {
    defer_source(s, defer_list, defer_count);
    return;
  }
  const size_t row = static_cast<size_t>(me) * k;
#pragma unroll
  for (int u = 0; u < KC; ++u) {
    if (u < k) {
      const int t = k - 1 - u;
      const bool real = J[u] != kEmptyJ;
      d.out_nn[row + t] = real ? J[u] : -1;
      if (d.out_d2) {
        d.out_d2[row + t] = real ? __longlong_as_double(D[u]) : nan_value();
      }
    }
  }
  if (d.out_reach) d.out_reach[me] = done_reach;
}

// Queue every active source of the frames that have no register-heap
// bucket (k above the largest bucket, or the exact-only mode).
__global__ __launch_bounds__(256) void k_route_exact(
    const FrameDesc* __restrict__ fr, const long long* __restrict__ fcb,
    int F, const int* __restrict__ scell, const int* __restrict__ off,
    long long ncells, int* __restrict__ defer_list,
    int* __restrict__ defer_count) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int nact = off[ncells];
  if (s >= nact) return;
  const int f = find_frame(fcb, F, scell[s]);
  if (fr[f].kc == 0) defer_source(s, defer_list, defer_count);
}

// ---------------------------------------------------------- exact walk

// f64::total_cmp key
__device__ __forceinline__ long long total_key(double x) {
  const long long b = __double_as_longlong(x);
  return b ^ static_cast<long long>(static_cast<unsigned long long>(b >> 63) >> 1);
}

// Exact-order replay of the reference walk for deferred sources: the
// lexicographic stencil (src/knearest.rs:435-465), each cell in
// decreasing index order, KHeap::push / worst / finish
// (src/knearest.rs:106-158) with the heap in a strided global scratch.
// One lane per source; a grid-stride loop over the deferred list.
__global__ __launch_bounds__(64) void k_walk_exact(
    const FrameDesc* __restrict__ fr, const long long* __restrict__ fcb,
    int F, const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, const int* __restrict__ sidx,
    const int* __restrict__ scell, const int* __restrict__ off,
    const int* __restrict__ defer_list, const int* __restrict__ defer_count,
    double* __restrict__ hd2, int* __restrict__ hidx) {
  const long long tid = blockIdx.x * static_cast<long long>(blockDim.x) +
                        threadIdx.x;
  const long long nthreads = static_cast<long long>(gridDim.x) * blockDim.x;
  const int nq = *defer_count;
  for (long long qi = tid; qi < nq; qi += nthreads) {
    const int s = defer_list[qi];
    const int g = scell[s];
    const int f = find_frame(fcb, F, g);
    const FrameDesc& d = fr[f];
    const CellGeom gm = geom_of(d);
    const int k = d.k;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const long long cb = d.cell_base;
    const int c = static_cast<int>(g - cb);
    const int hx = c % nx, hy = (c / nx) % ny, hz = c / (nx * ny);
    const double px = sx[s], py = sy[s], pz = sz[s];
    const int me = sidx[s];
    double* hd = hd2 + tid;
    int* hj = hidx + tid;
    const long long st = nthreads;
    int n = 0;
    int worst = 0;  // first argmax slot, valid when n == k
    int final_reach = d.max_reach;
    for (int reach = 1; reach <= d.max_reach; ++reach) {
      for (int dx = -reach; dx <= reach; ++dx) {
        for (int dy = -reach; dy <= reach; ++dy) {
          for (int dz = -reach; dz <= reach; ++dz) {
            const bool shell = reach == 1 || abs(dx) == reach ||
                               abs(dy) == reach || abs(dz) == reach;
            if (!shell && reach > 1) continue;
            const int jx = hx + dx, jy = hy + dy, jz = hz + dz;
            const long long cq = (static_cast<long long>(rem_euclid(jz, nz)) * ny +
                                  rem_euclid(jy, ny)) * nx +
                                 rem_euclid(jx, nx);
            double shx, shy, shz;
            lattice_shift(gm, div_euclid(jx, nx), div_euclid(jy, ny),
                          div_euclid(jz, nz), shx, shy, shz);
            const int a0 = off[cb + cq];
            const int a1 = off[cb + cq + 1];
            for (int t = a1 - 1; t >= a0; --t) {
              const int j = sidx[t];
              if (j == me) continue;
              const double d2 = dist2_shifted(px, py, pz, sx[t], sy[t], sz[t],
                                              shx, shy, shz);
              // KHeap::push. With a full heap, d2 >= the worst slot can
              // change nothing: an index already present holds at most
              // that value, and the replacement needs d2 < worst.
              if (n == k && !(d2 < hd[worst * st])) continue;
              bool found = false;
              for (int u = 0; u < n; ++u) {
                if (hj[u * st] == j) {
                  if (d2 < hd[u * st]) hd[u * st] = d2;
                  found = true;
                  break;
                }
              }
              if (!found) {
                if (n < k) {
                  hd[n * st] = d2;
                  hj[n * st] = j;
                  n += 1;
                } else {
                  hd[worst * st] = d2;
                  hj[worst * st] = j;
                }
              }
              if (n == k) {
                worst = 0;
                for (int u = 1; u < n; ++u)
                  if (hd[u * st] > hd[worst * st]) worst = u;
              }
            }
          }
        }
      }
      if (n >= k) {
        double w = hd[0];
        for (int u = 1; u < n; ++u) {
          const double v = hd[u * st];
          if (v > w) w = v;
        }
        const double bound = static_cast<double>(reach) * d.cell_min;
        if (w <= bound * bound) {
          final_reach = reach;
          break;
        }
      }
    }
    // KHeap::finish: sort by (total_cmp(d2), index)
    for (int a = 1; a < n; ++a) {
      const double dv = hd[a * st];
      const int jv = hj[a * st];
      const long long kv = total_key(dv);
      int p = a;
      while (p > 0) {
        const long long kp = total_key(hd[(p - 1) * st]);
        if (kp > kv || (kp == kv && hj[(p - 1) * st] > jv)) {
          hd[p * st] = hd[(p - 1) * st];
          hj[p * st] = hj[(p - 1) * st];
          --p;
        } else {
          break;
        }
      }
      hd[p * st] = dv;
      hj[p * st] = jv;
    }
    const size_t row = static_cast<size_t>(me) * k;
    for (int t = 0; t < k; ++t) {
      const bool real = t < n;
      d.out_nn[row + t] = real ? hj[t * st] : -1;
      if (d.out_d2) d.out_d2[row + t] = real ? hd[t * st] : nan_value();
    }
    if (d.out_reach) d.out_reach[me] = final_reach;
  }
}

template <typename T>
T* as(void* p) {
  return static_cast<T*>(p);
}

unsigned blocks_for(long long n, int threads) {
  const long long b = (n + threads - 1) / threads;
  return static_cast<unsigned>(b < 1 ? 1 : b);
}

}  // namespace

int fast_bucket(int k) {
  if (k <= 4) return 4;
  if (k <= 8) return 8;
  if (k <= 16) return 16;
  if (k <= 32) return 32;
  return 0;
}

Engine::Engine() {
  check(hipGetDevice(&device_), "hipGetDevice");
}

Engine::~Engine() {
  if (stream_ != nullptr) {
    (void)hipStreamSynchronize(stream_);
  }
  free_all();
  if (staging_ != nullptr) (void)hipHostFree(staging_);
  if (staging_ev_ != nullptr) (void)hipEventDestroy(staging_ev_);
  for (auto& e : ev_) {
    if (e != nullptr) (void)hipEventDestroy(e);
  }
  if (stream_ != nullptr) (void)hipStreamDestroy(stream_);
}

hipStream_t Engine::stream() {
  if (stream_ == nullptr) {
    check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking),
          "hipStreamCreate");
  }
  return stream_;
}

void Engine::sync() {
  if (stream_ != nullptr) check(hipStreamSynchronize(stream_), "stream wait");
}

void Engine::grow(Buf& b, size_t bytes) {
  if (bytes <= b.bytes && b.p != nullptr) return;
  if (b.p != nullptr) {
    // a previous call may still read the old buffer
    check(hipStreamSynchronize(stream()), "stream wait");
    check(hipFree(b.p), "hipFree");
    b.p = nullptr;
    b.bytes = 0;
  }
  const size_t want = std::max<size_t>(bytes, 256);
  check(hipMalloc(&b.p, want), "hipMalloc");
  b.bytes = want;
}

void Engine::free_all() {
  Buf* all[] = {&frames_, &fpt_,     &fcb_,      &count_,  &status_,
                &counters_, &off_,   &item_off_, &items_,  &cellslot_,
                &tmp_idx_, &sx_,     &sy_,       &sz_,     &sidx_,
                &scell_,   &defer_,  &heap_d2_,  &heap_idx_, &flags_};
  for (Buf* b : all) {
    if (b->p != nullptr) (void)hipFree(b->p);
    b->p = nullptr;
    b->bytes = 0;
  }
}

const Timings& Engine::last_timings() {
  if (timings_pending_) {
    check(hipEventSynchronize(ev_[7]), "event wait");
    auto el = [&](int a, int b) {
      float ms = 0;
      check(hipEventElapsedTime(&ms, ev_[a], ev_[b]), "elapsed");
      return ms;
    };
    timings_.fold = el(0, 1);
    timings_.scan = el(1, 2);
    timings_.scatter = el(2, 3);
    timings_.sort = el(3, 4);
    timings_.walk = el(4, 5);
    timings_.exact = el(5, 6);
    timings_.total = el(0, 7);
    int dc = 0;
    check(hipMemcpy(&dc, as<int>(counters_.p) + 1, sizeof(int),
                    hipMemcpyDeviceToHost),
          "DtoH deferred");
    timings_.deferred = dc;
    timings_pending_ = false;
  }
  return timings_;
}

void Engine::run(const std::vector<FrameSpec>& frames, bool wait) {
  const int F = static_cast<int>(frames.size());
  if (F == 0) return;
  hipStream_t st = stream();
  std::vector<FrameDesc> desc(static_cast<size_t>(F));
  std::vector<long long> fpt(static_cast<size_t>(F) + 1);
  std::vector<long long> fcb(static_cast<size_t>(F) + 1);
  long long ntot = 0, ctot = 0;
  int kmax_exact = 1;
  bool any_fast[33] = {};
  for (int f = 0; f < F; ++f) {
    const FrameSpec& s = frames[static_cast<size_t>(f)];
    FrameDesc& d = desc[static_cast<size_t>(f)];
    std::memset(&d, 0, sizeof d);
    d.xyz = s.xyz;
    d.mask = s.mask;
    d.out_nn = s.out_nn;
    d.out_d2 = s.out_d2;
    d.out_reach = s.out_reach;
    std::memcpy(d.h, s.geom.h, sizeof d.h);
    std::memcpy(d.hinv, s.geom.hinv, sizeof d.hinv);
    std::memcpy(d.origin, s.geom.origin, sizeof d.origin);
    std::memcpy(d.widths, s.geom.widths, sizeof d.widths);
    d.ortho = s.geom.ortho;
    d.cell_min = s.plan.cell_min;
    d.nx = s.plan.nx;
    d.ny = s.plan.ny;
    d.nz = s.plan.nz;
    d.max_reach = s.plan.max_reach;
    d.r_safe = safe_reach(s.plan);
    d.n = s.n;
    d.k = s.k;
    d.pt_base = ntot;
    d.cell_base = ctot;
    d.kc = mode_ == WalkMode::kExactOnly ? 0 : fast_bucket(s.k);
    if (d.kc) any_fast[d.kc] = true;
    kmax_exact = std::max(kmax_exact, s.k);
    fpt[static_cast<size_t>(f)] = ntot;
    fcb[static_cast<size_t>(f)] = ctot;
    ntot += s.n;
    ctot += s.plan.ncell;
  }
  fpt[static_cast<size_t>(F)] = ntot;
  fcb[static_cast<size_t>(F)] = ctot;
  if (ntot >= INT_MAX || ctot >= INT_MAX) {
    throw std::runtime_error("batch exceeds 2^31 - 1 points or cells");
  }
  const long long tiles = ctot / kScanTile + 1;

  grow(frames_, sizeof(FrameDesc) * static_cast<size_t>(F));
  grow(fpt_, sizeof(long long) * (static_cast<size_t>(F) + 1));
  grow(fcb_, sizeof(long long) * (static_cast<size_t>(F) + 1));
  grow(count_, sizeof(int) * static_cast<size_t>(ctot));
  grow(status_, sizeof(unsigned long long) * static_cast<size_t>(tiles));
  grow(counters_, sizeof(int) * 4);
  grow(off_, sizeof(int) * (static_cast<size_t>(ctot) + 1));
  grow(item_off_, sizeof(int) * (static_cast<size_t>(ctot) + 1));
  grow(items_, sizeof(int) * (static_cast<size_t>(ntot) + 1));
  grow(cellslot_, sizeof(int2) * static_cast<size_t>(ntot));
  grow(tmp_idx_, sizeof(int) * static_cast<size_t>(ntot));
  grow(sx_, sizeof(double) * static_cast<size_t>(ntot));
  grow(sy_, sizeof(double) * static_cast<size_t>(ntot));
  grow(sz_, sizeof(double) * static_cast<size_t>(ntot));
  grow(sidx_, sizeof(int) * static_cast<size_t>(ntot));
  grow(scell_, sizeof(int) * static_cast<size_t>(ntot));
  grow(defer_, sizeof(int) * static_cast<size_t>(ntot));
  grow(flags_, sizeof(int) * static_cast<size_t>(F));
  // exact-walk scratch: one heap of kmax slots per resident lane
  constexpr long long kExactBlocks = 1024;
  long long exact_threads = kExactBlocks * 64;
  const long long heap_budget = 512ll << 20;
  while (exact_threads > 64 &&
         exact_threads * kmax_exact * 12 > heap_budget) {
    exact_threads /= 2;
  }
  grow(heap_d2_, sizeof(double) * static_cast<size_t>(exact_threads) *
                     static_cast<size_t>(kmax_exact));
  grow(heap_idx_, sizeof(int) * static_cast<size_t>(exact_threads) *
                      static_cast<size_t>(kmax_exact));

  // descriptors go through pinned staging; wait until the previous call's
  // copy has consumed it
  const size_t need = sizeof(FrameDesc) * static_cast<size_t>(F) +
                      2 * sizeof(long long) * (static_cast<size_t>(F) + 1);
  if (staging_ev_ == nullptr) {
    check(hipEventCreateWithFlags(&staging_ev_, hipEventDisableTiming),
          "hipEventCreate");
  }
  if (staging_pending_) {
    check(hipEventSynchronize(staging_ev_), "staging wait");
    staging_pending_ = false;
  }
  if (need > staging_bytes_) {
    if (staging_ != nullptr) check(hipHostFree(staging_), "hipHostFree");
    staging_ = nullptr;
    check(hipHostMalloc(&staging_, need, hipHostMallocDefault),
          "hipHostMalloc");
    staging_bytes_ = need;
  }
  char* sp = static_cast<char*>(staging_);
  const size_t dbytes = sizeof(FrameDesc) * static_cast<size_t>(F);
  const size_t bbytes = sizeof(long long) * (static_cast<size_t>(F) + 1);
  std::memcpy(sp, desc.data(), dbytes);
  std::memcpy(sp + dbytes, fpt.data(), bbytes);
  std::memcpy(sp + dbytes + bbytes, fcb.data(), bbytes);
  check(hipMemcpyAsync(frames_.p, sp, dbytes, hipMemcpyHostToDevice, st),
        "HtoD frames");
  check(hipMemcpyAsync(fpt_.p, sp + dbytes, bbytes, hipMemcpyHostToDevice, st),
        "HtoD fpt");
  check(hipMemcpyAsync(fcb_.p, sp + dbytes + bbytes, bbytes,
                       hipMemcpyHostToDevice, st),
        "HtoD fcb");
  check(hipEventRecord(staging_ev_, st), "event record");
  staging_pending_ = true;

  if (timing_) {
    for (auto& e : ev_) {
      if (e == nullptr) check(hipEventCreate(&e), "hipEventCreate");
    }
    check(hipEventRecord(ev_[0], st), "event record");
  }

  // clears (fill kernels, not memset nodes, so the sequence can be
  // captured into a graph)
  k_zero_u32<<<blocks_for(ctot, 256), 256, 0, st>>>(as<unsigned>(count_.p),
                                                    ctot);
  k_zero_u32<<<blocks_for(tiles * 2, 256), 256, 0, st>>>(
      as<unsigned>(status_.p), tiles * 2);
  k_zero_u32<<<1, 64, 0, st>>>(as<unsigned>(counters_.p), 4);
  k_zero_u32<<<blocks_for(F, 256), 256, 0, st>>>(as<unsigned>(flags_.p), F);

  const FrameDesc* dfr = as<FrameDesc>(frames_.p);
  const long long* dfpt = as<long long>(fpt_.p);
  const long long* dfcb = as<long long>(fcb_.p);
  int* counters = as<int>(counters_.p);

  k_fold_count<<<blocks_for(ntot, 256), 256, 0, st>>>(
      dfr, dfpt, F, ntot, as<int>(count_.p), as<int2>(cellslot_.p),
      as<int>(flags_.p));
  check(hipGetLastError(), "fold_count launch");
  if (timing_) check(hipEventRecord(ev_[1], st), "event record");

  k_scan<<<static_cast<unsigned>(tiles), kScanThreads, 0, st>>>(
      as<int>(count_.p), ctot, as<int>(off_.p), as<int>(item_off_.p),
      as<unsigned long long>(status_.p), counters + 0);
  check(hipGetLastError(), "scan launch");
  if (timing_) check(hipEventRecord(ev_[2], st), "event record");

  k_emit_items<<<blocks_for(ctot, 256), 256, 0, st>>>(
      as<int>(count_.p), as<int>(item_off_.p), ctot, as<int>(items_.p));
  k_scatter<<<blocks_for(ntot, 256), 256, 0, st>>>(
      as<int2>(cellslot_.p), dfpt, F, ntot, as<int>(off_.p),
      as<int>(tmp_idx_.p));
  check(hipGetLastError(), "scatter launch");
  if (timing_) check(hipEventRecord(ev_[3], st), "event record");

  // items <= ntot (one chunk per non-empty cell at most per 64 points,
  // plus one per cell); launch for the bound and exit early
  const long long max_items = std::min<long long>(ntot, ctot + ntot / 64 + 1);
  k_cell_sort<<<blocks_for(max_items * kWave, 256), 256, 0, st>>>(
      as<int>(items_.p), as<int>(item_off_.p), ctot, as<int>(off_.p),
      as<int>(tmp_idx_.p), dfr, dfcb, F, as<double>(sx_.p), as<double>(sy_.p),
      as<double>(sz_.p), as<int>(sidx_.p), as<int>(scell_.p));
  check(hipGetLastError(), "cell_sort launch");
  if (timing_) check(hipEventRecord(ev_[4], st), "event record");

  const unsigned wblocks = blocks_for(ntot, 256);
#define LC_WALK(KC)                                                         \
  if (any_fast[KC]) {                                                       \
    k_walk_fast<KC><<<wblocks, 256, 0, st>>>(                               \
        dfr, dfcb, F, as<int>(flags_.p), as<double>(sx_.p),                 \
        as<double>(sy_.p), as<double>(sz_.p), as<int>(sidx_.p),             \
        as<int>(scell_.p), as<int>(off_.p), ctot, as<int>(defer_.p),        \
        counters + 1);                                                      \
    check(hipGetLastError(), "walk launch");                                \
  }
  LC_WALK(4)
  LC_WALK(8)
  LC_WALK(16)
  LC_WALK(32)
#undef LC_WALK
  // frames outside every register bucket go straight to the exact walk
  bool any_exact_frame = false;
  for (int f = 0; f < F; ++f) {
    any_exact_frame |= desc[static_cast<size_t>(f)].kc == 0;
  }
  if (any_exact_frame) {
    k_route_exact<<<wblocks, 256, 0, st>>>(dfr, dfcb, F, as<int>(scell_.p),
                                           as<int>(off_.p), ctot,
                                           as<int>(defer_.p), counters + 1);
    check(hipGetLastError(), "route launch");
  }
  if (timing_) check(hipEventRecord(ev_[5], st), "event record");
  k_walk_exact<<<static_cast<unsigned>(exact_threads / 64), 64, 0, st>>>(
      dfr, dfcb, F, as<double>(sx_.p), as<double>(sy_.p), as<double>(sz_.p),
      as<int>(sidx_.p), as<int>(scell_.p), as<int>(off_.p), as<int>(defer_.p),
      counters + 1, as<double>(heap_d2_.p), as<int>(heap_idx_.p));
  check(hipGetLastError(), "exact walk launch");
  if (timing_) {
    check(hipEventRecord(ev_[6], st), "event record");
    check(hipEventRecord(ev_[7], st), "event record");
    timings_pending_ = true;
  }
  if (wait) sync();
}

}  // namespace lcamd
