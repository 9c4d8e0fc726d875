// Linked-cell k-nearest pipeline on gfx950 (MI355X).
//
// Preparation (all single-pass streaming kernels):
//   fold_count  fold every point into its primary cell, bin it, count
//               occupants with one atomic per point, or per run of lanes in
//               one cell (the slot is the point's arrival rank in the
//               cell); the 9-bit Hilbert index of the point's 8x8x8
//               sub-cell rides along
//   scan x2     single-pass decoupled look-back exclusive scans over cell
//               counts (DPP row shifts and row broadcasts within the wave);
//               one packed 64-bit status word per 2048-cell tile carries the
//               occupant prefix and a second prefix (16-point tiles, then
//               64-source chunks)
//   scatter     arrival-order scatter of (key, id) into cell segments
//   cell_sort   one wave per 64-member chunk: rank by (sub-cell key, id) for
//               the walk layout and by id for the exact-order replay,
//               re-fold, write 32-byte candidate records (x, y, z, id)
//   boxes       exact bounding boxes of every 16-point tile (and of its two
//               8-point halves) and every cell
// Walks:
//   walk_rows   the default: one 16-lane row per 16-source tile, four rows
//               per wave; candidate tiles listed per shell with exact box
//               tests, drained nearest first, candidates broadcast within
//               the row through DPP, accepted ones buffered in LDS and
//               merged into per-lane register heaps
//   walk_dense  one wave per 64 sources of one dense home cell; candidates
//               arrive through scalar loads (uniform across the wave), tiles
//               whose exact lower-bound distance exceeds every lane's k-th
//               best are skipped, accepted candidates are buffered in LDS and
//               merged into per-lane register heaps in batches
//   walk_sparse one lane per source of a sparse home cell, same heap, the
//               same tile bounds per lane
//   walk_exact  exact-order replay of the reference walk and KHeap for the
//               sources the fast walks defer (distance ties at or inside the
//               k-th, stencil wrap, k above the register buckets, non-finite
//               input)
//
// Reference: src/knearest.rs:351-486 (the CPU walk, the parity target) and
// src/gpu/knearest.cpp:192-441 (the CUDA kernels this file replaces).
#include "lc_engine.hpp"
#include "lc_sched.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <mutex>
#include <type_traits>

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
// The walks' lambdas capture their state (heaps, thresholds, buffers) by
// reference; one left out of line would put that state in scratch memory.
#define LC_INLINE __attribute__((always_inline))
constexpr long long kInfBits = 0x7ff0000000000000LL;
constexpr long long kUnusedBits = static_cast<long long>(0xfff0000000000000ULL);  // -inf
constexpr unsigned long long kNanBits = 0x7ff8000000000000ull;
constexpr int kTile = 16;      // points per bounding-box tile
constexpr int kTileShift = 4;
constexpr int kChunkShift = 6;  // sources per dense work item (one wave)
constexpr int kPad = 16;        // slack elements after SoA arrays (group reads)
constexpr int kNeverDense = 1 << 30;  // dense_min that keeps every cell sparse
constexpr int kRowWalk = -1;          // split value selecting the row walk
constexpr int kRowWalkPersist = -2;   // the row walk as one persistent launch per heap bucket
constexpr int kExactWaveMaxK = 2 * kWave;  // wave-per-source exact walk: k up to two slots per lane

// Cell-ordered candidate record: folded coordinates and the point's index
// within its frame (64-bit, so a record is 32 bytes: one s_load_dwordx8
// when the address is wave-uniform, two dwordx4 loads otherwise).
struct __align__(32) Cand {
  double x, y, z;
  long long id;
};

// Exact bounding box of a 16-point tile or of a cell, padded to 64 bytes.
struct __align__(64) Box {
  double lx, hx, ly, hy, lz, hz, pad0, pad1;
};

// r = 1 stencil in near-first order: home, 6 faces, 12 edges, 8 corners.
// The canonical heap makes the result independent of the visit order
// inside a shell; the order only makes the k-th distance tight early.
constexpr signed char kNear27[27][3] = {
    {0, 0, 0},   {-1, 0, 0},  {1, 0, 0},   {0, -1, 0},  {0, 1, 0},
    {0, 0, -1},  {0, 0, 1},   {-1, -1, 0}, {-1, 1, 0},  {1, -1, 0},
    {1, 1, 0},   {-1, 0, -1}, {-1, 0, 1},  {1, 0, -1},  {1, 0, 1},
    {0, -1, -1}, {0, -1, 1},  {0, 1, -1},  {0, 1, 1},   {-1, -1, -1},
    {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1},  {1, -1, -1}, {1, -1, 1},
    {1, 1, -1},  {1, 1, 1}};

// The table packed into three 64-bit words, 10 entries of 6 bits each
// ((dx+1) | (dy+1) << 2 | (dz+1) << 4), so a lane reads its entry with
// shifts instead of a memory load.
constexpr unsigned long long near27_word(int w) {
  unsigned long long v = 0;
  for (int j = 0; j < 10 && 10 * w + j < 27; ++j) {
    const int i = 10 * w + j;
    const unsigned long long e =
        static_cast<unsigned long long>((kNear27[i][0] + 1) | ((kNear27[i][1] + 1) << 2) |
                                        ((kNear27[i][2] + 1) << 4));
    v |= e << (6 * j);
  }
  return v;
}
__device__ __forceinline__ void near27(int i, int& dx, int& dy, int& dz) {
  constexpr unsigned long long w0 = near27_word(0), w1 = near27_word(1),
                               w2 = near27_word(2);
  const unsigned long long w = i < 10 ? w0 : (i < 20 ? w1 : w2);
  const int j = i < 10 ? i : (i < 20 ? i - 10 : i - 20);
  const int e = static_cast<int>((w >> (6 * j)) & 63ull);
  dx = (e & 3) - 1;
  dy = ((e >> 2) & 3) - 1;
  dz = ((e >> 4) & 3) - 1;
}

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

__device__ __forceinline__ double as_f64(long long b) {
  return __longlong_as_double(b);
}

__device__ __forceinline__ int uni(int v) {
  return __builtin_amdgcn_readfirstlane(v);
}

// Sub-cell position of a point inside its cell, LC_SUB_BITS bits per axis
// (2^LC_SUB_BITS sub-cells per axis), and its key: the Hilbert index below
// by default, or the Morton key interleaving the bits (x lowest, then y,
// then z, per bit) with LC_HILBERT=0. The key orders a cell's members for
// the walk layout, so 16 consecutive members (a tile) are spatially compact
// even in cells of several hundred points.
#ifndef LC_SUB_BITS
#define LC_SUB_BITS 3
#endif
__device__ __forceinline__ int sub_bits(double s, int n, int bin) {
  constexpr int kSub = 1 << LC_SUB_BITS;
  const double u = s * static_cast<double>(n) - static_cast<double>(bin);
  int v = static_cast<int>(u * static_cast<double>(kSub));
  v = v < 0 ? 0 : (v > kSub - 1 ? kSub - 1 : v);
  return v;
}

__device__ __forceinline__ int morton_key(int x, int y, int z) {
  int key = 0;
#pragma unroll
  for (int b = 0; b < LC_SUB_BITS; ++b) {
    key |= ((x >> b) & 1) << (3 * b);
    key |= ((y >> b) & 1) << (3 * b + 1);
    key |= ((z >> b) & 1) << (3 * b + 2);
  }
  return key;
}

// Hilbert index of a sub-cell (J. Skilling, "Programming the Hilbert
// curve", AIP Conf. Proc. 707, 2004: axes to transposed index, then the
// bits of the three transposed words interleaved from the top level down).
// Consecutive indices are face-adjacent sub-cells, so a run of 16 members
// stays more compact than along the Morton order's jumps.
__device__ __forceinline__ int hilbert_key(int x, int y, int z) {
  int X[3] = {x, y, z};
#pragma unroll
  for (int q = 1 << (LC_SUB_BITS - 1); q > 1; q >>= 1) {
    const int p = q - 1;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (X[i] & q) {
        X[0] ^= p;
      } else {
        const int t = (X[0] ^ X[i]) & p;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  int t = 0;
#pragma unroll
  for (int q = 1 << (LC_SUB_BITS - 1); q > 1; q >>= 1) {
    if (X[2] & q) t ^= q - 1;
  }
  int key = 0;
#pragma unroll
  for (int b = LC_SUB_BITS - 1; b >= 0; --b) {
#pragma unroll
    for (int i = 0; i < 3; ++i) key = (key << 1) | (((X[i] ^ t) >> b) & 1);
  }
  return key;
}

#ifndef LC_HILBERT
#define LC_HILBERT 1
#endif
__device__ __forceinline__ int subcell_key(int x, int y, int z) {
  return LC_HILBERT ? hilbert_key(x, y, z) : morton_key(x, y, z);
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
    int4* __restrict__ cellslot, int* __restrict__ fflags) {
  const long long p = blockIdx.x * static_cast<long long>(blockDim.x) +
                      threadIdx.x;
  if (p >= ntot) return;
  const int f = find_frame(fpt, F, p);
  const FrameDesc& d = fr[f];
  const int i = static_cast<int>(p - d.pt_base);
  if (d.skip != 0 || (d.mask != nullptr && d.mask[i] == 0)) {
    cellslot[p] = make_int4(-1, 0, 0, 0);
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
  const int key = subcell_key(sub_bits(s0, d.nx, bx), sub_bits(s1, d.ny, by),
                             sub_bits(s2, d.nz, bz));
  const long long gcell =
      d.cell_base + (static_cast<long long>(bz) * d.ny + by) * d.nx + bx;
  // Points given in spatial order (lattices, molecules) put runs of
  // consecutive lanes in one cell: when every lane is active and the runs
  // average two lanes or more, each run takes its slots with one atomic from
  // its first lane. In random order each lane takes its own.
  const int lane = threadIdx.x & (kWave - 1);
  int slot = 0;
  bool runs = false;
  if (__ballot(true) == ~0ull) {
    // the previous lane's cell (DPP wave_shr:1; lane 0 reads -1)
    const int gc = static_cast<int>(gcell);
    const int gprev = __builtin_amdgcn_update_dpp(-1, gc, 0x138, 0xf, 0xf, false);
    const bool head = gprev != gc;
    const unsigned long long heads = __ballot(head);
    runs = __popcll(heads) <= kWave / 2;
    if (runs) {
      const unsigned long long upto = heads & ((2ull << lane) - 1ull);
      const int start = 63 - __clzll(upto);
      const unsigned long long above = heads & ~((2ull << lane) - 1ull);
      const int end = above != 0 ? __ffsll(static_cast<long long>(above)) - 1 : kWave;
      int base = 0;
      if (head) base = atomicAdd(&count[gcell], end - lane);
      slot = __shfl(base, start, kWave) + (lane - start);
    }
  }
  if (!runs) slot = atomicAdd(&count[gcell], 1);
  cellslot[p] = make_int4(static_cast<int>(gcell), slot, key, 0);
}

// Device-side plan of a device-resident cell: Cell::from_vectors and the
// mesh plan (src/cell.rs:97-123, src/knearest.rs:377-397) with the host
// plan's arithmetic (one thread; IEEE division and square root), then the
// descriptors of every frame (the role of the reference's invert_cell,
// src/gpu/knearest.cpp:74-190, without its read-back of the plan). status:
// 0 planned, 1 bad cell, 2 too many cells (the reference's limits, or more
// than cap_cells cells); a failed frame is marked skip over a 1x1x1 mesh.
__global__ __launch_bounds__(64) void k_plan_dcell(
    const double* __restrict__ cell, int cell_n, double hint, FrameDesc tmpl, int F,
    long long n, long long k, long long cap_cells, FrameDesc* __restrict__ out,
    long long* __restrict__ fpt, long long* __restrict__ fcb, int* __restrict__ status) {
  __shared__ CellGeom sg;
  __shared__ MeshPlan sp;
  __shared__ int sst;
  if (threadIdx.x == 0) {
    double a[3], b[3], c[3], o[3] = {0.0, 0.0, 0.0};
    if (cell_n == 3) {
      a[0] = cell[0], a[1] = 0.0, a[2] = 0.0;
      b[0] = 0.0, b[1] = cell[1], b[2] = 0.0;
      c[0] = 0.0, c[1] = 0.0, c[2] = cell[2];
    } else {
      for (int t = 0; t < 3; ++t) {
        a[t] = cell[t];
        b[t] = cell[3 + t];
        c[t] = cell[6 + t];
        if (cell_n == 12) o[t] = cell[9 + t];
      }
    }
    int st = cell_from_vectors(a, b, c, o, sg) == kOk ? 0 : 1;
    if (st == 0 && plan_mesh(sg, hint, sp) != kOk) st = 2;
    if (st == 0 && sp.ncell > cap_cells) st = 2;
    sst = st;
    *status = st;
  }
  __syncthreads();
  const bool ok = sst == 0;
  const long long nc = ok ? sp.ncell : 1;
  for (int f = threadIdx.x; f <= F; f += blockDim.x) {
    fpt[f] = f * n;
    fcb[f] = f * nc;
    if (f == F) continue;
    FrameDesc d = tmpl;
    d.xyz = tmpl.xyz + f * n * 3;
    d.out_nn = tmpl.out_nn + f * n * k;
    d.out_d2 = tmpl.out_d2 != nullptr ? tmpl.out_d2 + f * n * k : nullptr;
    d.out_reach = tmpl.out_reach != nullptr ? tmpl.out_reach + f * n : nullptr;
    for (int t = 0; t < 9; ++t) {
      d.h[t] = ok ? sg.h[t] : (t % 4 == 0 ? 1.0 : 0.0);
      d.hinv[t] = ok ? sg.hinv[t] : (t % 4 == 0 ? 1.0 : 0.0);
    }
    for (int t = 0; t < 3; ++t) {
      d.origin[t] = ok ? sg.origin[t] : 0.0;
      d.widths[t] = ok ? sg.widths[t] : 1.0;
    }
    d.ortho = ok ? sg.ortho : 1;
    d.cell_min = ok ? sp.cell_min : 1.0;
    d.nx = ok ? sp.nx : 1;
    d.ny = ok ? sp.ny : 1;
    d.nz = ok ? sp.nz : 1;
    d.max_reach = ok ? sp.max_reach : 1;
    d.r_safe = ok ? safe_reach(sp) : 0;
    d.pt_base = f * n;
    d.cell_base = f * nc;
    d.skip = ok ? 0 : 1;
    out[f] = d;
  }
}

// ---------------------------------------------------------------- scan

__device__ __forceinline__ unsigned long long pack_cell(int cnt, int sh) {
  const unsigned long long second =
      static_cast<unsigned long long>((cnt + (1 << sh) - 1) >> sh);
  return static_cast<unsigned long long>(cnt) | (second << 31);
}

__device__ __forceinline__ unsigned long long wave_sum_u64(
    unsigned long long v) {
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, kWave);
  return v;
}

// v from the lane DPP control CTRL names (0 where it names none, or outside
// the rows ROWS selects), both 32-bit halves moved by DPP
template <int CTRL, int ROWS>
__device__ __forceinline__ unsigned long long dpp_u64(unsigned long long v) {
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(v), CTRL, ROWS, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(v >> 32), CTRL, ROWS, 0xf, false);
  return (static_cast<unsigned long long>(static_cast<unsigned>(hi)) << 32) |
         static_cast<unsigned long long>(static_cast<unsigned>(lo));
}

// Inclusive scan over the 64 lanes of a wave through DPP: row_shr by 1, 2,
// 4 and 8 scans each 16-lane row (lanes shifted in from outside the row
// read 0), row_bcast:15 adds row 0's total to row 1 and row 2's to row 3,
// row_bcast:31 adds the total of rows 0-1 to rows 2 and 3. Every lane must
// be active.
__device__ __forceinline__ unsigned long long wave_incl_scan_u64(
    unsigned long long v) {
  v += dpp_u64<0x111, 0xf>(v);
  v += dpp_u64<0x112, 0xf>(v);
  v += dpp_u64<0x114, 0xf>(v);
  v += dpp_u64<0x118, 0xf>(v);
  v += dpp_u64<0x142, 0xa>(v);
  v += dpp_u64<0x143, 0xc>(v);
  return v;
}

// Single-pass exclusive scan with decoupled look-back. Tiles take ids
// from an atomic counter, so a tile only ever waits on tiles that are
// already resident. Each tile publishes its aggregate at once and its
// inclusive prefix after the look-back; the whole first wave inspects 64
// predecessors per step with one ballot. The scanned value packs the
// occupant count (bits 0-30) with ceil(count / 2^sh) (bits 31-61).
__global__ __launch_bounds__(kScanThreads) void k_scan(
    const int* __restrict__ count, long long ncells, int sh,
    int* __restrict__ off, int* __restrict__ second_off,
    unsigned long long* status, int* tile_counter) {
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
    sum += pack_cell(c, sh);
  }
  const unsigned long long incl = wave_incl_scan_u64(sum);
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
      if (off) off[idx] = static_cast<int>(e & 0x7fffffffull);
      second_off[idx] = static_cast<int>((e >> 31) & 0x7fffffffull);
    }
  }
}

// ------------------------------------------------------------- scatter

__global__ __launch_bounds__(256) void k_emit_items(
    const int* __restrict__ count, const int* __restrict__ item_off,
    const int* __restrict__ tile_off, long long ncells,
    int* __restrict__ items, int* __restrict__ tile_cell) {
  const long long g = blockIdx.x * static_cast<long long>(blockDim.x) +
                      threadIdx.x;
  if (g >= ncells) return;
  const int c = count[g];
  if (c == 0) return;
  const int nch = (c + (1 << kChunkShift) - 1) >> kChunkShift;
  const int b = item_off[g];
  for (int q = 0; q < nch; ++q) items[b + q] = static_cast<int>(g);
  const int nt = (c + kTile - 1) >> kTileShift;
  const int tb = tile_off[g];
  for (int q = 0; q < nt; ++q) tile_cell[tb + q] = static_cast<int>(g);
}

__global__ __launch_bounds__(256) void k_scatter(
    const int4* __restrict__ cellslot, const long long* __restrict__ fpt,
    int F, long long ntot, const int* __restrict__ off,
    long long* __restrict__ tmp_key) {
  const long long p = blockIdx.x * static_cast<long long>(blockDim.x) +
                      threadIdx.x;
  if (p >= ntot) return;
  const int4 cs = cellslot[p];
  if (cs.x < 0) return;
  const int f = find_frame(fpt, F, p);
  const long long i = p - fpt[f];
  tmp_key[off[cs.x] + cs.y] = (static_cast<long long>(cs.z) << 32) | i;
}

#ifndef LC_SORT_BATCH
#define LC_SORT_BATCH 32
#endif

// One wave per (cell, 64-member chunk). Each lane ranks its member within
// the segment twice: by (sub-cell key, id) for the walk layout, whose
// consecutive members are spatially close, and by id alone for byidx[],
// through which the exact-order walk visits a cell in decreasing index
// order as the reference linked list does (src/knearest.rs:423-428).
__global__ __launch_bounds__(256) void k_cell_sort(
    const int* __restrict__ items, const int* __restrict__ item_off,
    long long ncells, const int* __restrict__ off,
    const long long* __restrict__ tmp_key, const FrameDesc* __restrict__ fr,
    const long long* __restrict__ fcb, int F, Cand* __restrict__ cand,
    int* __restrict__ scell, int* __restrict__ byidx) {
  const int w = uni(static_cast<int>(blockIdx.x * (blockDim.x / kWave) +
                                     threadIdx.x / kWave));
  const int lane = threadIdx.x & (kWave - 1);
  const int total = item_off[ncells];
  if (w >= total) return;
  const int g = uni(items[w]);
  const int q = w - uni(item_off[g]);
  const int o0 = uni(off[g]);
  const int o1 = uni(off[g + 1]);
  const int a = o0 + q * kWave + lane;
  const bool valid = a < o1;
  const long long mine = valid ? tmp_key[a] : LLONG_MAX;
  const int mid = static_cast<int>(mine & 0xffffffffll);
  int rk = 0, ri = 0;
  // the segment's keys LC_SORT_BATCH at a time (wave-uniform addresses:
  // wide scalar loads, one wait per batch), then eight at a time, then the
  // remainder
  int b = o0;
#pragma unroll 1
  for (; b + LC_SORT_BATCH <= o1; b += LC_SORT_BATCH) {
    long long v[LC_SORT_BATCH];
#pragma unroll
    for (int u = 0; u < LC_SORT_BATCH; ++u) v[u] = tmp_key[b + u];
#pragma unroll
    for (int u = 0; u < LC_SORT_BATCH; ++u) {
      rk += v[u] < mine ? 1 : 0;
      ri += static_cast<int>(v[u] & 0xffffffffll) < mid ? 1 : 0;
    }
  }
#pragma unroll 1
  for (; b + 8 <= o1; b += 8) {
    long long v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = tmp_key[b + u];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      rk += v[u] < mine ? 1 : 0;
      ri += static_cast<int>(v[u] & 0xffffffffll) < mid ? 1 : 0;
    }
  }
  for (; b < o1; ++b) {
    const long long v = tmp_key[b];
    rk += v < mine ? 1 : 0;
    ri += static_cast<int>(v & 0xffffffffll) < mid ? 1 : 0;
  }
  if (!valid) return;
  const int f = find_frame(fcb, F, g);
  const FrameDesc& d = fr[f];
  const CellGeom gm = geom_of(d);
  const double* r = d.xyz + 3 * static_cast<size_t>(mid);
  double s0, s1, s2, fx, fy, fz;
  fold_point(gm, r[0], r[1], r[2], s0, s1, s2, fx, fy, fz);
  const int dst = o0 + rk;
  Cand cr;
  cr.x = fx;
  cr.y = fy;
  cr.z = fz;
  cr.id = mid;
  cand[dst] = cr;
  scell[dst] = g;
  byidx[o0 + ri] = dst;
}

// Lane permutation within a 16-lane row through DPP on the two 32-bit
// halves (64-bit DPP on gfx950 takes only the row broadcast).
template <int CTRL>
__device__ __forceinline__ double row_dpp(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, 0xf, 0xf, true);
  return __builtin_bit_cast(double, (static_cast<long long>(hi) << 32) |
                                        static_cast<long long>(static_cast<unsigned>(lo)));
}
// Maximum / minimum over the 16 lanes of a row: quad_perm [1,0,3,2] and
// [2,3,0,1] reduce each quad, row_half_mirror joins the quads of each half
// row, row_mirror joins the halves (max and min ignore the order).
__device__ __forceinline__ double row_max16(double v) {
  v = fmax(v, row_dpp<0xB1>(v));
  v = fmax(v, row_dpp<0x4E>(v));
  v = fmax(v, row_dpp<0x141>(v));
  return fmax(v, row_dpp<0x140>(v));
}
__device__ __forceinline__ double row_min16(double v) {
  v = fmin(v, row_dpp<0xB1>(v));
  v = fmin(v, row_dpp<0x4E>(v));
  v = fmin(v, row_dpp<0x141>(v));
  return fmin(v, row_dpp<0x140>(v));
}
__device__ __forceinline__ unsigned row_min16_u32(unsigned v) {
  auto d = [](unsigned x, auto ctrl) {
    return static_cast<unsigned>(__builtin_amdgcn_update_dpp(
        0, static_cast<int>(x), decltype(ctrl)::value, 0xf, 0xf, true));
  };
  v = min(v, d(v, std::integral_constant<int, 0xB1>{}));
  v = min(v, d(v, std::integral_constant<int, 0x4E>{}));
  v = min(v, d(v, std::integral_constant<int, 0x141>{}));
  return min(v, d(v, std::integral_constant<int, 0x140>{}));
}

// Exact bounding boxes. One wave per 64-member chunk = four 16-point
// tiles; each 16-lane row reduces min / max with DPP.
__global__ __launch_bounds__(256) void k_tile_box(
    const int* __restrict__ items, const int* __restrict__ item_off,
    long long ncells, const int* __restrict__ off,
    const int* __restrict__ tile_off, const Cand* __restrict__ cand,
    Box* __restrict__ tbox, Box* __restrict__ hbox) {
  const int w = uni(static_cast<int>(blockIdx.x * (blockDim.x / kWave) +
                                     threadIdx.x / kWave));
  const int lane = threadIdx.x & (kWave - 1);
  const int total = item_off[ncells];
  if (w >= total) return;
  const int g = uni(items[w]);
  const int q = w - uni(item_off[g]);
  const int o0 = uni(off[g]);
  const int o1 = uni(off[g + 1]);
  const int t0 = uni(tile_off[g]);
  const int a = o0 + q * kWave + lane;
  const bool valid = a < o1;
  const double inf = as_f64(kInfBits);
  const Cand cr = cand[valid ? a : o0];
  const double x = cr.x, y = cr.y, z = cr.z;
  // each half row (8 lanes) reduces, then row_mirror brings the other
  // half's result: lane 0 of the row holds both halves and the whole tile
  auto half_min = [](double v) {
    v = fmin(v, row_dpp<0xB1>(v));
    v = fmin(v, row_dpp<0x4E>(v));
    return fmin(v, row_dpp<0x141>(v));
  };
  auto half_max = [](double v) {
    v = fmax(v, row_dpp<0xB1>(v));
    v = fmax(v, row_dpp<0x4E>(v));
    return fmax(v, row_dpp<0x141>(v));
  };
  const double alx = half_min(valid ? x : inf), ahx = half_max(valid ? x : -inf);
  const double aly = half_min(valid ? y : inf), ahy = half_max(valid ? y : -inf);
  const double alz = half_min(valid ? z : inf), ahz = half_max(valid ? z : -inf);
  const double blx = row_dpp<0x140>(alx), bhx = row_dpp<0x140>(ahx);
  const double bly = row_dpp<0x140>(aly), bhy = row_dpp<0x140>(ahy);
  const double blz = row_dpp<0x140>(alz), bhz = row_dpp<0x140>(ahz);
  if ((lane & (kTile - 1)) == 0 && valid) {
    const long long t = t0 + (q * kWave + lane) / kTile;
    Box b;
    b.lx = fmin(alx, blx);
    b.hx = fmax(ahx, bhx);
    b.ly = fmin(aly, bly);
    b.hy = fmax(ahy, bhy);
    b.lz = fmin(alz, blz);
    b.hz = fmax(ahz, bhz);
    b.pad0 = 0.0;
    b.pad1 = 0.0;
    tbox[t] = b;
    if (hbox != nullptr) {
      // an empty second half keeps the (+inf, -inf) box: no point is
      // within any distance of it
      Box h;
      h.pad0 = 0.0;
      h.pad1 = 0.0;
      h.lx = alx;
      h.hx = ahx;
      h.ly = aly;
      h.hy = ahy;
      h.lz = alz;
      h.hz = ahz;
      hbox[2 * t] = h;
      h.lx = blx;
      h.hx = bhx;
      h.ly = bly;
      h.hy = bhy;
      h.lz = blz;
      h.hz = bhz;
      hbox[2 * t + 1] = h;
    }
  }
}

__global__ __launch_bounds__(256) void k_cell_box(
    const int* __restrict__ off, const int* __restrict__ tile_off,
    long long ncells, const Box* __restrict__ tbox, Box* __restrict__ cbox) {
  const long long g = blockIdx.x * static_cast<long long>(blockDim.x) +
                      threadIdx.x;
  if (g >= ncells) return;
  const int t0 = tile_off[g], t1 = tile_off[g + 1];
  if (t1 - t0 < 2) return;  // single-tile cells use the tile box
  Box b = tbox[t0];
  for (int t = t0 + 1; t < t1; ++t) {
    const Box o = tbox[t];
    b.lx = fmin(b.lx, o.lx);
    b.hx = fmax(b.hx, o.hx);
    b.ly = fmin(b.ly, o.ly);
    b.hy = fmax(b.hy, o.hy);
    b.lz = fmin(b.lz, o.lz);
    b.hz = fmax(b.hz, o.hz);
  }
  cbox[g] = b;
}

// ------------------------------------------------------------ heap

// Per-lane top-k in registers, ascending by d2 bits (d2 >= 0 or +inf for
// finite input, so the IEEE bit pattern orders like the value). Slots
// [0, KC - k) hold kUnusedBits (-inf: below every key both as a double and
// as a signed integer), and never leave; slots
// [KC - k, KC) are the k best so far, empty ones (+inf, kEmptyJ), so
// d[KC - 1] is the k-th best. Only d2 orders the heap: equal distances are
// resolved at the end (an equal pair inside the final k is ordered by
// index, an equal pair across the k-th defers the source to the
// exact-order walk).
template <int KC>
struct Top {
  long long d[KC];
  int j[KC];
};

template <int KC>
__device__ __forceinline__ void top_init(Top<KC>& h, int k) {
#pragma unroll
  for (int t = 0; t < KC; ++t) {
    h.d[t] = t < KC - k ? kUnusedBits : kInfBits;
    h.j[t] = kEmptyJ;
  }
}

// Key minimum and maximum (v_min_f64 / v_max_f64). Heap keys are never
// NaN, so these order exactly like the compare they replace.
__device__ __forceinline__ double key_min(double a, double b) {
  return __builtin_fmin(a, b);
}
__device__ __forceinline__ double key_max(double a, double b) {
  return __builtin_fmax(a, b);
}

// Compare-exchange of two (d2, index) slots. The keys are ordered as
// doubles (v_min_f64 / v_max_f64 and one compare for the indices): for the
// values a heap holds (-inf, d2 >= +0, +inf; never NaN, never -0) that is
// the order of their bit patterns as signed integers.
__device__ __forceinline__ void cswap(long long& da, int& ja, long long& db,
                                      int& jb, bool up) {
  const double a = as_f64(da), b = as_f64(db);
  const bool sw = up ? (a > b) : (a < b);
  const double lo = key_min(a, b), hi = key_max(a, b);
  da = __double_as_longlong(up ? lo : hi);
  db = __double_as_longlong(up ? hi : lo);
  const int tj = ja;
  ja = sw ? jb : ja;
  jb = sw ? tj : jb;
}

template <int KC>
__device__ __forceinline__ void bitonic_sort(long long (&d)[KC], int (&j)[KC]) {
#pragma unroll
  for (int size = 2; size <= KC; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
#pragma unroll
      for (int i = 0; i < KC; ++i) {
        const int p = i ^ stride;
        if (p > i) cswap(d[i], j[i], d[p], j[p], (i & size) == 0);
      }
    }
  }
}

// Ascending sort of 8 slots with Batcher's odd-even merge sort: 19
// compare-exchanges in 6 layers (the bitonic network takes 24). The buffer
// merge uses it unless LC_SORT8_OE is 0.
#ifndef LC_SORT8_OE
#define LC_SORT8_OE 1
#endif
__device__ __forceinline__ void oddeven_sort8(long long (&d)[8], int (&j)[8]) {
  auto cs = [&](int a, int b) { cswap(d[a], j[a], d[b], j[b], true); };
  cs(0, 1), cs(2, 3), cs(4, 5), cs(6, 7);
  cs(0, 2), cs(1, 3), cs(4, 6), cs(5, 7);
  cs(1, 2), cs(5, 6);
  cs(0, 4), cs(1, 5), cs(2, 6), cs(3, 7);
  cs(2, 4), cs(3, 5);
  cs(1, 2), cs(3, 4), cs(5, 6);
}

// Merge n buffered candidates (LDS column of this lane, stride kWave) into
// the top-k: sort the batch, keep the KC smallest of (top, batch) with the
// bitonic min trick (top ascending against batch descending), clean the
// bitonic result. Real entries that drop out feed min_rej, the smallest d2
// ever rejected, used for boundary-tie detection.
template <int KC>
__device__ __forceinline__ void batch_merge(Top<KC>& h, const long long* bd,
                                            const int* bj, int n,
                                            double& min_rej) {
  long long b[KC];
  int bx[KC];
#pragma unroll
  for (int t = 0; t < KC; ++t) {
    const bool v = t < n;
    b[t] = v ? bd[t * kWave] : kInfBits;
    bx[t] = v ? bj[t * kWave] : kEmptyJ;
  }
  bitonic_sort<KC>(b, bx);
  // empty slots and pads are (+inf, kEmptyJ): a dropped +inf leaves
  // min_rej unchanged
#pragma unroll
  for (int t = 0; t < KC; ++t) {
    const double x = as_f64(h.d[t]), y = as_f64(b[KC - 1 - t]);
    const int xj = h.j[t], yj = bx[KC - 1 - t];
    const bool ty = y < x;
    h.d[t] = __double_as_longlong(key_min(x, y));
    h.j[t] = ty ? yj : xj;
    min_rej = key_min(min_rej, key_max(x, y));
  }
#pragma unroll
  for (int stride = KC >> 1; stride > 0; stride >>= 1) {
#pragma unroll
    for (int i = 0; i < KC; ++i) {
      const int p = i ^ stride;
      if (p > i) cswap(h.d[i], h.j[i], h.d[p], h.j[p], true);
    }
  }
}

// Merge n buffered candidates into the top-k one at a time: each is
// inserted into the ascending register array (evicting the largest) or
// rejected. Few registers beyond the heap itself; the same kept set and
// the same min_rej as a batch merge.
template <int KC>
__device__ __forceinline__ void insert_merge(Top<KC>& h, const long long* bd,
                                             const int* bj, int n,
                                             double& min_rej) {
#pragma unroll 1
  for (int e = 0; e < n; ++e) {
    const long long x = bd[e * kWave];
    const int xj = bj[e * kWave];
    const long long top = h.d[KC - 1];
    if (x < top) {
      if (h.j[KC - 1] != kEmptyJ) min_rej = fmin(min_rej, as_f64(top));
#pragma unroll
      for (int t = KC - 1; t > 0; --t) {
        const bool up = h.d[t - 1] > x;
        const bool here = !up && h.d[t] > x;
        h.d[t] = up ? h.d[t - 1] : (here ? x : h.d[t]);
        h.j[t] = up ? h.j[t - 1] : (here ? xj : h.j[t]);
      }
      const bool here0 = h.d[0] > x;
      h.d[0] = here0 ? x : h.d[0];
      h.j[0] = here0 ? xj : h.j[0];
    } else {
      min_rej = fmin(min_rej, as_f64(x));
    }
  }
}

// insert_merge over up to 8 candidates already in registers.
template <int KC>
__device__ __forceinline__ void insert_merge_regs(Top<KC>& h, const long long (&b)[8],
                                                  const int (&bx)[8], int n,
                                                  double& min_rej) {
#pragma unroll
  for (int e = 0; e < 8; ++e) {
    if (e < n) {
      const long long x = b[e];
      const int xj = bx[e];
      const long long top = h.d[KC - 1];
      if (x < top) {
        if (h.j[KC - 1] != kEmptyJ) min_rej = fmin(min_rej, as_f64(top));
#pragma unroll
        for (int t = KC - 1; t > 0; --t) {
          const bool up = h.d[t - 1] > x;
          const bool here = !up && h.d[t] > x;
          h.d[t] = up ? h.d[t - 1] : (here ? x : h.d[t]);
          h.j[t] = up ? h.j[t - 1] : (here ? xj : h.j[t]);
        }
        const bool here0 = h.d[0] > x;
        h.d[0] = here0 ? x : h.d[0];
        h.j[0] = here0 ? xj : h.j[0];
      } else {
        min_rej = fmin(min_rej, as_f64(x));
      }
    }
  }
}

// Merge up to 8 buffered candidates into the top-k (KC >= 8) with a fixed
// network: sort the 8 (odd-even network), take the element-wise min of the top
// (ascending) and the batch reversed and padded with +inf (top[KC - 8 + t]
// against batch[7 - t]); that min holds the KC smallest of both and is
// bitonic, so one bitonic clean over all KC slots sorts it. Same kept set
// and min_rej as insert_merge, at a cost that does not grow with the
// number of entries.
template <int KC>
__device__ __forceinline__ void merge8_regs(Top<KC>& h, long long (&b)[8],
                                            int (&bx)[8], double& min_rej) {
#if LC_SORT8_OE
  oddeven_sort8(b, bx);
#else
  bitonic_sort<8>(b, bx);
#endif
  // an empty slot or pad is (+inf, kEmptyJ), so the d2 a pair drops feeds
  // min_rej unconditionally: +inf leaves it unchanged
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const double x = as_f64(h.d[KC - 8 + t]), y = as_f64(b[7 - t]);
    const int xj = h.j[KC - 8 + t], yj = bx[7 - t];
    const bool ty = y < x;
    h.d[KC - 8 + t] = __double_as_longlong(key_min(x, y));
    h.j[KC - 8 + t] = ty ? yj : xj;
    min_rej = key_min(min_rej, key_max(x, y));
  }
#pragma unroll
  for (int stride = KC >> 1; stride > 0; stride >>= 1) {
#pragma unroll
    for (int i = 0; i < KC; ++i) {
      const int p = i ^ stride;
      if (p > i) cswap(h.d[i], h.j[i], h.d[p], h.j[p], true);
    }
  }
}

// Tagged heap keys (row walk, heaps of 8 slots and more). A key is the d2
// bit pattern with its low kBits bits replaced by a slot tag, so every key
// of a merge is distinct and the merge network moves keys alone: one
// v_min_f64 and one v_max_f64 per compare-exchange, instead of those plus
// the compare and two selects that carry an index along. Each lane keeps,
// per slot, (index << kBits | the d2's own low bits) in LDS; after the
// network every output slot gathers its entry by the tag it carries, and
// the slots are tagged again in order. Two d2 that agree above the tag
// bits (within 2^kBits ulps) may leave the network in tag order rather
// than d2 order: inside the kept k the emit restores (d2, index) order,
// and when the bucket of the k-th holds a dropped key too the source goes
// to the exact-order walk, as for an exact boundary tie. A key set only
// ever grows above the k-th by keys of its own bucket or larger, so a
// swapped pair is always seen at that test.
#ifndef LC_TAG_MERGE
#define LC_TAG_MERGE 1
#endif
template <int KC>
struct TagKey {
  static constexpr int kBits = KC + 8 <= 32 ? 5 : 6;
  static constexpr long long kMask = (1ll << kBits) - 1;
  // an empty slot: the largest finite double, tag bits cleared (above any
  // d2 of a frame the tagged walk takes, tag_frame_ok); an unused slot (k
  // below the bucket size): -1.0, below any d2
  static constexpr long long kEmpty = 0x7fefffffffffffffll & ~kMask;
  static constexpr long long kUnused = static_cast<long long>(0xbff0000000000000ull);
  // indices stored with the low bits: frames of fewer points than this
  static constexpr long long kMaxN = 1ll << (32 - kBits);
};

template <int KC>
__device__ __forceinline__ void tag_init(Top<KC>& h, int k) {
#pragma unroll
  for (int t = 0; t < KC; ++t) {
    h.d[t] = (t < KC - k ? TagKey<KC>::kUnused : TagKey<KC>::kEmpty) | t;
  }
}

// Whether the tagged row walk can take a frame: every index fits beside
// the tag bits, and every d2 of the frame's walk (folded points and one
// lattice shift per axis, each coordinate difference below 9 times the
// largest cell entry) stays finite and below the empty key.
__device__ __forceinline__ bool tag_frame_ok(const FrameDesc& d, long long max_n) {
  double e = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) e = fmax(e, fabs(d.h[i]));
#pragma unroll
  for (int i = 0; i < 3; ++i) e = fmax(e, fabs(d.widths[i]));
  return static_cast<long long>(d.n) < max_n && e < 1e150;
}

// merge8_regs on tagged keys: the odd-even sort of the 8 buffered keys
// (tagged KC + t; pads carry the empty key), the element-wise min against
// the top 8 slots, the bitonic clean. A key's tag names its entry's slot,
// and entries stay where they are: a buffered key that displaces a heap key
// takes that key's slot (reloc copies its entry there, while the clean
// runs), so the kept keys' tags are always the slots 0 .. KC - 1. min_rej
// takes every key a pair drops.
// v_min_f64 / v_max_f64 of two keys, written out so that the compiler adds
// no quieting of signalling NaNs before them (keys are never NaN). Measured
// (walk, tagged keys): clustered 4M frame 3.86 against 3.93 ms (persistent
// launch), FCC 8M frame (k = 32) 8.86 against 8.95 ms.
#ifndef LC_TAG_ASM
#define LC_TAG_ASM 1
#endif
__device__ __forceinline__ double raw_min(double a, double b) {
#if LC_TAG_ASM
  double r;
  asm("v_min_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
#else
  return key_min(a, b);
#endif
}
__device__ __forceinline__ double raw_max(double a, double b) {
#if LC_TAG_ASM
  double r;
  asm("v_max_f64 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
#else
  return key_max(a, b);
#endif
}

// row_max16 of thresholds (never NaN: -inf, a key or a cap) through the
// raw v_max_f64, without the quieting the compiler puts before fmax
__device__ __forceinline__ double row_max16_thr(double v) {
  v = raw_max(v, row_dpp<0xB1>(v));
  v = raw_max(v, row_dpp<0x4E>(v));
  v = raw_max(v, row_dpp<0x141>(v));
  return raw_max(v, row_dpp<0x140>(v));
}

template <int KC, typename Reloc>
__device__ __forceinline__ void merge8_keys(long long (&hd)[KC], long long (&b)[8],
                                            double& min_rej, Reloc&& reloc) {
  constexpr long long M = TagKey<KC>::kMask;
  auto cs = [](long long& x, long long& y) LC_INLINE {
    const double a = as_f64(x), c = as_f64(y);
    x = __double_as_longlong(raw_min(a, c));
    y = __double_as_longlong(raw_max(a, c));
  };
  cs(b[0], b[1]), cs(b[2], b[3]), cs(b[4], b[5]), cs(b[6], b[7]);
  cs(b[0], b[2]), cs(b[1], b[3]), cs(b[4], b[6]), cs(b[5], b[7]);
  cs(b[1], b[2]), cs(b[5], b[6]);
  cs(b[0], b[4]), cs(b[1], b[5]), cs(b[2], b[6]), cs(b[3], b[7]);
  cs(b[2], b[4]), cs(b[3], b[5]);
  cs(b[1], b[2]), cs(b[3], b[4]), cs(b[5], b[6]);
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const long long x = hd[KC - 8 + t], y = b[7 - t];
    const double xf = as_f64(x), yf = as_f64(y);
    const bool take = yf < xf;  // keys are distinct: never equal
    min_rej = raw_min(min_rej, raw_max(xf, yf));
    const int ty = static_cast<int>(y & M), tx = static_cast<int>(x & M);
    hd[KC - 8 + t] = take ? ((y & ~M) | tx) : x;
    // branch-free: when the heap key stays, the buffered entry is copied
    // onto itself
    reloc(ty, take ? tx : ty);
  }
#pragma unroll
  for (int stride = KC >> 1; stride > 0; stride >>= 1) {
#pragma unroll
    for (int i = 0; i < KC; ++i) {
      const int p = i ^ stride;
      if (p > i) cs(hd[i], hd[p]);
    }
  }
}

template <int KC>
__device__ __forceinline__ void merge8(Top<KC>& h, const long long* bd,
                                       const int* bj, int n, double& min_rej) {
  long long b[8];
  int bx[8];
#pragma unroll
  for (int t = 0; t < 8; ++t) {
    const bool v = t < n;
    b[t] = v ? bd[t * kWave] : kInfBits;
    bx[t] = v ? bj[t * kWave] : kEmptyJ;
  }
  merge8_regs<KC>(h, b, bx, min_rej);
}

// Finish a source with a full heap: the row is exact unless a rejected or
// evicted d2 equals the k-th (a boundary tie, where the reference keeps
// whichever index it met first). Equal d2 inside the row is ordered by
// index as KHeap::finish does (src/knearest.rs:146-158). Returns false to
// defer the source.
// Output rows are written once and never read by the walk: non-temporal
// stores keep them from evicting candidate tiles (C3 walk 5.02 -> 4.97 ms).
#ifndef LC_NT_OUT
#define LC_NT_OUT 1
#endif
typedef int v4i __attribute__((ext_vector_type(4)));
typedef double v2d __attribute__((ext_vector_type(2)));

template <int KC>
__device__ __forceinline__ bool top_emit(Top<KC>& h, int k, double min_rej,
                                         int* out_nn, double* out_d2,
                                         size_t row) {
  if (min_rej == as_f64(h.d[KC - 1])) return false;
  bool tie = false;
#pragma unroll
  for (int t = 0; t + 1 < KC; ++t) {
    tie |= (t >= KC - k) && h.d[t] == h.d[t + 1];
  }
  if (tie) {
#pragma unroll 1
    for (int pass = 0; pass < KC; ++pass) {
#pragma unroll
      for (int t = 0; t + 1 < KC; ++t) {
        const bool sw = (t >= KC - k) && h.d[t] == h.d[t + 1] &&
                        h.j[t] > h.j[t + 1];
        const int a = h.j[t];
        h.j[t] = sw ? h.j[t + 1] : a;
        h.j[t + 1] = sw ? a : h.j[t + 1];
      }
    }
  }
  // k filling the bucket with 16-byte aligned rows (any k of 4, 8, 16 or
  // 32 with aligned outputs): whole rows as 16-byte vector stores
  int* const nn_row = out_nn + row;
  double* const d2_row = out_d2 != nullptr ? out_d2 + row : nullptr;
  const bool vec = (KC % 4 == 0) && k == KC &&
                   (reinterpret_cast<uintptr_t>(nn_row) & 15) == 0 &&
                   (reinterpret_cast<uintptr_t>(d2_row) & 15) == 0;
  if (vec) {
#pragma unroll
    for (int q = 0; q < KC / 4; ++q) {
      int v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int t = 4 * q + u;
        v[u] = h.j[t] != kEmptyJ ? h.j[t] : -1;
      }
#if LC_NT_OUT
      const v4i vv = {v[0], v[1], v[2], v[3]};
      __builtin_nontemporal_store(vv, reinterpret_cast<v4i*>(nn_row) + q);
#else
      reinterpret_cast<int4*>(nn_row)[q] = make_int4(v[0], v[1], v[2], v[3]);
#endif
    }
    if (d2_row != nullptr) {
#pragma unroll
      for (int q = 0; q < KC / 2; ++q) {
        double v[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const int t = 2 * q + u;
          v[u] = h.j[t] != kEmptyJ ? as_f64(h.d[t]) : nan_value();
        }
#if LC_NT_OUT
        const v2d vv = {v[0], v[1]};
        __builtin_nontemporal_store(vv, reinterpret_cast<v2d*>(d2_row) + q);
#else
        reinterpret_cast<double2*>(d2_row)[q] = make_double2(v[0], v[1]);
#endif
      }
    }
    return true;
  }
#pragma unroll
  for (int t = 0; t < KC; ++t) {
    if (t >= KC - k) {
      const int o = t - (KC - k);
      const bool real = h.j[t] != kEmptyJ;
      nn_row[o] = real ? h.j[t] : -1;
      if (d2_row != nullptr) d2_row[o] = real ? as_f64(h.d[t]) : nan_value();
    }
  }
  return true;
}

// Finish a source of the tagged walk: exact d2 and indices from the keys
// and the lane's LDS entries (pk[t]: the entry slot t's tag names), the boundary test on the
// tag bucket of the k-th, (d2, index) order inside the kept k, then
// top_emit. Returns false to defer the source.
template <int KC>
__device__ __forceinline__ bool tag_emit(const Top<KC>& h, const unsigned (&pk)[KC], int k,
                                         double min_rej, int* out_nn, double* out_d2,
                                         size_t row) {
  constexpr long long M = TagKey<KC>::kMask;
  if ((__double_as_longlong(min_rej) & ~M) == (h.d[KC - 1] & ~M)) return false;
  Top<KC> e;
#pragma unroll
  for (int t = 0; t < KC; ++t) {
    const long long key = h.d[t] & ~M;
    const bool unused = key < 0;
    const bool empty = key == TagKey<KC>::kEmpty;
    e.d[t] = unused ? kUnusedBits : (empty ? kInfBits : (key | (pk[t] & M)));
    e.j[t] = (unused || empty) ? kEmptyJ : static_cast<int>(pk[t] >> TagKey<KC>::kBits);
    // pk[t]: the entry of slot t's tag
  }
  bool dis = false;
#pragma unroll
  for (int t = 0; t + 1 < KC; ++t) dis |= (t >= KC - k) && e.d[t] > e.d[t + 1];
  if (dis) {
    // keys of one tag bucket in tag order: sort the kept k by d2 (the
    // bit patterns of d2 >= 0 order as integers)
#pragma unroll 1
    for (int pass = 0; pass < KC; ++pass) {
#pragma unroll
      for (int t = 0; t + 1 < KC; ++t) {
        const bool sw = (t >= KC - k) && e.d[t] > e.d[t + 1];
        const long long a = e.d[t];
        const int aj = e.j[t];
        e.d[t] = sw ? e.d[t + 1] : a;
        e.j[t] = sw ? e.j[t + 1] : aj;
        e.d[t + 1] = sw ? a : e.d[t + 1];
        e.j[t + 1] = sw ? aj : e.j[t + 1];
      }
    }
  }
  return top_emit<KC>(e, k, -__builtin_inf(), out_nn, out_d2, row);
}


__device__ __forceinline__ void defer_source(int s, int* defer_list,
                                             int* defer_count) {
  const int pos = atomicAdd(defer_count, 1);
  defer_list[pos] = s;
}

// Lower bound of (q + shift - p)^2 over a box [lo, hi] of folded points,
// evaluated with the walk's own rounded operations. Rounding is monotone,
// so for every q in the box fl(fl(q + s) - p) lies between the rounded
// corner differences and the bound never exceeds the d2 the walk would
// compute: pruning on it cannot change any result.
__device__ __forceinline__ double box_lb(double px, double py, double pz,
                                         const Box& b, double shx, double shy,
                                         double shz, bool zero) {
  double lx = b.lx, hx = b.hx, ly = b.ly, hy = b.hy, lz = b.lz, hz = b.hz;
  if (!zero) {
    lx = lx + shx;
    hx = hx + shx;
    ly = ly + shy;
    hy = hy + shy;
    lz = lz + shz;
    hz = hz + shz;
  }
  const double gx = fmax(fmax(lx - px, px - hx), 0.0);
  const double gy = fmax(fmax(ly - py, py - hy), 0.0);
  const double gz = fmax(fmax(lz - pz, pz - hz), 0.0);
  return gx * gx + gy * gy + gz * gz;
}

// Depth schedule buckets: register-heap bucket (4 values) x predicted
// depth 1..kSchedDepths, deepest first inside each heap bucket.
constexpr int kSchedDepths = 4;
constexpr int kSchedBuckets = 4 * kSchedDepths;

__host__ __device__ constexpr int sched_kc_index(int kc) {
  return kc == 4 ? 0 : (kc == 8 ? 1 : (kc == 16 ? 2 : 3));
}

struct WalkArgs {
  const FrameDesc* fr;
  const long long* fcb;
  int F;
  const int* fflags;
  const Cand* cand;
  const int* scell;
  const int* off;
  const int* tile_off;
  const int* items;
  const int* item_off;
  long long ncells;
  const Box* tbox;
  const Box* hbox;  // the two half-tile boxes of tile t at 2t, 2t + 1 (or null)
  const Box* cbox;
  int dense_min;
  int* defer_list;
  int* defer_count;
  // optional work counters (null = off): see kStat* indices
  unsigned long long* stats;
  // optional depth schedule of the lane-per-source walk (null = cell
  // order): source slots grouped by (heap bucket, predicted depth)
  const int* sched;
  const int* sched_count;  // kSchedBuckets run lengths
  const int* tile_cell;    // home cell of every 16-slot source tile
  // optional predicted k-th-neighbour distance per home cell and per
  // source tile (null = plain walk): the row walk's threshold cap
  const float* pred;
  const float* pred_tile;
  long long pred_tile_n;  // tiles beyond it are not capped
  // optional predicted k-th-neighbour distance per source slot (the
  // walk's cell-sorted order): the source's own threshold cap
  const float* pred_src;
  long long pred_src_n;
  // capped row walk: rows walked a second time without their cap are
  // counted here (null: not counted)
  int* retry_count;
  // depth schedule of the row walk (learned): rows list the cells of
  // shells 1 .. R in one pass, R = ceil(cap / cell_min) of the tile's
  // cap, and the stop test decides from shell R on; tile_perm (optional)
  // packs the tiles of every 16 by that depth
  int depth;
  const int* tile_perm;
  // persistent row walk (null: one wave per group of four tiles): waves
  // take groups from this counter until n_waves groups are taken
  int* queue;
  int n_waves;
};

enum {
  kStatDenseWaves = 0,
  kStatDenseTileTests,
  kStatDenseTiles,
  kStatDenseSteps,
  kStatDenseMerges,
  kStatDenseAppends,
  kStatSparseSources,
  kStatSparseTileTests,
  kStatSparseTiles,
  kStatSparseEvals,
  kStatSparseMerges,
  kStatRowSources,
  kStatRowTileTests,
  kStatRowTiles,
  kStatRowMerges,
  kStatRowRounds,   // wave-level tile rounds of the row walk
  kStatRowCells,    // wave-level candidate-cell visits of the row walk
  kStatRowRefines,  // row-level exact tile refinements
  kStatRowUseful,   // row-level tiles swept (the rest of the wave's rounds idle)
  kStatRowCycFill,  // wave-level cycles listing tiles (stats build only)
  kStatRowCycDrain, // wave-level cycles draining lists, merges included
  kStatRowCycMerge, // wave-level cycles merging buffers into the heaps
  kStatRowCycAll,   // wave-level cycles of the whole walk
  kStatRowFlushes,  // wave-level buffer merges
  kStatRowWavesR2,  // waves whose walk reaches shell 2
  kStatRowRoundsR2, // wave-level tile rounds in shells >= 2
  kStatRowUsefulR2, // row-level tiles swept in shells >= 2
  kStatRowCycR2,    // wave-level cycles in shells >= 2 (stats build only)
  kStatRowDead,     // row-level tiles swept in which no source accepted
  kStatRowAccepts,  // lane-level candidates accepted into the buffers
  kStatRowDeadHalf, // row-level 8-candidate halves of swept tiles in which no source accepted
  kStatRowCycOpen,  // wave-level cycles of the opening sweeps (own and next tile)
  kStatRowCycInit,  // wave-level cycles before the opening
  kStatRowIdleDone, // row-level drain rounds idle because the row has no live source
  kStatRowIdleEmpty,  // row-level drain rounds idle because the live row listed nothing
  kStatCount = 35
};

__device__ __forceinline__ void stat_add(unsigned long long* st, int i,
                                         unsigned long long v) {
  if (st != nullptr && v != 0) atomicAdd(&st[i], v);
}

// idx-th cell of Chebyshev shell r: r == 1 walks the near-first table,
// r > 1 the (2r+1)^3 cube lexicographically, skipping the interior.
// Returns false for interior cells.
__device__ __forceinline__ bool shell_offset(int r, int idx, int& dx, int& dy,
                                             int& dz) {
  if (r == 1) {
    near27(idx, dx, dy, dz);
    return true;
  }
  const int side = 2 * r + 1;
  dx = idx / (side * side) - r;
  dy = (idx / side) % side - r;
  dz = idx % side - r;
  return dx == -r || dx == r || dy == -r || dy == r || dz == -r || dz == r;
}

// Waves per block of the fast walks: the per-lane LDS buffers hold KC
// entries of 12 bytes, so the largest bucket runs two waves per block.
template <int KC>
struct WalkShape {
  static constexpr int kWaves = KC >= 32 ? 2 : 4;
  static constexpr int kGroup = KC >= 8 ? 4 : 2;  // candidates per buffer check
};

// ---------------------------------------------------------- dense walk

// One wave = up to 64 sources of one home cell (sources in sub-cell order,
// so the wave covers a compact part of the cell). Per candidate cell and
// tile, each lane tests the tile's exact lower bound against its k-th
// best; the wave skips the tile unless some lane needs it. Candidates are
// wave-uniform (scalar loads); a lane that accepts one appends it to its
// LDS buffer, and the buffers are merged into the register top-k in
// batches (all lanes at once, whenever one is nearly full).
template <int KC>
__global__ __launch_bounds__(WalkShape<KC>::kWaves* kWave) void k_walk_dense(
    WalkArgs A) {
  constexpr int W = WalkShape<KC>::kWaves;
  constexpr int G = WalkShape<KC>::kGroup;
  __shared__ long long s_bd[W][KC][kWave];
  __shared__ int s_bj[W][KC][kWave];
  const int wv = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const int w = uni(static_cast<int>(blockIdx.x * W + wv));
  const int total = uni(A.item_off[A.ncells]);
  if (w >= total) return;
  const int g = uni(A.items[w]);
  const int o0 = uni(A.off[g]);
  const int cnt = uni(A.off[g + 1]) - o0;
  if (cnt < A.dense_min) return;
  const int q = w - uni(A.item_off[g]);
  const int f = uni(find_frame(A.fcb, A.F, g));
  const FrameDesc& d = A.fr[f];
  if (d.kc != KC) return;
  const int ls = q * kWave + lane;
  const bool valid = ls < cnt;
  const int s = o0 + (valid ? ls : 0);
  if ((A.fflags[f] & kFrameNonFinite) != 0 || d.r_safe < 1) {
    if (valid) defer_source(s, A.defer_list, A.defer_count);
    return;
  }
  const int k = d.k;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const long long cb = d.cell_base;
  const int c = g - static_cast<int>(cb);
  const int hx = c % nx;
  const int hy = (c / nx) % ny;
  const int hz = c / (nx * ny);
  const Cand mc = A.cand[s];
  const double px = mc.x, py = mc.y, pz = mc.z;
  const int me = static_cast<int>(mc.id);
  const int ortho = d.ortho;
  const double w0 = d.widths[0], w1 = d.widths[1], w2 = d.widths[2];

  Top<KC> h;
  top_init<KC>(h, k);
  double min_rej = as_f64(kInfBits);
  double thr = as_f64(kInfBits);
  bool live = valid;
  int nbuf = 0;
  int done_reach = 0;
  long long* bd = &s_bd[wv][0][lane];
  int* bj = &s_bj[wv][0][lane];
  unsigned st_tests = 0, st_tiles = 0, st_steps = 0, st_merges = 0,
           st_appends = 0;

  auto flush = [&]() LC_INLINE {
    st_merges += 1;
    if (nbuf > 0) batch_merge<KC>(h, bd, bj, nbuf, min_rej);
    nbuf = 0;
    thr = as_f64(h.d[KC - 1]);
  };

  // Tiles of one candidate cell, starting at tile `start`. The next
  // tile's box and the next group of candidates are loaded before the
  // current ones are used, so the scalar loads overlap the arithmetic.
  // Z: the cell needs no lattice shift (q + 0 == q for the d2 bits).
  auto walk_tiles = [&](auto ztag, int a0, int n2, int t0, int nt, int start,
                        bool home, double shx, double shy, double shz) LC_INLINE {
    constexpr bool Z = decltype(ztag)::value;
    int t = start;
    Box bx = A.tbox[t0 + t];
#pragma unroll 1
    for (int tt = 0; tt < nt; ++tt) {
      int tn = t + 1;
      tn = tn >= nt ? tn - nt : tn;
      const Box bnext = A.tbox[t0 + tn];
      const double lb = box_lb(px, py, pz, bx, shx, shy, shz, Z);
      const bool need = live && lb <= thr;
      st_tests += 1;
      if (__ballot(need)) {
        st_tiles += 1;
        const int base = a0 + t * kTile;
        const int e = min(kTile, n2 - t * kTile);
        const int last = base + e - 1;
        Cand c[G];
#pragma unroll
        for (int u = 0; u < G; ++u) c[u] = A.cand[min(base + u, last)];
#pragma unroll 1
        for (int hh = 0; hh < e; hh += G) {
          Cand nxt[G];
#pragma unroll
          for (int u = 0; u < G; ++u) nxt[u] = A.cand[min(base + hh + G + u, last)];
          const int lim = e - hh;
          st_steps += lim < G ? lim : G;
#pragma unroll
          for (int u = 0; u < G; ++u) {
            const double qx = Z ? c[u].x : c[u].x + shx;
            const double qy = Z ? c[u].y : c[u].y + shy;
            const double qz = Z ? c[u].z : c[u].z + shz;
            const int j = static_cast<int>(c[u].id);
            const double ddx = qx - px, ddy = qy - py, ddz = qz - pz;
            const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
            const bool acc = u < lim && need && d2 <= thr && !(home && j == me);
            if (acc) {
              bd[nbuf * kWave] = __double_as_longlong(d2);
              bj[nbuf * kWave] = j;
              nbuf += 1;
              st_appends += 1;
            }
          }
          if (__ballot(nbuf > KC - G)) flush();
#pragma unroll
          for (int u = 0; u < G; ++u) c[u] = nxt[u];
        }
      }
      t = tn;
      bx = bnext;
    }
  };

  // Visit one candidate cell (unwrapped offset dx, dy, dz from home).
  auto visit = [&](int dx, int dy, int dz, bool home) LC_INLINE {
    const int jx = hx + dx, jy = hy + dy, jz = hz + dz;
    const int cx = jx < 0 ? jx + nx : (jx >= nx ? jx - nx : jx);
    const int cy = jy < 0 ? jy + ny : (jy >= ny ? jy - ny : jy);
    const int cz = jz < 0 ? jz + nz : (jz >= nz ? jz - nz : jz);
    const int na = jx < 0 ? -1 : (jx >= nx ? 1 : 0);
    const int nb = jy < 0 ? -1 : (jy >= ny ? 1 : 0);
    const int nc = jz < 0 ? -1 : (jz >= nz ? 1 : 0);
    const int gq = uni(static_cast<int>(cb) + (cz * ny + cy) * nx + cx);
    const int a0 = uni(A.off[gq]);
    const int n2 = uni(A.off[gq + 1]) - a0;
    if (n2 == 0) return;
    const bool zero = (na | nb | nc) == 0;
    double shx = 0.0, shy = 0.0, shz = 0.0;
    if (!zero) {
      if (ortho) {
        shx = static_cast<double>(na) * w0;
        shy = static_cast<double>(nb) * w1;
        shz = static_cast<double>(nc) * w2;
      } else {
        const double fa = na, fb = nb, fc = nc;
        shx = fa * d.h[0] + fb * d.h[3] + fc * d.h[6];
        shy = fa * d.h[1] + fb * d.h[4] + fc * d.h[7];
        shz = fa * d.h[2] + fb * d.h[5] + fc * d.h[8];
      }
    }
    const int nt = (n2 + kTile - 1) >> kTileShift;
    if (nt > 1) {
      const double lb = box_lb(px, py, pz, A.cbox[gq], shx, shy, shz, zero);
      if (!__ballot(live && lb <= thr)) return;
    }
    const int t0 = uni(A.tile_off[gq]);
    int start = 0;
    if (home) {
      start = q * (kWave / kTile);
      while (start >= nt) start -= nt;
    }
    if (zero) {
      walk_tiles(std::true_type{}, a0, n2, t0, nt, start, home, shx, shy, shz);
    } else {
      walk_tiles(std::false_type{}, a0, n2, t0, nt, start, home, shx, shy, shz);
    }
  };

#pragma unroll 1
  for (int r = 1; r <= d.r_safe; ++r) {
    const int side = 2 * r + 1;
    const int ncube = side * side * side;
#pragma unroll 1
    for (int idx = 0; idx < ncube; ++idx) {
      int dx, dy, dz;
      if (!shell_offset(r, idx, dx, dy, dz)) continue;
      visit(dx, dy, dz, r == 1 && idx == 0);
    }
    if (__ballot(nbuf > 0)) flush();
    // stop test, src/knearest.rs:466-471
    if (live && h.j[KC - 1] != kEmptyJ) {
      const double bound = static_cast<double>(r) * d.cell_min;
      if (as_f64(h.d[KC - 1]) <= bound * bound) {
        live = false;
        done_reach = r;
      }
    }
    if (!__ballot(live)) break;
  }
  if (A.stats != nullptr) {
    unsigned long long app = st_appends;
    for (int dd = 32; dd >= 1; dd >>= 1) app += __shfl_xor(app, dd, kWave);
    if (lane == 0) {
      stat_add(A.stats, kStatDenseWaves, 1);
      stat_add(A.stats, kStatDenseTileTests, st_tests);
      stat_add(A.stats, kStatDenseTiles, st_tiles);
      stat_add(A.stats, kStatDenseSteps, st_steps);
      stat_add(A.stats, kStatDenseMerges, st_merges);
      stat_add(A.stats, kStatDenseAppends, app);
    }
  }
  if (!valid) return;
  const size_t row = static_cast<size_t>(me) * k;
  if (done_reach == 0 ||
      !top_emit<KC>(h, k, min_rej, d.out_nn, d.out_d2, row)) {
    defer_source(s, A.defer_list, A.defer_count);
    return;
  }
  if (d.out_reach) d.out_reach[me] = done_reach;
}

// --------------------------------------------------------- sparse walk

// One lane per source of a cell below the dense threshold, in cell order
// (neighbouring lanes share home or adjacent cells). Each lane walks its
// own stencil: the home cell from its own 16-point tile outwards, then
// the near-first neighbour order; cells and tiles whose exact lower bound
// exceeds the lane's k-th best are skipped. Accepted candidates go to the
// lane's LDS buffer; whenever any lane's buffer is nearly full, every lane
// with entries merges its batch.
template <int KC>
__global__ __launch_bounds__(WalkShape<KC>::kWaves* kWave) void k_walk_sparse(
    WalkArgs A) {
  constexpr int W = WalkShape<KC>::kWaves;
  constexpr int G = WalkShape<KC>::kGroup;
  __shared__ long long s_bd[W][KC][kWave];
  __shared__ int s_bj[W][KC][kWave];
  const int wv = threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  const unsigned long long t_begin = wall_clock64();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (A.stats != nullptr && lane == 0) {
    // per-wave timeline (100 MHz wall clock) after the counters; the
    // end stamp is written by the wave's last lane to finish
    const long long wid = blockIdx.x * static_cast<long long>(W) + wv;
    A.stats[kStatCount + 2 * wid] = t_begin;
  }
  int s = i;
  if (A.sched != nullptr) {
    // this heap bucket's run of the depth schedule
    const int q0 = sched_kc_index(KC) * kSchedDepths;
    int lo = 0, len = 0;
    for (int q = 0; q < q0 + kSchedDepths; ++q) {
      const int c = A.sched_count[q];
      if (q < q0) lo += c; else len += c;
    }
    if (i >= len) return;
    s = A.sched[lo + i];
  } else if (i >= A.off[A.ncells]) {
    return;
  }
  const int g = A.scell[s];
  const int o0 = A.off[g];
  const int cnt = A.off[g + 1] - o0;
  if (cnt >= A.dense_min) return;
  const int f = find_frame(A.fcb, A.F, g);
  const FrameDesc& d = A.fr[f];
  if (d.kc != KC) return;
  if ((A.fflags[f] & kFrameNonFinite) != 0 || d.r_safe < 1) {
    defer_source(s, A.defer_list, A.defer_count);
    return;
  }
  const int k = d.k;
  const int nx = d.nx, ny = d.ny, nz = d.nz;
  const long long cb = d.cell_base;
  const int c = g - static_cast<int>(cb);
  const int hx = c % nx;
  const int hy = (c / nx) % ny;
  const int hz = c / (nx * ny);
  const Cand mc = A.cand[s];
  const double px = mc.x, py = mc.y, pz = mc.z;
  const int me = static_cast<int>(mc.id);
  const int own_tile = (s - o0) >> kTileShift;
  const int ortho = d.ortho;
  const double w0 = d.widths[0], w1 = d.widths[1], w2 = d.widths[2];
  double hm[9];
#pragma unroll
  for (int t = 0; t < 9; ++t) hm[t] = d.h[t];

  Top<KC> h;
  top_init<KC>(h, k);
  double min_rej = as_f64(kInfBits);
  double thr = as_f64(kInfBits);
  int nbuf = 0;
  int done_reach = 0;
  long long* bd = &s_bd[wv][0][lane];
  int* bj = &s_bj[wv][0][lane];
  unsigned st_tests = 0, st_tiles = 0, st_evals = 0, st_merges = 0;

  auto flush = [&]() LC_INLINE {
    if (nbuf > 0) {
      st_merges += 1;
      batch_merge<KC>(h, bd, bj, nbuf, min_rej);
      nbuf = 0;
      thr = as_f64(h.d[KC - 1]);
    }
  };

  // One lane's tiles of one candidate cell from tile `start`; the next
  // group of candidates is loaded before the current group is used.
  auto walk_tiles = [&](auto ztag, int a0, int n2, int t0, int nt, int start,
                        bool home, double shx, double shy, double shz) LC_INLINE {
    constexpr bool Z = decltype(ztag)::value;
#pragma unroll 1
    for (int tt = 0; tt < nt; ++tt) {
      int t = start + tt;
      t = t >= nt ? t - nt : t;
      if (nt > 1 || !home) {
        const double lb = box_lb(px, py, pz, A.tbox[t0 + t], shx, shy, shz, Z);
        st_tests += 1;
        if (lb > thr) continue;
      }
      st_tiles += 1;
      const int base = a0 + t * kTile;
      const int e = min(kTile, n2 - t * kTile);
      const int last = base + e - 1;
      Cand c[G];
#pragma unroll
      for (int u = 0; u < G; ++u) c[u] = A.cand[min(base + u, last)];
#pragma unroll 1
      for (int hh = 0; hh < e; hh += G) {
        Cand nxt[G];
#pragma unroll
        for (int u = 0; u < G; ++u) nxt[u] = A.cand[min(base + hh + G + u, last)];
        const int lim = e - hh;
#pragma unroll
        for (int u = 0; u < G; ++u) {
          const double qx = Z ? c[u].x : c[u].x + shx;
          const double qy = Z ? c[u].y : c[u].y + shy;
          const double qz = Z ? c[u].z : c[u].z + shz;
          const int j = static_cast<int>(c[u].id);
          const double ddx = qx - px, ddy = qy - py, ddz = qz - pz;
          const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
          const bool acc = u < lim && d2 <= thr && j != me;
          if (acc) {
            bd[nbuf * kWave] = __double_as_longlong(d2);
            bj[nbuf * kWave] = j;
            nbuf += 1;
          }
        }
        st_evals += lim < G ? lim : G;
        if (__ballot(nbuf > KC - G)) flush();
#pragma unroll
        for (int u = 0; u < G; ++u) c[u] = nxt[u];
      }
    }
  };

  auto visit = [&](int dx, int dy, int dz, bool home) LC_INLINE {
    const int jx = hx + dx, jy = hy + dy, jz = hz + dz;
    const int cx = jx < 0 ? jx + nx : (jx >= nx ? jx - nx : jx);
    const int cy = jy < 0 ? jy + ny : (jy >= ny ? jy - ny : jy);
    const int cz = jz < 0 ? jz + nz : (jz >= nz ? jz - nz : jz);
    const int na = jx < 0 ? -1 : (jx >= nx ? 1 : 0);
    const int nb = jy < 0 ? -1 : (jy >= ny ? 1 : 0);
    const int nc = jz < 0 ? -1 : (jz >= nz ? 1 : 0);
    const int gq = static_cast<int>(cb) + (cz * ny + cy) * nx + cx;
    const int a0 = A.off[gq];
    const int n2 = A.off[gq + 1] - a0;
    if (n2 == 0) return;
    const bool zero = (na | nb | nc) == 0;
    double shx = 0.0, shy = 0.0, shz = 0.0;
    if (!zero) {
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
    }
    const int nt = (n2 + kTile - 1) >> kTileShift;
    if (nt > 1) {
      const double lb = box_lb(px, py, pz, A.cbox[gq], shx, shy, shz, zero);
      if (lb > thr) return;
    }
    const int t0 = A.tile_off[gq];
    const int start = home ? own_tile : 0;
    if (zero) {
      walk_tiles(std::true_type{}, a0, n2, t0, nt, start, home, shx, shy, shz);
    } else {
      walk_tiles(std::false_type{}, a0, n2, t0, nt, start, home, shx, shy, shz);
    }
  };

#pragma unroll 1
  for (int r = 1; r <= d.r_safe; ++r) {
    const int side = 2 * r + 1;
    const int ncube = side * side * side;
#pragma unroll 1
    for (int idx = 0; idx < ncube; ++idx) {
      int dx, dy, dz;
      if (!shell_offset(r, idx, dx, dy, dz)) continue;
      visit(dx, dy, dz, r == 1 && idx == 0);
    }
    flush();
    if (h.j[KC - 1] != kEmptyJ) {
      const double bound = static_cast<double>(r) * d.cell_min;
      if (as_f64(h.d[KC - 1]) <= bound * bound) {
        done_reach = r;
        break;
      }
    }
  }
  if (A.stats != nullptr) {
    const long long wid = blockIdx.x * static_cast<long long>(W) + wv;
    atomicMax(&A.stats[kStatCount + 2 * wid + 1], wall_clock64());
    stat_add(A.stats, kStatSparseSources, 1);
    stat_add(A.stats, kStatSparseTileTests, st_tests);
    stat_add(A.stats, kStatSparseTiles, st_tiles);
    stat_add(A.stats, kStatSparseEvals, st_evals);
    stat_add(A.stats, kStatSparseMerges, st_merges);
  }
  const size_t row = static_cast<size_t>(me) * k;
  if (done_reach == 0 ||
      !top_emit<KC>(h, k, min_rej, d.out_nn, d.out_d2, row)) {
    defer_source(s, A.defer_list, A.defer_count);
    return;
  }
  if (d.out_reach) d.out_reach[me] = done_reach;
}

// ------------------------------------------------------------- row walk

// Lane I of each 16-lane row, broadcast to the whole row (DPP
// row_newbcast: one VALU move per 64-bit value, no LDS traffic).
template <int I>
__device__ __forceinline__ double row_bcast(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const long long zero = 0;
  const long long r =
      __builtin_amdgcn_update_dpp(zero, b, 0x150 + I, 0xf, 0xf, true);
  return __builtin_bit_cast(double, r);
}
template <int I>
__device__ __forceinline__ int row_bcast(int v) {
  const int zero = 0;
  return __builtin_amdgcn_update_dpp(zero, v, 0x150 + I, 0xf, 0xf, true);
}

// FP32 listing test. The per-source tile test of the row walk's fill only
// prunes: a tile it keeps is swept with the exact FP64 arithmetic, so it
// may keep more than the FP64 test and never less. In FP32 the source
// broadcast folds into the subtractions (v_sub_f32_dpp / v_subrev_f32_dpp
// with row_newbcast) and each axis is one v_max3_f32, about 15 VALU per
// source against 27 in FP64. It keeps a tile whenever the FP64 test does:
//  * every coordinate the test sees (folded points, shifted boxes) is
//    below C = sum |origin| + 2 (sum |h| + sum |widths|) in magnitude, and
//    the box is widened by eps = C 2^-21, four float ulps at that scale, so
//    a gap fl32(fl32(lo) - eps - fl32(p)) is at least two ulps below the
//    real gap lo - p and below the FP64 gap: gx32 <= gx64 per axis;
//  * the squared sum then carries at most 3 2^-24 relative error
//    (ls32 <= ls64 (1 + 2^-21)), and the threshold is the float at or above
//    it times (1 + 2^-19), so ls32 above it puts ls64 above the threshold;
//  * the list key is the smallest ls32 of the sources that keep the tile,
//    at most FLT_MAX, times (1 - 2^-20): at or below the FP64 bound it
//    replaces (a sum that overflows to +inf in FP32 has an FP64 bound of
//    FLT_MAX or more; the key must not be +inf, which the drain would take
//    for a tile beyond every threshold).
// Coordinates beyond float range are clamped to +-FLT_MAX after the
// conversion: the clamp is monotone and shrinks every difference, so gaps
// only fall, and no infinity - infinity NaN can arise. Flushing FP32
// subnormals to zero only lowers a gap or a sum. Measured (walk): clustered
// 4M frame 3.65 against 3.85 ms (persistent launch), FCC 8M frame (k = 32)
// 8.56 against 8.87 ms, with 0.1% more tiles swept.
#ifndef LC_LIST32
#define LC_LIST32 1
#endif
template <int I>
__device__ __forceinline__ float row_bcast_f32(float v) {
  return __builtin_bit_cast(
      float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x150 + I, 0xf, 0xf, true));
}
// a coordinate as a float, clamped to the finite range (v_med3_f32)
__device__ __forceinline__ float coord32(double x) {
  return __builtin_amdgcn_fmed3f(static_cast<float>(x), -3.40282347e38f, 3.40282347e38f);
}
// a bound as the list key's double: at most FLT_MAX, scaled down
__device__ __forceinline__ double key_of32(float b) {
  return fmin(static_cast<double>(b), 3.4028234663852886e38) * (1.0 - 0x1p-20);
}
// the float at or above t (t >= 0 or -inf), times (1 + 2^-19)
__device__ __forceinline__ float thr_up32(double t) {
  float f = static_cast<float>(t);
  if (static_cast<double>(f) < t) f = __int_as_float(__float_as_int(f) + 1);
  return f * (1.0f + 0x1p-19f);
}
// eps of the frame's FP32 listing test (see above)
__device__ __forceinline__ float list_eps32(const FrameDesc& d) {
  double c = 0.0;
#pragma unroll
  for (int i = 0; i < 9; ++i) c += fabs(d.h[i]);
#pragma unroll
  for (int i = 0; i < 3; ++i) c += fabs(d.widths[i]);
  c = 2.0 * c;
#pragma unroll
  for (int i = 0; i < 3; ++i) c += fabs(d.origin[i]);
  return static_cast<float>(c * 0x1p-21) * 1.0625f;
}

// Lower bound of d2 between any point of the source box [sl, sh] and any
// point of box b shifted by s, with the walk's rounded operations (see
// box_lb: fl(fl(q + s) - p) is monotone in q and in p, so the bound never
// exceeds a d2 the walk computes for a pair inside the boxes).
__device__ __forceinline__ double hull_lb(const double (&sl)[3],
                                          const double (&sh)[3], const Box& b,
                                          double shx, double shy, double shz) {
  const double lx = b.lx + shx, hx = b.hx + shx;
  const double ly = b.ly + shy, hy = b.hy + shy;
  const double lz = b.lz + shz, hz = b.hz + shz;
  const double gx = fmax(fmax(lx - sh[0], sl[0] - hx), 0.0);
  const double gy = fmax(fmax(ly - sh[1], sl[1] - hy), 0.0);
  const double gz = fmax(fmax(lz - sh[2], sl[2] - hz), 0.0);
  return gx * gx + gy * gy + gz * gz;
}

// i-th cell of Chebyshev shell r: shell 1 from the near-first table, the
// surface of shell r > 1 in lexicographic (dx, dy, dz) order.
__device__ __forceinline__ int shell_cells(int r) {
  const int s = 2 * r + 1;
  return r == 1 ? 27 : s * s * s - (s - 2) * (s - 2) * (s - 2);
}
__device__ __forceinline__ void surface_cell(int r, int i, int& dx, int& dy,
                                             int& dz) {
  if (r == 1) {
    near27(i, dx, dy, dz);
    return;
  }
  const int s = 2 * r + 1;
  const int face = s * s;
  const int ring = 4 * s - 4;
  if (i < face) {
    dx = -r;
    dy = -r + i / s;
    dz = -r + i % s;
    return;
  }
  i -= face;
  if (i < (s - 2) * ring) {
    dx = -r + 1 + i / ring;
    int q = i % ring;
    if (q < s) {
      dy = -r;
      dz = -r + q;
    } else if (q < s + 2 * (s - 2)) {
      q -= s;
      dy = -r + 1 + q / 2;
      dz = (q & 1) ? r : -r;
    } else {
      q -= s + 2 * (s - 2);
      dy = r;
      dz = -r + q;
    }
    return;
  }
  i -= (s - 2) * ring;
  dx = r;
  dy = -r + i / s;
  dz = -r + i % s;
}

// i-th tile of a cell of nt tiles in order of distance from tile tl in the
// cell's sub-cell order: tl, tl + 1, tl - 1, tl + 2, ... while both sides
// last, then the longer side (a bijection of 0 .. nt - 1).
#ifndef LC_HOME_SPIRAL
#define LC_HOME_SPIRAL 1
#endif
__device__ __forceinline__ int spiral_tile(int i, int tl, int nt) {
  const int left = tl, right = nt - 1 - tl;
  const int a = left < right ? left : right;
  if (i <= 2 * a) return tl + ((i & 1) ? (i + 1) / 2 : -(i / 2));
  return right > left ? tl + (i - a) : tl - (i - a);
}

// First wave of frame f in the row walk: frames start on wave boundaries
// so that every wave's frame data is uniform. ceil(t/4) + f leaves at
// most one spare wave per frame and needs no scan.
__device__ __forceinline__ int row_wave0(const int* tile_off,
                                         const long long* fcb, int f) {
  return (tile_off[fcb[f]] + 3) / 4 + f;
}

// One 16-lane row per source tile (16 consecutive sources of one home
// cell, in sub-cell order), four rows per wave, every wave inside one
// frame. Per shell,
// a row lists the candidate tiles some of its sources still need: 16 cells
// (one per lane) get their headers at once, then 16 of their tiles (one per
// lane) get a bounding-box test against the row's sources and, when that
// passes, an exact per-source test (each source broadcast to the row). The
// list, in LDS, is then drained tile by tile: each lane loads one
// candidate of the tile, adds the lattice shift, and the 16 candidates
// reach the row's lanes through DPP broadcasts, while the next tile's loads
// are in flight. The row's own tile, already in registers, and the two
// tiles of its cell on either side of it in sub-cell order open the walk,
// so the tests start from a finite k-th distance.
// A row stops listing tiles once its list holds LC_ROW_FILL entries, and
// the wave drains. LC_ROW_SELECT drains each row's list nearest tile first,
// by a lower bound of the tile's distance to the row's sources that can
// still take it, and drops the rest of the list once the nearest remaining
// bound exceeds the row's threshold: the tiles that fill the heap come
// first, whatever the listing order, and the far ones are never swept.
// Measured (walk, clustered 4M frame / FCC 8M frame): list order with a
// fill of 4, 4.57 / 10.93 ms; nearest first with a fill of 4, 4.27 / 10.56;
// 8, 4.27 / 10.66; 16, 4.07 / 10.36; 24, 4.16 / 10.40; 32, 4.22 / 11.08;
// 48, 4.23 / 10.70. A list holds at most LC_ROW_FILL - 1 + 16 entries,
// the row walk's list capacity.
#ifndef LC_ROW_FILL
#define LC_ROW_FILL 16
#endif
// Cells of more than one tile are tested against the row's bounding box
// before their tiles are (one more dependent load per cell chunk). Without
// the test (walk): clustered 4M frame 3.75 against 3.68 ms, FCC 8M frame
// 8.50 against 8.58 ms, whose cells rarely fail it.
#ifndef LC_CELL_BOX_TEST
#define LC_CELL_BOX_TEST 1
#endif
#ifndef LC_ROW_SELECT
#define LC_ROW_SELECT 1
#endif
// Heaps of LC_ROW_HALF_KC slots and more list half tiles (8 members, each
// with its own box) and sweep two per round, lanes 0-7 and 8-15 of the row:
// a half whose box lies beyond every source's threshold is never swept.
// The list then holds up to LC_ROW_FILL_HALF - 1 + 32 half tiles. Measured
// (walk): FCC 8M frame, k = 32, 10.33 ms with whole tiles, 9.79 with half
// tiles (28% fewer rounds); clustered 4M frame, k = 16, 4.08 ms with whole
// tiles, 4.23 with half tiles (20 registers spilled at 3 waves per SIMD)
// and 5.41 with half tiles at 2 waves per SIMD, so the 16-slot heap keeps
// whole tiles. LC_ROW_HALF 0 turns them off.
#ifndef LC_ROW_HALF
#define LC_ROW_HALF 1
#endif
#ifndef LC_ROW_HALF_KC
#define LC_ROW_HALF_KC 32
#endif
#ifndef LC_ROW_FILL_HALF
#define LC_ROW_FILL_HALF 32
#endif
// Selection keys held in registers through a drain (read from LDS once, at
// its start) rather than read from LDS every round: 4.08 against 4.09 ms
// (clustered 4M), 9.79 against 10.07 (FCC 8M, half tiles).
#ifndef LC_KEYS_REG
#define LC_KEYS_REG 1
#endif
#ifndef LC_RAW_THR
#define LC_RAW_THR 1
#endif
#if LC_RAW_THR && LC_TAG_ASM
#define LC_ROW_MAX_THR row_max16_thr
#define LC_THR_MIN raw_min
#else
#define LC_ROW_MAX_THR row_max16
#define LC_THR_MIN fmin
#endif
#ifndef LC_APPEND_ALL
#define LC_APPEND_ALL 1
#endif
// Resident 256-thread blocks per CU the register budget is cut for: 3 (3
// waves per SIMD, 168 VGPRs) for the heaps of up to 16, 2 (256 VGPRs) for
// the 32-slot heap. Measured on the clustered 4M (k = 16), the uneven batch
// and the FCC 8M (k = 32) workloads: at 4 the spills of the larger heaps
// cost more than the extra wave hides, at 2 the k <= 16 walks hide too
// little latency; the 32-slot heap spills 215 VGPRs at 3 (FCC 16.3 ms)
// and 2 at 2 (12.7 ms).
#ifndef LC_ROW_MINB32
#define LC_ROW_MINB32 2
#endif
template <int KC>
struct RowOcc {
  static constexpr int kMinBlocks = KC == 32 ? LC_ROW_MINB32 : 3;
};
#ifdef LC_ROW_MINB
#define LC_ROW_BLOCKS(KC) LC_ROW_MINB
#else
#define LC_ROW_BLOCKS(KC) RowOcc<KC>::kMinBlocks
#endif
// Per-wave candidate buffers of the row walk (4 waves, 8 buffered candidates
// per lane): 16-byte records (d2 bits, index) for the plain heap; for the
// tagged heap, the d2 bits and each lane's slot entries, the KC heap slots
// (index << tag bits | the d2's low bits) and then the 8 buffer slots
// (indices until a merge packs them). Measured in LDS per block: 32 KB,
// or 16 KB + (KC + 8) KB for the tagged heap (40 KB at KC = 16, which
// still leaves 3 blocks per CU).
template <int KC, bool kTag>
struct RowBufs;
template <int KC>
struct RowBufs<KC, false> {
  int4 buf[4][8][kWave];
};
template <int KC>
struct RowBufs<KC, true> {
  // per wave, slots of one dword per lane: the KC + 8 entries, then the low
  // and the high halves of the 8 buffered d2 (one slot stride for all, so
  // an append's three stores share one address)
  unsigned dw[4][KC + 24][kWave];
};

// kDepth: the depth schedule (rows list shells 1 .. predicted depth in one
// pass); without it every row of a wave is in the same shell. kPersist: the
// grid fills the GPU once and every wave takes groups of four tiles from a
// device counter until none is left (one persistent launch per heap
// bucket); without it each wave walks the group its position names.
template <int KC, bool kStats, bool kDepth, bool kPersist>
__global__ __launch_bounds__(256, LC_ROW_BLOCKS(KC)) void k_walk_rows(WalkArgs A) {
  constexpr int W = 4;   // waves per block
  constexpr int BC = 8;  // buffered candidates per lane between merges
  // a wave-wide "any lane accepts" branch around each step's append: it
  // pays for the heaps of up to 16 (C3 walk 5.27 ms against 5.40 without)
  // and not for the 32-slot heap (FCC 11.15 ms against 10.99)
  constexpr bool kStepBranch = KC < 32;
  constexpr bool kHalf = LC_ROW_HALF != 0 && KC >= LC_ROW_HALF_KC;
  static_assert(!kHalf || LC_ROW_SELECT, "half tiles are drained by selection");
  // list length at which a drain starts
  constexpr int kFill = kHalf ? LC_ROW_FILL_HALF : LC_ROW_FILL;
  // tile list entries per row: a fill round starts below kFill entries and
  // appends at most 16 (32 half tiles)
  constexpr int LC = kFill - 1 + (kHalf ? 32 : 16);
  static_assert(kFill >= 1 && LC <= 64, "tile list capacity");
  constexpr int NK = (LC + 15) / 16;  // list keys per lane when selecting
  // tagged heap keys for heaps of 8 slots and more
  constexpr bool kTag = LC_TAG_MERGE != 0 && KC >= 8;
  constexpr long long TM = TagKey<KC>::kMask;
  // buffered candidates: one 16-byte record (d2 bits, index) per slot and
  // lane, so an append is one address and one ds_write_b128; or, tagged,
  // the d2 bits and the index in the lane's slot entries
  static_assert(W == 4 && BC == 8, "RowBufs shape");
  __shared__ RowBufs<KC, kTag> s_rb;
  // tile list entries: (first slot, wrap code | members << 5), one 8-byte
  // record so a fetch is one LDS read
  __shared__ int2 s_list[W][4][LC];
  // per list entry: the lower bound of its tile's distance to the row's
  // sources, rounded down to float, in the high bits, the entry's index in
  // the low six (nearest-first draining)
  __shared__ unsigned s_key[W][4][LC];
  __shared__ double s_hull[W][4][6];
  __shared__ double s_shift[W][27][4];  // lattice shift per wrap code (x, y, z, pad)
  const int wv = uni(static_cast<int>(threadIdx.x) / kWave);
  const int lane = threadIdx.x & (kWave - 1);
  const int li = lane & 15;
  const int rw = lane >> 4;
  const int rbase = lane & 48;
#pragma unroll 1
  do {
    int wgq = 0;
    if constexpr (kPersist) {
      int t = 0;
      if (lane == 0) t = atomicAdd(A.queue, 1);
      wgq = uni(t);
      if (wgq >= A.n_waves) break;
    } else {
      wgq = uni(static_cast<int>(blockIdx.x) * W + wv);
    }
    const int wg = wgq;
    // the frame whose waves contain wg
    int f = 0;
    {
      int lo = 0, hi = A.F - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (row_wave0(A.tile_off, A.fcb, mid) <= wg) lo = mid;
        else hi = mid - 1;
      }
      f = uni(lo);
    }
    const FrameDesc& d = A.fr[f];
    if (d.kc != KC) continue;
    const int tb = A.tile_off[A.fcb[f]];
    const int nt_frame = A.tile_off[A.fcb[f + 1]] - tb;
    const int tf = (wg - row_wave0(A.tile_off, A.fcb, f)) * 4 + rw;
    if (__ballot(tf < nt_frame) == 0) continue;
    const bool has_tile = tf < nt_frame;
    int T = tb + (has_tile ? tf : 0);
    if (A.tile_perm != nullptr && has_tile) T = A.tile_perm[T];
    const int g = A.tile_cell[T];
    const int o0 = A.off[g];
    const int cnt = A.off[g + 1] - o0;
    const int tl = T - A.tile_off[g];
    const int srel = tl * kTile + li;
    bool valid = has_tile && srel < cnt;
    const int s = o0 + (valid ? srel : 0);
    const bool frame_ok = (A.fflags[f] & kFrameNonFinite) == 0 && d.r_safe >= 1;
    if (!frame_ok) {
      if (valid) defer_source(s, A.defer_list, A.defer_count);
      continue;
    }
    if (kTag && !tag_frame_ok(d, TagKey<KC>::kMaxN)) {
      // indices or distances beyond what the tagged keys hold (frames of
      // 2^27 points and more, 2^26 for the 32-slot heap, or cells of 1e150
      // and more): the exact-order walk takes them
      if (valid) defer_source(s, A.defer_list, A.defer_count);
      continue;
    }
    const int k = d.k;
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const int cb = static_cast<int>(d.cell_base);
    const int c = g - cb;
    const int hx = c % nx;
    const int hy = (c / nx) % ny;
    const int hz = c / (nx * ny);
    const Cand mc = A.cand[s];
    const double px = mc.x, py = mc.y, pz = mc.z;
    const int me = static_cast<int>(mc.id);
    const double inf = as_f64(kInfBits);
    // the listing's FP32 copies (see LC_LIST32)
    const float px32 = coord32(px), py32 = coord32(py), pz32 = coord32(pz);
    const float eps32 = list_eps32(d);
    // bounding box of the row's sources, kept in LDS (read only while
    // listing tiles)
    {
      const double v0 = row_min16(valid ? px : inf);
      const double v1 = row_max16(valid ? px : -inf);
      const double v2 = row_min16(valid ? py : inf);
      const double v3 = row_max16(valid ? py : -inf);
      const double v4 = row_min16(valid ? pz : inf);
      const double v5 = row_max16(valid ? pz : -inf);
      if (li == 0) {
        double* hb = s_hull[wv][rw];
        hb[0] = v0;
        hb[1] = v1;
        hb[2] = v2;
        hb[3] = v3;
        hb[4] = v4;
        hb[5] = v5;
      }
    }
    const double* hull = s_hull[wv][rw];

    // lattice shift of a stencil cell from its wrap code (na+1)*9+(nb+1)*3+nc+1
    // The 27 lattice shifts of the frame, from the wrap code
    // (na+1)*9+(nb+1)*3+nc+1: computed once per wave into LDS (lane = code)
    // by the expressions of Cell::lattice_shift (src/cell.rs:227-247), so a
    // lookup gives the same bits without re-reading the frame's cell.
    {
      if (lane < 27) {
        const int code = lane;
        const int na = code / 9 - 1, nb = (code / 3) % 3 - 1, nc = code % 3 - 1;
        double x, y, z;
        if (code == 13) {
          x = y = z = 0.0;
        } else if (d.ortho) {
          x = static_cast<double>(na) * d.widths[0];
          y = static_cast<double>(nb) * d.widths[1];
          z = static_cast<double>(nc) * d.widths[2];
        } else {
          const double fa = na, fb = nb, fc = nc;
          x = fa * d.h[0] + fb * d.h[3] + fc * d.h[6];
          y = fa * d.h[1] + fb * d.h[4] + fc * d.h[7];
          z = fa * d.h[2] + fb * d.h[5] + fc * d.h[8];
        }
        s_shift[wv][code][0] = x;
        s_shift[wv][code][1] = y;
        s_shift[wv][code][2] = z;
      }
      __builtin_amdgcn_wave_barrier();
    }
    auto shift_of = [&](int code, double& x, double& y, double& z) LC_INLINE {
      x = s_shift[wv][code][0];
      y = s_shift[wv][code][1];
      z = s_shift[wv][code][2];
    };

    Top<KC> h;
    if constexpr (kTag) tag_init<KC>(h, k); else top_init<KC>(h, k);
    double min_rej = inf;
    // threshold cap from the predicted k-th distance of the home cell (+inf
    // without a prediction). Candidates beyond it are never kept; the shell-
    // end check below proves the kept set exact or defers the source.
    double thr_cap = inf;
    if (A.pred != nullptr) {
      const double q = static_cast<double>(A.pred[g]);
      thr_cap = q * q;
    }
    if (A.pred_tile != nullptr && T < A.pred_tile_n) {
      const double q = static_cast<double>(A.pred_tile[T]);
      thr_cap = fmin(thr_cap, q * q);
    }
    if (A.pred_src != nullptr && s < A.pred_src_n) {
      const double q = static_cast<double>(A.pred_src[s]);
      thr_cap = fmin(thr_cap, q * q);
    }
    // predicted depth of the row (1 without a depth schedule, or where the
    // per-source reach is asked for: the reach is the reference's stop
    // shell, which only a shell-by-shell walk observes)
    int rp = 1;
    if (kDepth && A.pred_tile != nullptr && T < A.pred_tile_n && d.out_reach == nullptr) {
      const double q = ceil(static_cast<double>(A.pred_tile[T]) / d.cell_min);
      rp = q >= static_cast<double>(d.r_safe) ? d.r_safe : (q > 1.0 ? static_cast<int>(q) : 1);
    }
    bool live = valid;
    double thr = live ? thr_cap : -inf;  // -inf: finished (never accepts)
    bool cap_fail = false;
    double rthr = LC_ROW_MAX_THR(live ? thr : -inf);  // -inf: no live source
    int nbuf = 0;
    int done_reach = 0;
    unsigned st_tests = 0, st_tiles = 0, st_merges = 0, st_rounds = 0,
             st_cells = 0, st_exact = 0, st_useful = 0;
    unsigned long long cyc_fill = 0, cyc_drain = 0, cyc_merge = 0;
    unsigned st_flush = 0;
    unsigned st_rounds2 = 0, st_useful2 = 0, st_wave2 = 0;
    unsigned st_dead = 0, st_acc = 0, st_deadh = 0;
    unsigned st_idle_done = 0, st_idle_empty = 0;
    bool tile_acc = false, tile_acc_lo = false, tile_acc_hi = false;
    unsigned long long cyc_r2 = 0, cyc_open = 0, cyc_init = 0;
    int cur_r = 1;
    const unsigned long long cyc0 = kStats ? __builtin_readcyclecounter() : 0;

    auto flush = [&]() LC_INLINE {
      const unsigned long long t0 = kStats ? __builtin_readcyclecounter() : 0;
      if (kStats) st_flush += 1;
      if (nbuf > 0) {
        if (kStats) st_merges += 1;
        if constexpr (kTag) {
          // buffered keys tagged KC + t, their entries packed; pads carry
          // the empty key
          long long b[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) {
            const bool v = t < nbuf;
            const unsigned elo = s_rb.dw[wv][KC + 8 + t][lane];
            const unsigned ehi = s_rb.dw[wv][KC + 16 + t][lane];
            const long long e = (static_cast<long long>(ehi) << 32) | elo;
            const unsigned jx = s_rb.dw[wv][KC + t][lane];
            // (a pad's entry is packed too: no key names it)
            s_rb.dw[wv][KC + t][lane] =
                (jx << TagKey<KC>::kBits) | (elo & static_cast<unsigned>(TM));
            b[t] = (v ? (e & ~TM) : TagKey<KC>::kEmpty) | (KC + t);
          }
          merge8_keys<KC>(h.d, b, min_rej, [&](int from, int to) LC_INLINE {
            s_rb.dw[wv][to][lane] = s_rb.dw[wv][from][lane];
          });
        } else if constexpr (KC >= 8) {
          // the buffer is read straight from the shared arrays (LDS loads)
          long long b[8];
          int bx[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) {
            const bool v = t < nbuf;
            const int4 e = s_rb.buf[wv][t][lane];
            b[t] = v ? (static_cast<long long>(e.y) << 32) |
                           static_cast<long long>(static_cast<unsigned>(e.x))
                     : kInfBits;
            bx[t] = v ? e.z : kEmptyJ;
          }
          merge8_regs<KC>(h, b, bx, min_rej);
        } else {
          long long b[8];
          int bx[8];
#pragma unroll
          for (int t = 0; t < 8; ++t) {
            const int4 e = s_rb.buf[wv][t][lane];
            b[t] = (static_cast<long long>(e.y) << 32) |
                   static_cast<long long>(static_cast<unsigned>(e.x));
            bx[t] = e.z;
          }
          insert_merge_regs<KC>(h, b, bx, nbuf, min_rej);
        }
        nbuf = 0;
        // tagged: the k-th key with its tag bits set bounds the k-th d2
        // from above (the accept test then takes a few more candidates of
        // the k-th's own tag bucket, never fewer)
        const long long wk = kTag ? (h.d[KC - 1] | TM) : h.d[KC - 1];
        thr = live ? LC_THR_MIN(as_f64(wk), thr_cap) : -inf;
      }
      rthr = LC_ROW_MAX_THR(live ? thr : -inf);
      if (kStats) cyc_merge += __builtin_readcyclecounter() - t0;
    };

    double cx = 0.0, cy = 0.0, cz = 0.0;
    int cj = -1;
    // One candidate (lane I of each row) against every source of the row.
    // Empty candidate slots carry NaN coordinates and finished sources a
    // threshold of -inf: neither is ever accepted. The index is broadcast only
    // when some lane of the wave accepts (outside the per-lane branch, where
    // every lane is active for the DPP read). SELF: the candidates are the
    // row's own sources, and the source itself is skipped; no other tile of
    // the walk holds it (its own tile is swept only in the opening).
    auto step = [&](auto I, auto selftag) LC_INLINE {
      constexpr int ii = decltype(I)::value;
      constexpr bool SELF = decltype(selftag)::value;
      const double qx = row_bcast<ii>(cx);
      const double qy = row_bcast<ii>(cy);
      const double qz = row_bcast<ii>(cz);
      const double ddx = qx - px, ddy = qy - py, ddz = qz - pz;
      const double d2 = ddx * ddx + ddy * ddy + ddz * ddz;
      bool acc = d2 <= thr;
      int qj;
      if constexpr (kStepBranch) {
        // the index is broadcast only when some lane of the wave accepts.
        // SELF reads it before the test, with every lane active: a DPP read
        // from a lane that is switched off returns 0
        if (SELF) {
          const int sj = row_bcast<ii>(cj);
          acc = acc && sj != me;
        }
        if (!__ballot(acc)) return;
        qj = row_bcast<ii>(cj);
      } else {
        // no wave-wide test: the 32-slot heap's rows accept on most steps
        qj = row_bcast<ii>(cj);
        if (SELF) acc = acc && qj != me;
      }
      if constexpr (kTag && LC_APPEND_ALL) {
        // every lane writes its next free slot, and only an accepting lane
        // claims it: a sweep pair starts with at most BC - 2 entries in
        // every buffer, so the slot written is at most BC - 1 and stays
        // in the buffer
        const long long db = __double_as_longlong(d2);
        s_rb.dw[wv][KC + 8 + nbuf][lane] = static_cast<unsigned>(db);
        s_rb.dw[wv][KC + 16 + nbuf][lane] = static_cast<unsigned>(db >> 32);
        s_rb.dw[wv][KC + nbuf][lane] = static_cast<unsigned>(qj);
        nbuf += acc ? 1 : 0;
        if (kStats && acc) {
          st_acc += 1;
          tile_acc = true;
          if (ii < 8) tile_acc_lo = true;
          else tile_acc_hi = true;
        }
      } else if (acc) {
        // indexed on the shared arrays themselves, so the stores stay LDS
        // stores whatever the register allocator does with pointers
        const long long db = __double_as_longlong(d2);
        if constexpr (kTag) {
          s_rb.dw[wv][KC + 8 + nbuf][lane] = static_cast<unsigned>(db);
          s_rb.dw[wv][KC + 16 + nbuf][lane] = static_cast<unsigned>(db >> 32);
          s_rb.dw[wv][KC + nbuf][lane] = static_cast<unsigned>(qj);
        } else {
          s_rb.buf[wv][nbuf][lane] =
              make_int4(static_cast<int>(db), static_cast<int>(db >> 32), qj, 0);
        }
        nbuf += 1;
        if (kStats) {
          st_acc += 1;
          tile_acc = true;
          if (ii < 8) tile_acc_lo = true;
          else tile_acc_hi = true;
        }
      }
    };
    auto pair = [&](auto B, auto selftag) LC_INLINE {
      constexpr int b0 = decltype(B)::value;
      step(std::integral_constant<int, b0>{}, selftag);
      step(std::integral_constant<int, b0 + 1>{}, selftag);
    };
    // the 16 candidates held in cx..cj (one per lane of each row) against
    // every source of the row; one merge site, checked after every two
    // candidates, so buffers fill to BC - 1 or BC before a merge
    // The switch is entered at pair pr and falls through the following
    // pairs until a buffer is nearly full; only then does control leave it
    // for the (single) merge site and come back at the next pair.
    auto sweep = [&](auto selftag) LC_INLINE {
      if (kStats) st_rounds += 1;
      int pr = 0;
#pragma unroll 1
      while (pr < 8) {
        switch (pr) {
#define LC_PAIR(P)                                              \
    case P:                                                       \
      pair(std::integral_constant<int, 2 * (P)>{}, selftag);      \
      pr = (P) + 1;                                               \
      if (__ballot(nbuf > BC - 2)) break;                         \
      [[fallthrough]];
          LC_PAIR(0)
          LC_PAIR(1)
          LC_PAIR(2)
          LC_PAIR(3)
          LC_PAIR(4)
          LC_PAIR(5)
          LC_PAIR(6)
          LC_PAIR(7)
#undef LC_PAIR
          default: break;
        }
        if (__ballot(nbuf > BC - 2)) flush();
      }
    };

    const int nt_home = (cnt + kTile - 1) >> kTileShift;
    const int tl_nb = tl + 1 < nt_home ? tl + 1 : tl - 1;  // -1: none
    // a third opening tile, the nearest other neighbour of the row's own
    // tile in sub-cell order (tl - 1 when tl + 1 opened, else tl - 2): FCC
    // 8M frame (k = 32) walk 9.73 -> 9.47 ms; clustered 4M frame unchanged
    // (4.040 against 4.042 ms)
#ifndef LC_OPEN3
#define LC_OPEN3 1
#endif
    const int tl_nb2 = !LC_OPEN3 ? -1
                       : (tl_nb == tl + 1 ? (tl >= 1 ? tl - 1 : (tl + 2 < nt_home ? tl + 2 : -1))
                                          : (tl >= 2 ? tl - 2 : -1));
    const double nan = __builtin_nan("");
    // Walk, at most twice: a row some of whose sources the cap may have kept
    // from the reference's stop (heap not full at a shell end with the cap
    // below the stop bound) is walked again, whole and uncapped, by the same
    // wave once every other row has finished.
#pragma unroll 1
    for (int attempt = 0; attempt < 2; ++attempt) {
    // opening: the row's own tile (its sources are the candidates, no
    // shift), then the next tiles of the home cell in sub-cell order
    const unsigned long long to0 = kStats ? __builtin_readcyclecounter() : 0;
    if (kStats && attempt == 0) cyc_init = to0 - cyc0;
    // the next tile's candidates are loaded before the own tile is swept, so
    // the load overlaps that sweep
    // (loaded unconditionally, from the source's own slot when the lane
    // has no candidate, so that the loads sit in one block and the waits
    // before each sweep count exactly the loads it needs)
    double nbx, nby, nbz;
    int nbj;
    {
      const bool ok = has_tile && tl_nb >= 0 && tl_nb * kTile + li < cnt;
      const Cand q = A.cand[ok ? o0 + tl_nb * kTile + li : s];
      nbx = ok ? q.x : nan;
      nby = ok ? q.y : nan;
      nbz = ok ? q.z : nan;
      nbj = ok ? static_cast<int>(q.id) : -1;
    }
    cx = valid ? px : nan;
    cy = valid ? py : nan;
    cz = valid ? pz : nan;
    cj = valid ? me : -1;
    sweep(std::true_type{});
    cx = nbx;
    cy = nby;
    cz = nbz;
    cj = nbj;
    if (LC_OPEN3) {
      // the third tile's loads overlap the second tile's sweep
      const bool ok = has_tile && tl_nb2 >= 0 && tl_nb2 * kTile + li < cnt;
      const Cand q = A.cand[ok ? o0 + tl_nb2 * kTile + li : s];
      nbx = ok ? q.x : nan;
      nby = ok ? q.y : nan;
      nbz = ok ? q.z : nan;
      nbj = ok ? static_cast<int>(q.id) : -1;
    }
    if (__ballot(tl_nb >= 0 && has_tile)) sweep(std::false_type{});
    if (LC_OPEN3) {
      cx = nbx;
      cy = nby;
      cz = nbz;
      cj = nbj;
      if (__ballot(tl_nb2 >= 0 && has_tile)) sweep(std::false_type{});
    }
    if (__ballot(nbuf > 0)) flush();
    if (kStats) cyc_open += __builtin_readcyclecounter() - to0;

    int lcount = 0;  // entries in this row's list
    // drain the list: tile i's candidates are loaded while tile i - 1 is
    // swept
    auto drain = [&]() LC_INLINE {
      const unsigned long long t0 = kStats ? __builtin_readcyclecounter() : 0;
      double nxq = nan, nyq = nan, nzq = nan;
      int njq = -1, ncode = 13;
      // half tiles: lanes 0-7 load entry e, lanes 8-15 entry e2
      auto fetch = [&](int e, int e2) LC_INLINE {
        nxq = nyq = nzq = nan;
        njq = -1;
        ncode = 13;
        const int el = (kHalf && li >= 8) ? e2 : e;
        if (el >= 0) {
          const int2 ent = s_list[wv][rw][el];
          const int sl0 = ent.x;
          const int meta = ent.y;
          ncode = meta & 31;
          const int m = kHalf ? (li & 7) : li;
          if (m < (meta >> 5)) {
            const Cand q = A.cand[sl0 + m];
            nxq = q.x;
            nyq = q.y;
            nzq = q.z;
            njq = static_cast<int>(q.id);
          }
        }
      };
#if LC_ROW_SELECT && LC_KEYS_REG
      // nearest first, keys in registers: lane li holds the keys of entries
      // li, li + 16, ...; taken entries become ~0
      unsigned kr[NK];
#pragma unroll
      for (int j = 0; j < NK; ++j) {
        const int e = li + 16 * j;
        kr[j] = e < lcount ? s_key[wv][rw][e] : ~0u;
      }
      auto take = [&](unsigned key) LC_INLINE {
        const int e = static_cast<int>(key & 63u);
#pragma unroll
        for (int j = 0; j < NK; ++j) kr[j] = (li == (e & 15) && j == (e >> 4)) ? ~0u : kr[j];
      };
      auto beyond = [&](unsigned key) LC_INLINE {
        return static_cast<double>(__uint_as_float(key & ~63u)) > rthr;
      };
      // the row's smallest key (and, for half tiles, the second smallest):
      // -1 once the list is empty or the nearest bound is beyond the row's
      // threshold (every later entry is then beyond it too)
      int e2n = -1;
      auto next = [&](int) LC_INLINE {
        unsigned m1 = ~0u, m2 = ~0u;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
          const unsigned v = kr[j];
          if (kHalf) m2 = min(m2, max(m1, v));
          m1 = min(m1, v);
        }
        auto pair_step = [&](auto ctrl) LC_INLINE {
          constexpr int c = decltype(ctrl)::value;
          const unsigned b1 = static_cast<unsigned>(
              __builtin_amdgcn_update_dpp(0, static_cast<int>(m1), c, 0xf, 0xf, true));
          if (kHalf) {
            const unsigned b2 = static_cast<unsigned>(
                __builtin_amdgcn_update_dpp(0, static_cast<int>(m2), c, 0xf, 0xf, true));
            m2 = min(min(m2, b2), max(m1, b1));
          }
          m1 = min(m1, b1);
        };
        pair_step(std::integral_constant<int, 0xB1>{});
        pair_step(std::integral_constant<int, 0x4E>{});
        pair_step(std::integral_constant<int, 0x141>{});
        pair_step(std::integral_constant<int, 0x140>{});
        e2n = -1;
        if (m1 == ~0u || beyond(m1)) return -1;
        take(m1);
        if (kHalf && m2 != ~0u && !beyond(m2)) {
          take(m2);
          e2n = static_cast<int>(m2 & 63u);
        }
        return static_cast<int>(m1 & 63u);
      };
#elif LC_ROW_SELECT
      // nearest first: the entry of the row's smallest key, taken out of the
      // list; -1 once the list is empty or its nearest bound is beyond the
      // row's threshold (every later entry is then beyond it too)
      auto next = [&](int) LC_INLINE {
        unsigned kmin = ~0u;
#pragma unroll
        for (int j = 0; j < NK; ++j) {
          const int e = li + 16 * j;
          if (e < lcount) kmin = min(kmin, s_key[wv][rw][e]);
        }
        kmin = row_min16_u32(kmin);
        if (kmin == ~0u) return -1;
        if (static_cast<double>(__uint_as_float(kmin & ~63u)) > rthr) return -1;
        const int e = static_cast<int>(kmin & 63u);
        if (li == (e & 15)) s_key[wv][rw][e] = ~0u;
        return e;
      };
#else
      // in list order
      auto next = [&](int i) LC_INLINE { return i < lcount ? i : -1; };
#endif
      int i = 0;
      int e = next(i);
#if LC_ROW_SELECT && LC_KEYS_REG
      int e2 = e2n;
#else
      int e2 = (kHalf && e >= 0) ? next(i) : -1;
#endif
      fetch(e, e2);
#pragma unroll 1
      while (__ballot(e >= 0)) {
        double shx, shy, shz;
        shift_of(ncode, shx, shy, shz);
        cx = nxq + shx;
        cy = nyq + shy;
        cz = nzq + shz;
        cj = njq;
        if (kStats) st_useful += e >= 0 ? 1 : 0;
        if (kStats && e < 0) {
          // an idle row: finished, or live with an empty list this drain
          // (the rest ran out of listed tiles before the wave's last round)
          if (!(rthr >= 0.0)) st_idle_done += 1;
          else if (lcount == 0) st_idle_empty += 1;
        }
        if (kStats && cur_r >= 2) {
          st_rounds2 += 1;
          st_useful2 += e >= 0 ? 1 : 0;
        }
        const bool swept = e >= 0;
        i += 1;
        e = swept ? next(i) : -1;
#if LC_ROW_SELECT && LC_KEYS_REG
        e2 = swept ? e2n : -1;
#else
        if (kHalf) e2 = e >= 0 ? next(i) : -1;
#endif
        fetch(e, e2);
        if (kStats) st_tiles += 1;
        if (kStats) tile_acc = tile_acc_lo = tile_acc_hi = false;
        sweep(std::false_type{});
        if (kStats) {
          const bool row_any = ((__ballot(tile_acc) >> rbase) & 0xffffull) != 0;
          const bool row_lo = ((__ballot(tile_acc_lo) >> rbase) & 0xffffull) != 0;
          const bool row_hi = ((__ballot(tile_acc_hi) >> rbase) & 0xffffull) != 0;
          st_dead += (swept && !row_any) ? 1 : 0;
          st_deadh += swept ? (!row_lo ? 1 : 0) + (!row_hi ? 1 : 0) : 0;
        }
      }
      lcount = 0;
      if (kStats) cyc_drain += __builtin_readcyclecounter() - t0;
    };

    // the row's shell (rows of one wave may be in different shells under a
    // depth schedule)
    int r = 1;
#pragma unroll 1
    for (;; ++r) {
      int ncr = shell_cells(r);
      cur_r = r;
      if (kStats && !st_wave2 && __ballot(r >= 2) && attempt == 0) {
        st_wave2 = 1;
        cyc_r2 = __builtin_readcyclecounter();
      }
      // per-row enumeration state: a chunk of 16 cells, one per lane
      int cc = 0, p0 = 0, total = 0;
      int ca0 = 0, cn2 = 0, ct0 = 0, ccode = 13, cnt_c = 0, cpre = 0;
      bool more = rthr >= 0.0;
#pragma unroll 1
      while (true) {
        // fill the list while every filling row has room for 16 entries
        const unsigned long long tf0 = kStats ? __builtin_readcyclecounter() : 0;
#pragma unroll 1
        while (true) {
          const bool fill = more && lcount < kFill;
          if (!__ballot(fill)) break;
          if (fill) {
            if (p0 >= total) {
              // below the predicted depth the next shell's cells follow in
              // the same pass, with no stop test between
              if (kDepth && cc >= ncr && r < rp) {
                r += 1;
                ncr = shell_cells(r);
                cc = 0;
              }
              if (cc >= ncr) {
                more = false;
              } else {
                // headers of 16 cells, one per lane
                if (kStats) st_cells += 1;
                const int ci = cc + li;
                cn2 = 0;
                ca0 = ct0 = 0;
                ccode = 13;
                if (ci < ncr) {
                  int dx, dy, dz;
                  surface_cell(r, ci, dx, dy, dz);
                  const int jx = hx + dx, jy = hy + dy, jz = hz + dz;
                  const int ux = jx < 0 ? jx + nx : (jx >= nx ? jx - nx : jx);
                  const int uy = jy < 0 ? jy + ny : (jy >= ny ? jy - ny : jy);
                  const int uz = jz < 0 ? jz + nz : (jz >= nz ? jz - nz : jz);
                  const int na = jx < 0 ? -1 : (jx >= nx ? 1 : 0);
                  const int nb = jy < 0 ? -1 : (jy >= ny ? 1 : 0);
                  const int nc = jz < 0 ? -1 : (jz >= nz ? 1 : 0);
                  const int gq = cb + (uz * ny + uy) * nx + ux;
                  ca0 = A.off[gq];
                  cn2 = A.off[gq + 1] - ca0;
                  ct0 = A.tile_off[gq];
                  ccode = (na + 1) * 9 + (nb + 1) * 3 + (nc + 1);
                  if (LC_CELL_BOX_TEST && cn2 > kTile) {
                    double shx, shy, shz;
                    shift_of(ccode, shx, shy, shz);
                    const double sl[3] = {hull[0], hull[2], hull[4]};
                    const double sh[3] = {hull[1], hull[3], hull[5]};
                    if (hull_lb(sl, sh, A.cbox[gq], shx, shy, shz) > rthr) cn2 = 0;
                  }
                }
                cnt_c = (cn2 + kTile - 1) >> kTileShift;
                // inclusive prefix of tile counts over the row: DPP row_shr
                // by 1, 2, 4 and 8 (lanes shifted in from outside the row
                // read 0), then lane 15's sum broadcast to the row
                cpre = cnt_c;
                cpre += __builtin_amdgcn_update_dpp(0, cpre, 0x111, 0xf, 0xf, true);
                cpre += __builtin_amdgcn_update_dpp(0, cpre, 0x112, 0xf, 0xf, true);
                cpre += __builtin_amdgcn_update_dpp(0, cpre, 0x114, 0xf, 0xf, true);
                cpre += __builtin_amdgcn_update_dpp(0, cpre, 0x118, 0xf, 0xf, true);
                total = row_bcast<15>(cpre);
                p0 = 0;
                cc += 16;
              }
            }
            if (more && p0 < total) {
              // 16 tiles of the chunk, one per lane
              const int p = p0 + li;
              const bool pv = p < total;
              int own = 0;
#pragma unroll
              for (int step2 = 8; step2 >= 1; step2 >>= 1) {
                const int probe = own + step2 - 1;
                const int pr = __shfl(cpre, rbase + (probe < 16 ? probe : 15));
                if (probe < 16 && pr <= p) own += step2;
              }
              own = own < 16 ? own : 15;
              const int oa0 = __shfl(ca0, rbase + own);
              const int on2 = __shfl(cn2, rbase + own);
              const int ot0 = __shfl(ct0, rbase + own);
              const int ocode = __shfl(ccode, rbase + own);
              const int oex = __shfl(cpre - cnt_c, rbase + own);
              // the home cell's tiles go outward from the row's own tile in
              // sub-cell order (tl, tl + 1, tl - 1, ...): the nearest first,
              // so the thresholds tighten before the far tiles are tested
              const bool home = r == 1 && cc == 16 && own == 0;
              const int t = (LC_HOME_SPIRAL && home) ? spiral_tile(p - oex, tl, nt_home)
                                                     : p - oex;
              const bool opening = home && (t == tl || t == tl_nb || t == tl_nb2);
              bool need = false;
              double lbt = inf;  // lower bound of the tile's distance to the row's sources
              double bx0 = 0, bx1 = 0, by0 = 0, by1 = 0, bz0 = 0, bz1 = 0;
              if (pv && !opening) {
                double shx, shy, shz;
                shift_of(ocode, shx, shy, shz);
                const Box b = A.tbox[ot0 + t];
                bx0 = b.lx + shx;
                bx1 = b.hx + shx;
                by0 = b.ly + shy;
                by1 = b.hy + shy;
                bz0 = b.lz + shz;
                bz1 = b.hz + shz;
                if (kStats) st_tests += 1;
                const double gx = fmax(fmax(bx0 - hull[1], hull[0] - bx1), 0.0);
                const double gy = fmax(fmax(by0 - hull[3], hull[2] - by1), 0.0);
                const double gz = fmax(fmax(bz0 - hull[5], hull[4] - bz1), 0.0);
                lbt = gx * gx + gy * gy + gz * gz;
                need = lbt <= rthr;
              }
              if constexpr (kHalf) {
                // exact per-source test of each half of the tiles that passed
                // the row box, with the half tiles' own boxes
                bool need_a = false, need_b = false;
                double lb_a = inf, lb_b = inf;
                const int e = on2 - t * kTile;  // members of the tile
                if (__ballot(need)) {
                  if (kStats) st_exact += 1;
                  double ax0 = 0, ax1 = 0, ay0 = 0, ay1 = 0, az0 = 0, az1 = 0;
                  double cx0 = 0, cx1 = 0, cy0 = 0, cy1 = 0, cz0 = 0, cz1 = 0;
                  if (need) {
                    double shx, shy, shz;
                    shift_of(ocode, shx, shy, shz);
                    const Box ha = A.hbox[2 * (ot0 + t)];
                    const Box hb = A.hbox[2 * (ot0 + t) + 1];
                    ax0 = ha.lx + shx;
                    ax1 = ha.hx + shx;
                    ay0 = ha.ly + shy;
                    ay1 = ha.hy + shy;
                    az0 = ha.lz + shz;
                    az1 = ha.hz + shz;
                    cx0 = hb.lx + shx;
                    cx1 = hb.hx + shx;
                    cy0 = hb.ly + shy;
                    cy1 = hb.hy + shy;
                    cz0 = hb.lz + shz;
                    cz1 = hb.hz + shz;
                  }
                  const double tq = live ? thr : -inf;
                  bool any_a = false, any_b = false;
#if LC_LIST32
                  const float tq32 = thr_up32(tq);
                  const float fa[6] = {coord32(ax0) - eps32, coord32(ax1) + eps32,
                                       coord32(ay0) - eps32, coord32(ay1) + eps32,
                                       coord32(az0) - eps32, coord32(az1) + eps32};
                  const float fb[6] = {coord32(cx0) - eps32, coord32(cx1) + eps32,
                                       coord32(cy0) - eps32, coord32(cy1) + eps32,
                                       coord32(cz0) - eps32, coord32(cz1) + eps32};
                  float lb_a32 = __builtin_inff(), lb_b32 = __builtin_inff();
                  auto probe_src = [&](auto I) LC_INLINE {
                    constexpr int ii = decltype(I)::value;
                    const float qx = row_bcast_f32<ii>(px32);
                    const float qy = row_bcast_f32<ii>(py32);
                    const float qz = row_bcast_f32<ii>(pz32);
                    const float qt = row_bcast_f32<ii>(tq32);
                    {
                      const float gx = fmaxf(fmaxf(fa[0] - qx, qx - fa[1]), 0.0f);
                      const float gy = fmaxf(fmaxf(fa[2] - qy, qy - fa[3]), 0.0f);
                      const float gz = fmaxf(fmaxf(fa[4] - qz, qz - fa[5]), 0.0f);
                      const float ls = __builtin_fmaf(gz, gz, __builtin_fmaf(gy, gy, gx * gx));
                      const bool ok = qt >= ls;
                      any_a |= ok;
                      lb_a32 = fminf(lb_a32, ok ? ls : __builtin_inff());
                    }
                    {
                      const float gx = fmaxf(fmaxf(fb[0] - qx, qx - fb[1]), 0.0f);
                      const float gy = fmaxf(fmaxf(fb[2] - qy, qy - fb[3]), 0.0f);
                      const float gz = fmaxf(fmaxf(fb[4] - qz, qz - fb[5]), 0.0f);
                      const float ls = __builtin_fmaf(gz, gz, __builtin_fmaf(gy, gy, gx * gx));
                      const bool ok = qt >= ls;
                      any_b |= ok;
                      lb_b32 = fminf(lb_b32, ok ? ls : __builtin_inff());
                    }
                  };
#else
                  auto probe_src = [&](auto I) LC_INLINE {
                    constexpr int ii = decltype(I)::value;
                    const double qx = row_bcast<ii>(px);
                    const double qy = row_bcast<ii>(py);
                    const double qz = row_bcast<ii>(pz);
                    const double qt = row_bcast<ii>(tq);
                    {
                      const double gx = fmax(fmax(ax0 - qx, qx - ax1), 0.0);
                      const double gy = fmax(fmax(ay0 - qy, qy - ay1), 0.0);
                      const double gz = fmax(fmax(az0 - qz, qz - az1), 0.0);
                      const double ls = gx * gx + gy * gy + gz * gz;
                      any_a |= ls <= qt;
                      lb_a = key_min(lb_a, ls <= qt ? ls : inf);
                    }
                    {
                      const double gx = fmax(fmax(cx0 - qx, qx - cx1), 0.0);
                      const double gy = fmax(fmax(cy0 - qy, qy - cy1), 0.0);
                      const double gz = fmax(fmax(cz0 - qz, qz - cz1), 0.0);
                      const double ls = gx * gx + gy * gy + gz * gz;
                      any_b |= ls <= qt;
                      lb_b = key_min(lb_b, ls <= qt ? ls : inf);
                    }
                  };
#endif
                  probe_src(std::integral_constant<int, 0>{});
                  probe_src(std::integral_constant<int, 1>{});
                  probe_src(std::integral_constant<int, 2>{});
                  probe_src(std::integral_constant<int, 3>{});
                  probe_src(std::integral_constant<int, 4>{});
                  probe_src(std::integral_constant<int, 5>{});
                  probe_src(std::integral_constant<int, 6>{});
                  probe_src(std::integral_constant<int, 7>{});
                  probe_src(std::integral_constant<int, 8>{});
                  probe_src(std::integral_constant<int, 9>{});
                  probe_src(std::integral_constant<int, 10>{});
                  probe_src(std::integral_constant<int, 11>{});
                  probe_src(std::integral_constant<int, 12>{});
                  probe_src(std::integral_constant<int, 13>{});
                  probe_src(std::integral_constant<int, 14>{});
                  probe_src(std::integral_constant<int, 15>{});
#if LC_LIST32
                  lb_a = key_of32(lb_a32);
                  lb_b = key_of32(lb_b32);
#endif
                  // a tile of 8 or fewer has an empty second half, whose
                  // (+inf, -inf) box passes an infinite threshold
                  need_a = need && any_a;
                  need_b = need && any_b && e > 8;
                }
                const unsigned rma = static_cast<unsigned>(__ballot(need_a) >> rbase) & 0xffffu;
                const unsigned rmb = static_cast<unsigned>(__ballot(need_b) >> rbase) & 0xffffu;
                const unsigned below = (1u << li) - 1u;
                if (need_a) {
                  const int pos = lcount + __popc(rma & below);
                  s_list[wv][rw][pos] = make_int2(oa0 + t * kTile, ocode | ((e < 8 ? e : 8) << 5));
                  s_key[wv][rw][pos] = (__float_as_uint(__double2float_rd(lb_a)) & ~63u) |
                                       static_cast<unsigned>(pos);
                }
                if (need_b) {
                  const int pos = lcount + __popc(rma) + __popc(rmb & below);
                  s_list[wv][rw][pos] = make_int2(oa0 + t * kTile + 8, ocode | ((e - 8) << 5));
                  s_key[wv][rw][pos] = (__float_as_uint(__double2float_rd(lb_b)) & ~63u) |
                                       static_cast<unsigned>(pos);
                }
                lcount += __popc(rma) + __popc(rmb);
              } else {
              // exact per-source test of the tiles that passed the row box
                if (__ballot(need)) {
                  if (kStats) st_exact += 1;
                  const double tq = live ? thr : -inf;
                  bool any = false;
                  double lbs = inf;  // smallest per-source bound (a tighter key)
#if LC_LIST32
                  const float tq32 = thr_up32(tq);
                  const float fx0 = coord32(bx0) - eps32, fx1 = coord32(bx1) + eps32;
                  const float fy0 = coord32(by0) - eps32, fy1 = coord32(by1) + eps32;
                  const float fz0 = coord32(bz0) - eps32, fz1 = coord32(bz1) + eps32;
                  float lbs32 = __builtin_inff();
                  auto probe_src = [&](auto I) LC_INLINE {
                    constexpr int ii = decltype(I)::value;
                    const float qx = row_bcast_f32<ii>(px32);
                    const float qy = row_bcast_f32<ii>(py32);
                    const float qz = row_bcast_f32<ii>(pz32);
                    const float gx = fmaxf(fmaxf(fx0 - qx, qx - fx1), 0.0f);
                    const float gy = fmaxf(fmaxf(fy0 - qy, qy - fy1), 0.0f);
                    const float gz = fmaxf(fmaxf(fz0 - qz, qz - fz1), 0.0f);
                    const float ls = __builtin_fmaf(gz, gz, __builtin_fmaf(gy, gy, gx * gx));
                    const bool ok = row_bcast_f32<ii>(tq32) >= ls;
                    any |= ok;
                    // only sources that can still take the tile bound its key
                    if (LC_ROW_SELECT) lbs32 = fminf(lbs32, ok ? ls : __builtin_inff());
                  };
#else
                  auto probe_src = [&](auto I) LC_INLINE {
                    constexpr int ii = decltype(I)::value;
                    const double qx = row_bcast<ii>(px);
                    const double qy = row_bcast<ii>(py);
                    const double qz = row_bcast<ii>(pz);
                    const double qt = row_bcast<ii>(tq);
                    const double gx = fmax(fmax(bx0 - qx, qx - bx1), 0.0);
                    const double gy = fmax(fmax(by0 - qy, qy - by1), 0.0);
                    const double gz = fmax(fmax(bz0 - qz, qz - bz1), 0.0);
                    const double ls = gx * gx + gy * gy + gz * gz;
                    any |= ls <= qt;
                    // only sources that can still take the tile bound its key
                    if (LC_ROW_SELECT) lbs = key_min(lbs, ls <= qt ? ls : inf);
                  };
#endif
                  probe_src(std::integral_constant<int, 0>{});
                  probe_src(std::integral_constant<int, 1>{});
                  probe_src(std::integral_constant<int, 2>{});
                  probe_src(std::integral_constant<int, 3>{});
                  probe_src(std::integral_constant<int, 4>{});
                  probe_src(std::integral_constant<int, 5>{});
                  probe_src(std::integral_constant<int, 6>{});
                  probe_src(std::integral_constant<int, 7>{});
                  probe_src(std::integral_constant<int, 8>{});
                  probe_src(std::integral_constant<int, 9>{});
                  probe_src(std::integral_constant<int, 10>{});
                  probe_src(std::integral_constant<int, 11>{});
                  probe_src(std::integral_constant<int, 12>{});
                  probe_src(std::integral_constant<int, 13>{});
                  probe_src(std::integral_constant<int, 14>{});
                  probe_src(std::integral_constant<int, 15>{});
#if LC_LIST32
                  lbs = key_of32(lbs32);
#endif
                  need = need && any;
                  if (LC_ROW_SELECT) lbt = key_max(lbt, lbs);
                }
                const unsigned long long bal = __ballot(need);
                const unsigned rm = static_cast<unsigned>(bal >> rbase) & 0xffffu;
                if (need) {
                  const int pos = lcount + __popc(rm & ((1u << li) - 1u));
                  const int e = on2 - t * kTile;
                  s_list[wv][rw][pos] = make_int2(oa0 + t * kTile, ocode | ((e < kTile ? e : kTile) << 5));
                  if (LC_ROW_SELECT) {
                    s_key[wv][rw][pos] =
                        (__float_as_uint(__double2float_rd(lbt)) & ~63u) | static_cast<unsigned>(pos);
                  }
                }
    lcount += __popc(rm);
              }
              p0 += 16;
            }
          }
        }
        if (kStats) cyc_fill += __builtin_readcyclecounter() - tf0;
        if (__ballot(lcount > 0)) drain();
        if (!__ballot(more)) break;
      }
      if (__ballot(nbuf > 0)) flush();
      // stop test, src/knearest.rs:466-471
      const double bound = static_cast<double>(r) * d.cell_min;
      // the k-th d2 exactly (tagged: its key with the d2's own low bits)
      // and whether the heap is full
      long long kth = h.d[KC - 1];
      bool full = false;
      if constexpr (kTag) {
        full = (kth & ~TM) != TagKey<KC>::kEmpty;
        kth = (kth & ~TM) |
              (s_rb.dw[wv][static_cast<int>(kth & TM)][lane] & static_cast<unsigned>(TM));
      } else {
        full = h.j[KC - 1] != kEmptyJ;
      }
      if (live && full) {
        // a full heap holds only d2 <= thr_cap, and everything skipped or
        // rejected lies beyond it: the heap is the exact top-k of the shells
        // walked so far
        if (as_f64(kth) <= bound * bound) {
          live = false;
          done_reach = r;
        }
      } else if (live && thr_cap < bound * bound) {
        // not full: the plain walk might have filled its heap with candidates
        // beyond the cap and stopped here; only an uncapped walk can tell
        live = false;
        cap_fail = true;
      }
      // beyond r_safe the stencil wraps onto itself: exact walk
      if (live && r + 1 > d.r_safe) live = false;
      if (!live) thr = -inf;
      rthr = LC_ROW_MAX_THR(live ? thr : -inf);
      if (!__ballot(live)) break;
    }
    const bool row_fail = ((__ballot(cap_fail) >> rbase) & 0xffffull) != 0;
    if (!__ballot(row_fail)) break;
    if (row_fail && li == 0 && A.retry_count != nullptr) atomicAdd(A.retry_count, 1);
    // the failed rows start over without the cap; every other row is done
    if (row_fail) {
      if constexpr (kTag) tag_init<KC>(h, k); else top_init<KC>(h, k);
      min_rej = inf;
      thr_cap = inf;
      rp = 1;
      live = valid;
      thr = live ? inf : -inf;
      done_reach = 0;
      cap_fail = false;
    }
    rthr = LC_ROW_MAX_THR(live ? thr : -inf);
    }
    if (kStats) {
      stat_add(A.stats, kStatRowTileTests, st_tests);
      stat_add(A.stats, kStatRowMerges, st_merges);
      if (valid) stat_add(A.stats, kStatRowSources, 1);
      if (li == 0) {
        stat_add(A.stats, kStatRowTiles, st_tiles);
        stat_add(A.stats, kStatRowRefines, st_exact);
        stat_add(A.stats, kStatRowUseful, st_useful);
      }
      if (lane == 0) {
        stat_add(A.stats, kStatRowRounds, st_rounds);
        stat_add(A.stats, kStatRowCells, st_cells);
        stat_add(A.stats, kStatRowCycFill, cyc_fill);
        stat_add(A.stats, kStatRowCycDrain, cyc_drain);
        stat_add(A.stats, kStatRowCycMerge, cyc_merge);
        stat_add(A.stats, kStatRowCycAll, __builtin_readcyclecounter() - cyc0);
        stat_add(A.stats, kStatRowFlushes, st_flush);
        stat_add(A.stats, kStatRowWavesR2, st_wave2);
        stat_add(A.stats, kStatRowRoundsR2, st_rounds2);
        if (st_wave2) stat_add(A.stats, kStatRowCycR2, __builtin_readcyclecounter() - cyc_r2);
        stat_add(A.stats, kStatRowCycOpen, cyc_open);
        stat_add(A.stats, kStatRowCycInit, cyc_init);
      }
      if (li == 0) stat_add(A.stats, kStatRowUsefulR2, st_useful2);
      if (li == 0) stat_add(A.stats, kStatRowDead, st_dead);
      if (li == 0) stat_add(A.stats, kStatRowDeadHalf, st_deadh);
      if (li == 0) stat_add(A.stats, kStatRowIdleDone, st_idle_done);
      if (li == 0) stat_add(A.stats, kStatRowIdleEmpty, st_idle_empty);
      stat_add(A.stats, kStatRowAccepts, st_acc);
    }
    if (!valid) continue;
    const size_t row = static_cast<size_t>(me) * k;
    bool emitted = false;
    if (done_reach != 0) {
      if constexpr (kTag) {
        unsigned pk[KC];
#pragma unroll
        for (int t = 0; t < KC; ++t) pk[t] = s_rb.dw[wv][static_cast<int>(h.d[t] & TM)][lane];
        emitted = tag_emit<KC>(h, pk, k, min_rej, d.out_nn, d.out_d2, row);
      } else {
        emitted = top_emit<KC>(h, k, min_rej, d.out_nn, d.out_d2, row);
      }
    }
    if (!emitted) {
      defer_source(s, A.defer_list, A.defer_count);
      continue;
    }
    if (d.out_reach) d.out_reach[me] = done_reach;
  } while (kPersist);
}

#ifndef LC_CAP_CELL_MAX
#define LC_CAP_CELL_MAX 16
#endif
// Tiles whose learned cap the walk uses (LC_CAP_CELL_MAX): the one tile of
// each cell of 1 .. cmax members goes to the list, to which inference is
// restricted; every other tile gets no cap (+inf). Caps help where the
// opening cannot fill the heap (home cells of at most 16 points, whose
// sources take every candidate until their first merge) and barely
// elsewhere.
// One atomic per 1024-thread block on the list's count (the list's order
// does not matter): atomics on one address serialize, and one per wave took
// 37 us on the clustered 4M frame's 192,000 cells.
constexpr int kCapThreads = 1024;
__global__ __launch_bounds__(kCapThreads) void k_cap_tiles(
    const int* __restrict__ count, const int* __restrict__ tile_off, long long ncells,
    int cmax, float* __restrict__ caps, int* __restrict__ list, int* __restrict__ list_n) {
  __shared__ int s_off[kCapThreads / kWave];
  __shared__ int s_base;
  const long long g = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
  const bool in = g < ncells;
  const int c = in ? count[g] : 0;
  const int t0 = in ? tile_off[g] : 0, t1 = in ? tile_off[g + 1] : 0;
  const bool take = in && c >= 1 && c <= cmax && t1 - t0 == 1;
  const unsigned long long m = __ballot(take);
  if (lane == 0) s_off[wv] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    int tot = 0;
    for (int w = 0; w < kCapThreads / kWave; ++w) {
      const int v = s_off[w];
      s_off[w] = tot;
      tot += v;
    }
    s_base = tot > 0 ? atomicAdd(list_n, tot) : 0;
  }
  __syncthreads();
  if (take) list[s_base + s_off[wv] + __popcll(m & ((1ull << lane) - 1ull))] = t0;
  if (in && !take) {
    for (int t = t0; t < t1; ++t) caps[t] = __builtin_inff();
  }
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

// Depth schedule of the lane-per-source walk. Every source of a cell below
// the dense threshold, in a frame with a register-heap bucket, goes to
// bucket (heap bucket, clamped predicted depth of its home cell). Pass 1
// counts, pass 2 places; one atomic per wave and bucket, so each wave's
// sources stay in cell order inside their bucket. Only the walk order
// changes: every source still runs the same walk and stop test.
template <bool kPlace>
__global__ __launch_bounds__(256) void k_sched(
    const FrameDesc* __restrict__ fr, const long long* __restrict__ fcb,
    int F, const int* __restrict__ scell, const int* __restrict__ off,
    long long ncells, const float* __restrict__ pred, int dense_min,
    int* __restrict__ count, int* __restrict__ cursor,
    int* __restrict__ sched) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & (kWave - 1);
  const int nact = off[ncells];
  int b = -1;
  if (s < nact) {
    const int g = scell[s];
    if (off[g + 1] - off[g] < dense_min) {
      const FrameDesc& d = fr[find_frame(fcb, F, g)];
      const int kc = d.kc;
      if (kc != 0) {
        // predicted depth: the first shell whose stop bound covers the
        // predicted k-th distance
        const double q = static_cast<double>(pred[g]) / d.cell_min;
        int dep = q < kSchedDepths ? static_cast<int>(ceil(q)) : kSchedDepths;
        dep = dep < 1 ? 1 : (dep > kSchedDepths ? kSchedDepths : dep);
        b = sched_kc_index(kc) * kSchedDepths + (kSchedDepths - dep);
      }
    }
  }
  const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll 1
  for (int q = 0; q < kSchedBuckets; ++q) {
    const unsigned long long m = __ballot(b == q);
    if (m == 0) continue;
    const int leader = __ffsll(static_cast<long long>(m)) - 1;
    int base = 0;
    if (lane == leader) {
      base = atomicAdd(kPlace ? &cursor[q] : &count[q], __popcll(m));
    }
    if (kPlace) {
      base = __shfl(base, leader);
      if (b == q) {
        int start = 0;
        for (int p = 0; p < q; ++p) start += count[p];
        sched[start + base + __popcll(m & below)] = s;
      }
    }
  }
}

// Label of every source tile of a finished search: the largest final
// k-th-neighbour distance among the tile's sources, read back from the
// frame's out_d2 and rounded up to float (+inf where a source has fewer than
// k neighbours or out_d2 is absent). out[] starts at zero; non-negative
// floats order like their bit patterns.
__global__ __launch_bounds__(256) void k_tile_labels(
    const FrameDesc* __restrict__ fr, const long long* __restrict__ fcb, int F,
    const Cand* __restrict__ cand, const int* __restrict__ scell,
    const int* __restrict__ off, const int* __restrict__ tile_off,
    long long ncells, int* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= off[ncells]) return;
  const int g = scell[s];
  const FrameDesc& d = fr[find_frame(fcb, F, g)];
  const int T = tile_off[g] + ((s - off[g]) >> kTileShift);
  float v = __builtin_inff();
  if (d.out_d2 != nullptr) {
    const int i = static_cast<int>(cand[s].id);
    const double kth = d.out_d2[static_cast<size_t>(i) * d.k + d.k - 1];
    if (kth == kth) v = __double2float_ru(sqrt(kth));
  }
  atomicMax(&out[T], __float_as_int(v));
}

// Predicted depth of a tile under the depth schedule: the first shell whose
// stop bound covers the tile's cap (1 .. r_safe).
__device__ __forceinline__ int cap_depth(float cap, const FrameDesc& d) {
  const double q = ceil(static_cast<double>(cap) / d.cell_min);
  return q >= static_cast<double>(d.r_safe) ? d.r_safe : (q > 1.0 ? static_cast<int>(q) : 1);
}

// Depth packing of the row walk: the tiles of every 16 (counted from the
// frame's first tile, so each wave's four stay in one group and frame) in
// order of predicted depth, then tile: perm[first + rank] = tile.
#ifndef LC_DEPTH_PACK
#define LC_DEPTH_PACK 1
#endif
__global__ __launch_bounds__(256) void k_depth_order(
    const FrameDesc* __restrict__ fr, const long long* __restrict__ fcb, int F,
    const int* __restrict__ tile_off, const int* __restrict__ tile_cell, long long ncells,
    const float* __restrict__ cap, int* __restrict__ perm) {
  const int T = blockIdx.x * blockDim.x + threadIdx.x;
  if (T >= tile_off[ncells]) return;
  const int f = find_frame(fcb, F, tile_cell[T]);
  const FrameDesc& d = fr[f];
  const int tb = tile_off[fcb[f]], te = tile_off[fcb[f + 1]];
  const int g0 = tb + ((T - tb) & ~15);
  const int g1 = g0 + 16 < te ? g0 + 16 : te;
  const int my = cap_depth(cap[T], d);
  int rank = 0;
  for (int u = g0; u < g1; ++u) {
    const int du = cap_depth(cap[u], d);
    rank += (du < my || (du == my && u < T)) ? 1 : 0;
  }
  perm[g0 + rank] = T;
}

// Final k-th-neighbour distance of every source slot of the last search
// (rounded up to float, +inf without one): per-source schedule labels.
__global__ __launch_bounds__(256) void k_src_labels(
    const FrameDesc* __restrict__ fr, const long long* __restrict__ fcb, int F,
    const Cand* __restrict__ cand, const int* __restrict__ scell,
    const int* __restrict__ off, long long ncells, float* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= off[ncells]) return;
  const int g = scell[s];
  const FrameDesc& d = fr[find_frame(fcb, F, g)];
  float v = __builtin_inff();
  if (d.out_d2 != nullptr) {
    const int i = static_cast<int>(cand[s].id);
    const double kth = d.out_d2[static_cast<size_t>(i) * d.k + d.k - 1];
    if (kth == kth) v = __double2float_ru(sqrt(kth));
  }
  out[s] = v;
}

// ---------------------------------------------------------- exact walk

// f64::total_cmp key
__device__ __forceinline__ long long total_key(double x) {
  const long long b = __double_as_longlong(x);
  return b ^ static_cast<long long>(static_cast<unsigned long long>(b >> 63) >> 1);
}

// Exact-order replay of the reference walk for deferred sources: the
// lexicographic stencil (src/knearest.rs:435-465), each cell in
// decreasing index order (through byidx), KHeap::push / worst / finish
// (src/knearest.rs:106-158) with the heap in a strided global scratch.
// One lane per source; a grid-stride loop over the deferred list.
__global__ __launch_bounds__(64) void k_walk_exact(
    const FrameDesc* __restrict__ fr, const long long* __restrict__ fcb,
    int F, const Cand* __restrict__ cand, const int* __restrict__ scell,
    const int* __restrict__ off, const int* __restrict__ byidx, const int* __restrict__ defer_list,
    const int* __restrict__ defer_count, double* __restrict__ hd2,
    int* __restrict__ hidx) {
  const long long tid = blockIdx.x * static_cast<long long>(blockDim.x) +
                        threadIdx.x;
  const long long nthreads = static_cast<long long>(gridDim.x) * blockDim.x;
  const int nq = *defer_count;
  for (long long qi = tid; qi < nq; qi += nthreads) {
    const int s = defer_list[qi];
    const int g = scell[s];
    const int f = find_frame(fcb, F, g);
    const FrameDesc& d = fr[f];
    const int k = d.k;
    if (k <= kExactWaveMaxK) continue;  // k_walk_exact_wave's
    const CellGeom gm = geom_of(d);
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const long long cb = d.cell_base;
    const int c = static_cast<int>(g - cb);
    const int hx = c % nx, hy = (c / nx) % ny, hz = c / (nx * ny);
    const Cand mc = cand[s];
    const double px = mc.x, py = mc.y, pz = mc.z;
    const int me = static_cast<int>(mc.id);
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
              const Cand cq = cand[byidx[t]];
              const int j = static_cast<int>(cq.id);
              if (j == me) continue;
              const double d2 = dist2_shifted(px, py, pz, cq.x, cq.y, cq.z,
                                              shx, shy, shz);
              // KHeap::push. With a full heap and a worst slot that is a
              // number, d2 >= worst can change nothing: an index already
              // present holds at most that value, and the replacement
              // needs d2 < worst. (A NaN worst, kept by the first-argmax
              // scan when slot 0 is NaN, bounds nothing: push fully.)
              const double wv = hd[worst * st];
              if (n == k && wv == wv && !(d2 < wv)) continue;
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
                } else if (d2 < hd[worst * st]) {
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

// Wave-level KHeap::worst over the heap slots u < n (slot u = lane for the
// first register pair, kWave + lane for the second): the sequential scan of
// src/knearest.rs:135-144 (w = 0; w = u whenever d[u] > d[w]) keeps slot 0
// when it holds NaN and otherwise never moves onto a NaN, so it is the first
// argmax of the non-NaN values unless slot 0 is NaN.
__device__ __forceinline__ void wave_worst(double d0, double d1, int n,
                                           int lane, double& wv, int& ws) {
  const double first = __shfl(d0, 0, kWave);
  if (first != first) {
    wv = first;
    ws = 0;
    return;
  }
  double v = -__builtin_inf();
  int u = 1 << 30;
  if (lane < n && d0 == d0) {
    v = d0;
    u = lane;
  }
  if (kWave + lane < n && d1 == d1 && (d1 > v || u == (1 << 30))) {
    v = d1;
    u = kWave + lane;
  }
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    const double ov = __shfl_xor(v, m, kWave);
    const int ou = __shfl_xor(u, m, kWave);
    if (ov > v || (ov == v && ou < u)) {
      v = ov;
      u = ou;
    }
  }
  wv = v;
  ws = u;
}

// Exact-order replay of the reference walk and KHeap for deferred sources,
// one wave per source (k <= kExactWaveMaxK): the lexicographic stencil
// (src/knearest.rs:435-465), each cell in decreasing index order (through
// byidx), the cell's candidates loaded and measured 64 at a time. A push
// can only change the heap while it is not full or when d2 < worst (an
// index already present holds at most worst, and a replacement needs d2 <
// worst; a NaN worst bounds nothing), so only those candidates are pushed,
// one after another in visit order: the heap sees the reference's sequence
// of effective pushes. The
// heap is spread over the wave, two slots per lane; KHeap::push, worst and
// finish (src/knearest.rs:106-158) run as wave-wide ballots and reductions.
__global__ __launch_bounds__(64) void k_walk_exact_wave(
    const FrameDesc* __restrict__ fr, const long long* __restrict__ fcb,
    int F, const Cand* __restrict__ cand, const int* __restrict__ scell,
    const int* __restrict__ off, const int* __restrict__ byidx,
    const int* __restrict__ defer_list, const int* __restrict__ defer_count) {
  __shared__ long long s_key[kExactWaveMaxK];
  __shared__ int s_j[kExactWaveMaxK];
  const int lane = threadIdx.x;
  const int nq = *defer_count;
  for (int qi = blockIdx.x; qi < nq; qi += gridDim.x) {
    const int s = uni(defer_list[qi]);
    const int g = uni(scell[s]);
    const int f = uni(find_frame(fcb, F, g));
    const FrameDesc& d = fr[f];
    const int k = d.k;
    if (k > kExactWaveMaxK) continue;  // the lane-per-source replay's
    const CellGeom gm = geom_of(d);
    const int nx = d.nx, ny = d.ny, nz = d.nz;
    const long long cb = d.cell_base;
    const int c = static_cast<int>(g - cb);
    const int hx = c % nx, hy = (c / nx) % ny, hz = c / (nx * ny);
    const Cand mc = cand[s];
    const double px = mc.x, py = mc.y, pz = mc.z;
    const int me = static_cast<int>(mc.id);
    double hd0 = __builtin_inf(), hd1 = __builtin_inf();
    int hj0 = -1, hj1 = -1;
    int n = 0;            // filled slots (uniform)
    double worst = 0.0;   // KHeap::worst value when n == k (uniform)
    int wslot = 0;        // its slot
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
            const int a0 = uni(off[cb + cq]);
            const int a1 = uni(off[cb + cq + 1]);
            for (int top = a1 - 1; top >= a0; top -= kWave) {
              const int t = top - lane;  // decreasing index order
              const bool valid = t >= a0;
              int j = -1;
              double d2 = __builtin_inf();
              if (valid) {
                const Cand q = cand[byidx[t]];
                j = static_cast<int>(q.id);
                d2 = dist2_shifted(px, py, pz, q.x, q.y, q.z, shx, shy, shz);
              }
              // a NaN worst bounds nothing (see k_walk_exact)
              const bool need =
                  valid && j != me && (n < k || d2 < worst || worst != worst);
              unsigned long long m = __ballot(need);
              while (m != 0) {
                const int src = __ffsll(static_cast<long long>(m)) - 1;
                m &= m - 1;
                const int cj = __shfl(j, src, kWave);
                const double cd = __shfl(d2, src, kWave);
                if (n == k && worst == worst && !(cd < worst)) continue;  // no effect
                const bool in0 = lane < n && hj0 == cj;
                const bool in1 = kWave + lane < n && hj1 == cj;
                if (__ballot(in0 || in1) != 0) {
                  if (in0 && cd < hd0) hd0 = cd;
                  if (in1 && cd < hd1) hd1 = cd;
                } else if (n < k) {
                  if (n < kWave) {
                    if (lane == n) {
                      hd0 = cd;
                      hj0 = cj;
                    }
                  } else if (lane == n - kWave) {
                    hd1 = cd;
                    hj1 = cj;
                  }
                  n += 1;
                } else if (!(cd < worst)) {
                  // full, absent, not below worst (a NaN worst): no change
                } else if (wslot < kWave) {
                  if (lane == wslot) {
                    hd0 = cd;
                    hj0 = cj;
                  }
                } else if (lane == wslot - kWave) {
                  hd1 = cd;
                  hj1 = cj;
                }
                if (n == k) wave_worst(hd0, hd1, n, lane, worst, wslot);
              }
            }
          }
        }
      }
      if (n >= k) {
        double w;
        int ws;
        wave_worst(hd0, hd1, n, lane, w, ws);
        const double bound = static_cast<double>(reach) * d.cell_min;
        if (w <= bound * bound) {
          final_reach = reach;
          break;
        }
      }
    }
    // KHeap::finish: order by (total_cmp(d2), index); indices are distinct
    if (lane < n) {
      s_key[lane] = total_key(hd0);
      s_j[lane] = hj0;
    }
    if (kWave + lane < n) {
      s_key[kWave + lane] = total_key(hd1);
      s_j[kWave + lane] = hj1;
    }
    __syncthreads();  // one wave per block
    int r0 = 0, r1 = 0;
    const long long k0 = total_key(hd0), k1 = total_key(hd1);
    for (int u = 0; u < n; ++u) {
      const long long ku = s_key[u];
      const int ju = s_j[u];
      r0 += (ku < k0 || (ku == k0 && ju < hj0)) ? 1 : 0;
      r1 += (ku < k1 || (ku == k1 && ju < hj1)) ? 1 : 0;
    }
    const size_t row = static_cast<size_t>(me) * k;
    if (lane < n) {
      d.out_nn[row + r0] = hj0;
      if (d.out_d2) d.out_d2[row + r0] = hd0;
    }
    if (kWave + lane < n) {
      d.out_nn[row + r1] = hj1;
      if (d.out_d2) d.out_d2[row + r1] = hd1;
    }
    for (int t = n + lane; t < k; t += kWave) {
      d.out_nn[row + t] = -1;
      if (d.out_d2) d.out_d2[row + t] = nan_value();
    }
    if (lane == 0 && d.out_reach) d.out_reach[me] = final_reach;
    __syncthreads();  // the next source reuses s_key / s_j
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

void Engine::read_stats(unsigned long long* out, int n) {
  sync();
  const long long m = std::min<long long>(n, stats_n_);
  if (stats_.p != nullptr && m > 0) {
    check(hipMemcpy(out, stats_.p, sizeof(unsigned long long) * static_cast<size_t>(m),
                    hipMemcpyDeviceToHost),
          "DtoH stats");
  }
  for (long long i = std::max<long long>(m, 0); i < n; ++i) out[i] = 0;
}

int fast_bucket(int k) {
  if (k <= 4) return 4;
  if (k <= 8) return 8;
  if (k <= 16) return 16;
  if (k <= 32) return 32;
  return 0;
}

Engine::Engine() {
  check(hipGetDevice(&device_), "hipGetDevice");
  // tuning override, in the spirit of the reference's LINKCELL_TPP /
  // LINKCELL_BLOCK (src/gpu/knearest.cpp:948-960)
  if (const char* v = std::getenv("LINKCELL_DENSE_MIN")) {
    const int x = std::atoi(v);
    if (x >= 1) dense_min_override_ = x;
  }
  // LINKCELL_TPP keeps its meaning of "threads per source": 1 sends every
  // source to the lane-per-source walk, 16 selects the row walk (16-lane
  // rows per 16-source tile), 64 sends every cell to the wave-per-cell
  // walk.
  if (const char* v = std::getenv("LINKCELL_TPP")) {
    const int x = std::atoi(v);
    if (x == 1) dense_min_override_ = kNeverDense;
    if (x == 16) dense_min_override_ = kRowWalk;
    if (x >= 64) dense_min_override_ = 1;
  }
}

Engine::~Engine() {
  if (stream_ != nullptr) {
    (void)hipStreamSynchronize(stream_);
  }
  free_all();
  if (graph_exec_ != nullptr) (void)hipGraphExecDestroy(graph_exec_);
  if (graph_staging_ != nullptr) (void)hipHostFree(graph_staging_);
  if (staging_ != nullptr) (void)hipHostFree(staging_);
  if (staging_ev_ != nullptr) (void)hipEventDestroy(staging_ev_);
  for (auto& slot : ring_) {
    for (auto& e : slot) {
      if (e != nullptr) (void)hipEventDestroy(e);
    }
  }
  if (stream_ != nullptr) (void)hipStreamDestroy(stream_);
}

hipStream_t Engine::stream() {
  if (stream_ == nullptr) {
    // default flags, as the reference's cudaStreamCreate
    // (src/gpu/knearest.cpp:586): the stream is ordered against the null
    // stream, so a caller's fill or producer kernel there (lc_gpu_fill_i32,
    // python/src/gpu.rs:293-346) runs before the search that follows it
    check(hipStreamCreate(&stream_), "hipStreamCreate");
  }
  return stream_;
}

void Engine::sync() {
  if (stream_ != nullptr) check(hipStreamSynchronize(stream_), "stream wait");
  check_dcell_status();
}

void Engine::grow(Buf& b, size_t bytes) {
  if (bytes <= b.bytes && b.p != nullptr) return;
  if (capturing_) {
    throw std::runtime_error(
        "graph capture needs buffers sized by a prior run of the same batch");
  }
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
  Buf* all[] = {&tile_perm_, &frames_,   &fpt_,      &fcb_,    &count_,  &status_,
                &counters_, &off_,      &tile_off_, &item_off_, &items_,
                &cellslot_, &tmp_key_,  &cand_,
                &scell_,    &byidx_,  &tbox_,   &cbox_,
                &defer_,    &heap_d2_,  &heap_idx_, &flags_, &stats_,
                &sched_,    &sched_cnt_, &tile_cell_,
                &pred_tile_, &labels_, &cap_list_};
  for (Buf* b : all) {
    if (b->p != nullptr) (void)hipFree(b->p);
    b->p = nullptr;
    b->bytes = 0;
  }
}

namespace {

Timings timings_of(hipEvent_t* ev) {
  auto el = [&](int a, int b) {
    float ms = 0;
    check(hipEventElapsedTime(&ms, ev[a], ev[b]), "elapsed");
    return ms;
  };
  Timings t;
  t.fold = el(0, 1);
  t.scan = el(1, 2);
  t.scatter = el(2, 3);
  t.sort = el(3, 4);
  t.walk = el(4, 5);
  t.exact = el(5, 6);
  t.total = el(0, 7);
  t.setup = el(4, 8);
  t.pass1 = el(8, 9);
  t.retry = el(9, 5);
  return t;
}

}  // namespace

const Timings& Engine::last_timings() {
  if (timings_pending_) {
    hipEvent_t* ev = ring_[(ring_pos_ - 1) % kTimingRing];
    check(hipEventSynchronize(ev[7]), "event wait");
    timings_ = timings_of(ev);
    int dc = 0;
    check(hipMemcpy(&dc, as<int>(counters_.p) + 1, sizeof(int),
                    hipMemcpyDeviceToHost),
          "DtoH deferred");
    timings_.deferred = dc;
    int rt = 0;
    check(hipMemcpy(&rt, as<int>(counters_.p) + 3, sizeof(int),
                    hipMemcpyDeviceToHost),
          "DtoH retried");
    timings_.retried = rt;
    timings_pending_ = false;
  }
  return timings_;
}

std::vector<Timings> Engine::timing_history(int n) {
  const Timings last = last_timings();
  const long long have = std::min<long long>(ring_pos_, kTimingRing);
  const long long m = std::min<long long>(std::max(n, 0), have);
  std::vector<Timings> out;
  for (long long i = ring_pos_ - m; i < ring_pos_; ++i) {
    out.push_back(i == ring_pos_ - 1 ? last : timings_of(ring_[i % kTimingRing]));
  }
  return out;
}

namespace {

// Sizes of one call's device buffers.
struct Sizes {
  int F = 0;
  long long ntot = 0, ctot = 0;
  long long tiles = 0;      // scan tiles
  long long tcap = 0;       // bound on 16-point source tiles
  long long max_items = 0;  // bound on 64-source work items
  long long exact_threads = 0;
  int kmax_exact = 1;
};

// Whether the row walk of some frame lists half tiles (a heap of
// LC_ROW_HALF_KC slots or more: k above half of it), so the preparation
// also builds the half-tile boxes.
bool half_boxes(const Sizes& z) {
  return LC_ROW_HALF != 0 && z.kmax_exact > LC_ROW_HALF_KC / 2;
}

Sizes sizes_of(int F, long long ntot, long long ctot, int kmax_exact,
               long long heap_budget) {
  Sizes z;
  z.F = F;
  z.ntot = ntot;
  z.ctot = ctot;
  z.kmax_exact = kmax_exact;
  z.tiles = ctot / kScanTile + 1;
  // 16-point tiles: at most one partial tile per non-empty cell
  z.tcap = ntot / kTile + std::min(ntot, ctot) + 1;
  z.max_items = std::min<long long>(ntot, ctot + ntot / kWave + 1) + 1;
  // exact-walk scratch: one heap of kmax slots per resident lane
  constexpr long long kExactBlocks = 1024;
  z.exact_threads = kExactBlocks * 64;
  while (z.exact_threads > 64 && z.exact_threads * kmax_exact * 12 > heap_budget) {
    z.exact_threads /= 2;
  }
  return z;
}

// Device buffers of one call, owned by an engine or carved from a fixed
// scratch region. Order and sizes: buffer_bytes.
struct Bufs {
  FrameDesc* frames;
  long long* fpt;
  long long* fcb;
  int* count;
  unsigned long long* status;
  int* counters;
  int* off;
  int* tile_off;
  int* item_off;
  int* items;
  int4* cellslot;
  long long* tmp_key;
  Cand* cand;
  int* scell;
  int* byidx;
  Box* tbox;
  int* tile_cell;
  Box* cbox;
  int* defer;
  int* flags;
  double* heap_d2;
  int* heap_idx;
};
constexpr int kNumBufs = sizeof(Bufs) / sizeof(void*);

void buffer_bytes(const Sizes& z, size_t (&b)[kNumBufs]) {
  const size_t F = static_cast<size_t>(z.F), N = static_cast<size_t>(z.ntot),
               C = static_cast<size_t>(z.ctot);
  const size_t hs = static_cast<size_t>(z.exact_threads) * static_cast<size_t>(z.kmax_exact);
  const size_t v[kNumBufs] = {
      sizeof(FrameDesc) * F,                         // frames
      sizeof(long long) * (F + 1),                   // fpt
      sizeof(long long) * (F + 1),                   // fcb
      sizeof(int) * C,                               // count
      sizeof(unsigned long long) * 2 * static_cast<size_t>(z.tiles),  // status
      sizeof(int) * 12,                              // counters (6: device-cell plan status, 8..11: walk queues)
      sizeof(int) * (C + 1),                         // off
      sizeof(int) * (C + 1),                         // tile_off
      sizeof(int) * (C + 1),                         // item_off
      sizeof(int) * static_cast<size_t>(z.max_items),  // items
      sizeof(int4) * N,                              // cellslot
      sizeof(long long) * N,                         // tmp_key
      sizeof(Cand) * (N + kPad),                     // cand
      sizeof(int) * N,                               // scell
      sizeof(int) * N,                               // byidx
      // tbox, then (half tiles) two boxes per tile
      sizeof(Box) * static_cast<size_t>(z.tcap) * (half_boxes(z) ? 3 : 1),
      sizeof(int) * static_cast<size_t>(z.tcap),     // tile_cell
      sizeof(Box) * C,                               // cbox
      sizeof(int) * N,                               // defer
      sizeof(int) * F,                               // flags
      sizeof(double) * hs,                           // heap_d2
      sizeof(int) * hs,                              // heap_idx
  };
  for (int i = 0; i < kNumBufs; ++i) b[i] = v[i];
}

// Lays the buffers out from base (256-byte aligned each); returns the bytes
// used. base may be 0 to measure.
size_t carve(const Sizes& z, uintptr_t base, Bufs& out) {
  size_t b[kNumBufs];
  buffer_bytes(z, b);
  void** slot = reinterpret_cast<void**>(&out);
  size_t used = 0;
  for (int i = 0; i < kNumBufs; ++i) {
    slot[i] = reinterpret_cast<void*>(base + used);
    used += (std::max<size_t>(b[i], 1) + 255) & ~static_cast<size_t>(255);
  }
  return used;
}

// Clears, fold/count, scans, scatter, cell sort and boxes. ev: the call's
// timing events (null = untimed); 0..4 are recorded here.
void launch_prep(const Sizes& z, const Bufs& B, hipStream_t st, hipEvent_t* ev) {
  const long long ctot = z.ctot, ntot = z.ntot, tiles = z.tiles;
  const int F = z.F;
  if (ev) check(hipEventRecord(ev[0], st), "event record");
  // clears (fill kernels, not memset nodes, so the sequence can be
  // captured into a graph)
  k_zero_u32<<<blocks_for(ctot, 256), 256, 0, st>>>(reinterpret_cast<unsigned*>(B.count),
                                                    ctot);
  k_zero_u32<<<blocks_for(tiles * 4, 256), 256, 0, st>>>(
      reinterpret_cast<unsigned*>(B.status), tiles * 4);
  // counters 0-5: scans, deferred sources, retried rows, learned-cap tiles
  k_zero_u32<<<1, 64, 0, st>>>(reinterpret_cast<unsigned*>(B.counters), 6);
  k_zero_u32<<<blocks_for(F, 256), 256, 0, st>>>(reinterpret_cast<unsigned*>(B.flags), F);

  k_fold_count<<<blocks_for(ntot, 256), 256, 0, st>>>(B.frames, B.fpt, F, ntot, B.count,
                                                      B.cellslot, B.flags);
  check(hipGetLastError(), "fold_count launch");
  if (ev) check(hipEventRecord(ev[1], st), "event record");

  k_scan<<<static_cast<unsigned>(tiles), kScanThreads, 0, st>>>(
      B.count, ctot, kTileShift, B.off, B.tile_off, B.status, B.counters + 0);
  k_scan<<<static_cast<unsigned>(tiles), kScanThreads, 0, st>>>(
      B.count, ctot, kChunkShift, nullptr, B.item_off, B.status + tiles, B.counters + 2);
  check(hipGetLastError(), "scan launch");
  if (ev) check(hipEventRecord(ev[2], st), "event record");

  k_emit_items<<<blocks_for(ctot, 256), 256, 0, st>>>(B.count, B.item_off, B.tile_off, ctot,
                                                      B.items, B.tile_cell);
  k_scatter<<<blocks_for(ntot, 256), 256, 0, st>>>(B.cellslot, B.fpt, F, ntot, B.off,
                                                   B.tmp_key);
  check(hipGetLastError(), "scatter launch");
  if (ev) check(hipEventRecord(ev[3], st), "event record");

  const unsigned item_blocks = blocks_for(z.max_items * kWave, 256);
  k_cell_sort<<<item_blocks, 256, 0, st>>>(B.items, B.item_off, ctot, B.off, B.tmp_key,
                                           B.frames, B.fcb, F, B.cand, B.scell, B.byidx);
  k_tile_box<<<item_blocks, 256, 0, st>>>(B.items, B.item_off, ctot, B.off, B.tile_off,
                                          B.cand, B.tbox,
                                          half_boxes(z) ? B.tbox + z.tcap : nullptr);
  k_cell_box<<<blocks_for(ctot, 256), 256, 0, st>>>(B.off, B.tile_off, ctot, B.tbox, B.cbox);
  check(hipGetLastError(), "cell_sort launch");
  if (ev) check(hipEventRecord(ev[4], st), "event record");
}

// Walk arguments over the call's buffers, plain schedule, no counters.
// Compute units of the current device (looked up once per device).
int device_cus() {
  static std::mutex mu;
  static int cus[64] = {};
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= 64) return 256;
  std::lock_guard<std::mutex> lock(mu);
  if (cus[dev] == 0) {
    int v = 0;
    check(hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev),
          "hipDeviceGetAttribute");
    cus[dev] = v > 0 ? v : 256;
  }
  return cus[dev];
}

WalkArgs walk_args(const Sizes& z, const Bufs& B, int dense_min) {
  WalkArgs wa;
  std::memset(&wa, 0, sizeof wa);
  wa.fr = B.frames;
  wa.fcb = B.fcb;
  wa.F = z.F;
  wa.fflags = B.flags;
  wa.cand = B.cand;
  wa.scell = B.scell;
  wa.off = B.off;
  wa.tile_off = B.tile_off;
  wa.items = B.items;
  wa.item_off = B.item_off;
  wa.ncells = z.ctot;
  wa.tbox = B.tbox;
  wa.hbox = half_boxes(z) ? B.tbox + z.tcap : nullptr;
  wa.cbox = B.cbox;
  wa.dense_min = dense_min;
  wa.tile_cell = B.tile_cell;
  wa.defer_list = B.defer;
  wa.defer_count = B.counters + 1;
  // capped row walk: rows whose cap may have hidden the reference's stop
  // are walked again, uncapped, by their own wave; counted here
  wa.retry_count = B.counters + 3;
  return wa;
}

// The fast walks of every register bucket in use, then the exact-order
// walk of the deferred and unbucketed sources. ev: events 8, 9, 5, 6, 7.
void launch_walk(const Sizes& z, const Bufs& B, const WalkArgs& wa,
                 const bool (&any_fast)[33], bool any_exact_frame, hipStream_t st,
                 hipEvent_t* ev) {
  const bool row_mode = wa.dense_min == kRowWalk || wa.dense_min == kRowWalkPersist;
  const unsigned sparse_blocks = blocks_for(z.ntot, 256);
  // row walk: one wave per 4 source tiles, frames on wave boundaries
  const unsigned row_blocks = blocks_for(z.tcap / 4 + z.F + 1, 4);
  // persistent row walk: as many blocks as the GPU holds at once, each
  // bucket's waves taking groups from their own counter (cleared here by a
  // fill kernel, so the sequence can be captured into a graph)
  const bool persist = row_mode && wa.queue != nullptr && wa.stats == nullptr;
  WalkArgs pa = wa;
  pa.n_waves = static_cast<int>(row_blocks) * 4;
  const unsigned cus = static_cast<unsigned>(device_cus());
  if (persist) {
    k_zero_u32<<<1, 64, 0, st>>>(reinterpret_cast<unsigned*>(wa.queue), 4);
    check(hipGetLastError(), "queue clear launch");
  }
  if (ev) check(hipEventRecord(ev[8], st), "event record");
  if (LC_ROW_HALF != 0 && row_mode && wa.hbox == nullptr) {
    for (int kc = LC_ROW_HALF_KC; kc <= 32; ++kc) {
      if (any_fast[kc]) throw std::runtime_error("row walk: half-tile boxes not built");
    }
  }
#define LC_WALK(KC)                                                             \
  if (any_fast[KC] && row_mode && persist) {                                    \
    constexpr int qi = KC == 4 ? 0 : KC == 8 ? 1 : KC == 16 ? 2 : 3;            \
    const unsigned pb = std::min(row_blocks, cus * LC_ROW_BLOCKS(KC));          \
    pa.queue = wa.queue + qi;                                                   \
    if (wa.depth) k_walk_rows<KC, false, true, true><<<pb, 256, 0, st>>>(pa);   \
    else k_walk_rows<KC, false, false, true><<<pb, 256, 0, st>>>(pa);           \
    check(hipGetLastError(), "walk launch");                                    \
  } else if (any_fast[KC] && row_mode) {                                        \
    if (wa.stats != nullptr) {                                                  \
      if (wa.depth) k_walk_rows<KC, true, true, false><<<row_blocks, 256, 0, st>>>(wa);  \
      else k_walk_rows<KC, true, false, false><<<row_blocks, 256, 0, st>>>(wa);         \
    } else {                                                                    \
      if (wa.depth) k_walk_rows<KC, false, true, false><<<row_blocks, 256, 0, st>>>(wa); \
      else k_walk_rows<KC, false, false, false><<<row_blocks, 256, 0, st>>>(wa);        \
    }                                                                           \
    check(hipGetLastError(), "walk launch");                                    \
  } else if (any_fast[KC]) {                                                    \
    constexpr int W = WalkShape<KC>::kWaves;                                    \
    k_walk_dense<KC><<<static_cast<unsigned>((z.max_items + W - 1) / W),        \
                       W * kWave, 0, st>>>(wa);                                 \
    k_walk_sparse<KC><<<blocks_for(z.ntot, W * kWave), W * kWave, 0, st>>>(wa); \
    check(hipGetLastError(), "walk launch");                                    \
  }
  LC_WALK(4)
  LC_WALK(8)
  LC_WALK(16)
  LC_WALK(32)
#undef LC_WALK
  if (ev) check(hipEventRecord(ev[9], st), "event record");
  // frames outside every register bucket go straight to the exact walk
  if (any_exact_frame) {
    k_route_exact<<<sparse_blocks, 256, 0, st>>>(B.frames, B.fcb, z.F, B.scell, B.off,
                                                 z.ctot, B.defer, B.counters + 1);
    check(hipGetLastError(), "route launch");
  }
  if (ev) check(hipEventRecord(ev[5], st), "event record");
  k_walk_exact_wave<<<2048, 64, 0, st>>>(B.frames, B.fcb, z.F, B.cand, B.scell, B.off,
                                          B.byidx, B.defer, B.counters + 1);
  if (z.kmax_exact > kExactWaveMaxK) {
    k_walk_exact<<<static_cast<unsigned>(z.exact_threads / 64), 64, 0, st>>>(
        B.frames, B.fcb, z.F, B.cand, B.scell, B.off, B.byidx, B.defer, B.counters + 1,
        B.heap_d2, B.heap_idx);
  }
  check(hipGetLastError(), "exact walk launch");
  if (ev) {
    check(hipEventRecord(ev[6], st), "event record");
    check(hipEventRecord(ev[7], st), "event record");
  }
}

// Frame descriptor of one frame spec (bases and bucket left to the caller).
void fill_desc(const FrameSpec& s, FrameDesc& d) {
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
}

}  // namespace

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
  bool any_exact_frame = false;
  for (int f = 0; f < F; ++f) {
    const FrameSpec& s = frames[static_cast<size_t>(f)];
    FrameDesc& d = desc[static_cast<size_t>(f)];
    fill_desc(s, d);
    d.pt_base = ntot;
    d.cell_base = ctot;
    d.kc = mode_ == WalkMode::kExactOnly ? 0 : fast_bucket(s.k);
    if (d.kc) any_fast[d.kc] = true;
    any_exact_frame |= d.kc == 0;
    kmax_exact = std::max(kmax_exact, s.k);
    fpt[static_cast<size_t>(f)] = ntot;
    fcb[static_cast<size_t>(f)] = ctot;
    ntot += s.n;
    ctot += s.plan.ncell;
  }
  fpt[static_cast<size_t>(F)] = ntot;
  fcb[static_cast<size_t>(F)] = ctot;
  if (ntot >= INT_MAX / 2 || ctot >= INT_MAX / 2) {
    throw std::runtime_error("batch exceeds 2^30 points or cells");
  }
  last_F_ = F;
  last_ntot_ = ntot;
  last_ctot_ = ctot;
  const Sizes z = sizes_of(F, ntot, ctot, kmax_exact, 512ll << 20);
  const long long tcap = z.tcap;
  {
    size_t bytes[kNumBufs];
    buffer_bytes(z, bytes);
    Buf* own[kNumBufs] = {&frames_,  &fpt_,   &fcb_,      &count_,    &status_, &counters_,
                          &off_,     &tile_off_, &item_off_, &items_,   &cellslot_,
                          &tmp_key_, &cand_,  &scell_,    &byidx_,    &tbox_,   &tile_cell_,
                          &cbox_,    &defer_, &flags_,    &heap_d2_,  &heap_idx_};
    for (int i = 0; i < kNumBufs; ++i) grow(*own[i], bytes[i]);
  }
  Bufs B;
  B.frames = as<FrameDesc>(frames_.p);
  B.fpt = as<long long>(fpt_.p);
  B.fcb = as<long long>(fcb_.p);
  B.count = as<int>(count_.p);
  B.status = as<unsigned long long>(status_.p);
  B.counters = as<int>(counters_.p);
  B.off = as<int>(off_.p);
  B.tile_off = as<int>(tile_off_.p);
  B.item_off = as<int>(item_off_.p);
  B.items = as<int>(items_.p);
  B.cellslot = as<int4>(cellslot_.p);
  B.tmp_key = as<long long>(tmp_key_.p);
  B.cand = as<Cand>(cand_.p);
  B.scell = as<int>(scell_.p);
  B.byidx = as<int>(byidx_.p);
  B.tbox = as<Box>(tbox_.p);
  B.tile_cell = as<int>(tile_cell_.p);
  B.cbox = as<Box>(cbox_.p);
  B.defer = as<int>(defer_.p);
  B.flags = as<int>(flags_.p);
  B.heap_d2 = as<double>(heap_d2_.p);
  B.heap_idx = as<int>(heap_idx_.p);

  // descriptors go through pinned staging; wait until the previous call's
  // copy has consumed it
  const size_t need = sizeof(FrameDesc) * static_cast<size_t>(F) +
                      2 * sizeof(long long) * (static_cast<size_t>(F) + 1);
  char* sp = nullptr;
  if (capturing_) {
    // the graph's memcpy nodes read this buffer at every replay
    if (need > graph_staging_bytes_) {
      throw std::runtime_error("graph staging buffer too small");
    }
    sp = static_cast<char*>(graph_staging_);
  } else {
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
    sp = static_cast<char*>(staging_);
  }
  const size_t dbytes = sizeof(FrameDesc) * static_cast<size_t>(F);
  const size_t bbytes = sizeof(long long) * (static_cast<size_t>(F) + 1);
  std::memcpy(sp, desc.data(), dbytes);
  std::memcpy(sp + dbytes, fpt.data(), bbytes);
  std::memcpy(sp + dbytes + bbytes, fcb.data(), bbytes);
  check(hipMemcpyAsync(B.frames, sp, dbytes, hipMemcpyHostToDevice, st), "HtoD frames");
  check(hipMemcpyAsync(B.fpt, sp + dbytes, bbytes, hipMemcpyHostToDevice, st), "HtoD fpt");
  check(hipMemcpyAsync(B.fcb, sp + dbytes + bbytes, bbytes, hipMemcpyHostToDevice, st),
        "HtoD fcb");
  if (!capturing_) {
    check(hipEventRecord(staging_ev_, st), "event record");
    staging_pending_ = true;
  }
  const bool timing = timing_ && !capturing_;
  if (timing) {
    // a slot comes round again kTimingRing timed runs later (recording an
    // event again while its last record is pending is allowed)
    ev_ = ring_[ring_pos_ % kTimingRing];
    for (int e = 0; e < 10; ++e) {
      if (ev_[e] == nullptr) check(hipEventCreate(&ev_[e]), "hipEventCreate");
    }
  }
  hipEvent_t* ev = timing ? ev_ : nullptr;

  launch_prep(z, B, st, ev);

  // walk split: kRowWalk = rows everywhere; positive = a wave per dense
  // cell, one lane per source elsewhere
  WalkArgs wa = walk_args(z, B, dense_min_for(frames));
  const bool row_mode = wa.dense_min == kRowWalk || wa.dense_min == kRowWalkPersist;
  if (wa.dense_min == kRowWalkPersist) wa.queue = B.counters + 8;
  if (stats_on_ && !capturing_) {
    const long long nstat = kStatCount + 2 * (ntot / kWave + 1);
    grow(stats_, sizeof(unsigned long long) * static_cast<size_t>(nstat));
    k_zero_u32<<<blocks_for(2 * nstat, 256), 256, 0, st>>>(
        as<unsigned>(stats_.p), 2 * nstat);
    stats_n_ = nstat;
    wa.stats = as<unsigned long long>(stats_.p);
  }
  const float* pred = nullptr;
  if ((sched_mode_ == 1 || sched_mode_ == 3) && pred_ != nullptr && pred_->trained() &&
      row_mode) {
    // learned caps of every source tile, from the counts and tile boxes
    // just built
    grow(pred_tile_, sizeof(float) * static_cast<size_t>(tcap));
    if (sched_mode_ == 1 && LC_CAP_CELL_MAX > 0) {
      // caps only for the tiles of sparse home cells, inferred for those
      // alone
      grow(cap_list_, sizeof(int) * static_cast<size_t>(tcap));
      int* const lst = as<int>(cap_list_.p);
      int* const lcnt = B.counters + 5;  // cleared with the preparation's counters
      k_cap_tiles<<<blocks_for(ctot, kCapThreads), kCapThreads, 0, st>>>(B.count, B.tile_off, ctot,
                                                         LC_CAP_CELL_MAX,
                                                         as<float>(pred_tile_.p), lst, lcnt);
      check(hipGetLastError(), "cap tiles launch");
      pred_->predict_caps(tile_src(ctot, F), tcap, as<float>(pred_tile_.p), st, nullptr, lst,
                          lcnt);
    } else {
      pred_->predict_caps(tile_src(ctot, F), tcap, as<float>(pred_tile_.p), st);
    }
    wa.pred_tile = as<float>(pred_tile_.p);
    wa.pred_tile_n = tcap;
    if (sched_mode_ == 3) {
      // depth schedule: shells below each tile's predicted depth in one
      // pass, tiles packed by that depth
      wa.depth = 1;
      if (LC_DEPTH_PACK) {
        grow(tile_perm_, sizeof(int) * static_cast<size_t>(tcap));
        k_depth_order<<<blocks_for(tcap, 256), 256, 0, st>>>(
            B.frames, B.fcb, F, B.tile_off, B.tile_cell, ctot, as<float>(pred_tile_.p),
            as<int>(tile_perm_.p));
        check(hipGetLastError(), "depth order launch");
        wa.tile_perm = as<int>(tile_perm_.p);
      }
    }
  }
  if (sched_mode_ == 2 && given_src_ != nullptr) {
    wa.pred_src = given_src_;
    wa.pred_src_n = given_src_n_;
  }
  if (sched_mode_ == 2 && given_src_ == nullptr) {
    if (given_tile_ != nullptr) {
      wa.pred_tile = given_tile_;
      wa.pred_tile_n = given_tile_n_;
    } else {
      if (given_pred_ == nullptr || given_pred_n_ != ctot) {
        throw std::runtime_error(
            "schedule prediction length must equal the batch's cell count");
      }
      pred = given_pred_;
    }
  }
  wa.pred = pred;
  // depth order of the lane-per-source walk from the per-cell prediction
  if (pred != nullptr && !row_mode) {
    grow(sched_, sizeof(int) * static_cast<size_t>(ntot));
    grow(sched_cnt_, sizeof(int) * 2 * kSchedBuckets);
    int* cnt = as<int>(sched_cnt_.p);
    k_zero_u32<<<1, 64, 0, st>>>(as<unsigned>(sched_cnt_.p), 2 * kSchedBuckets);
    k_sched<false><<<blocks_for(ntot, 256), 256, 0, st>>>(
        B.frames, B.fcb, F, B.scell, B.off, ctot, pred, wa.dense_min, cnt,
        cnt + kSchedBuckets, as<int>(sched_.p));
    k_sched<true><<<blocks_for(ntot, 256), 256, 0, st>>>(
        B.frames, B.fcb, F, B.scell, B.off, ctot, pred, wa.dense_min, cnt,
        cnt + kSchedBuckets, as<int>(sched_.p));
    check(hipGetLastError(), "schedule launch");
    wa.sched = as<int>(sched_.p);
    wa.sched_count = cnt;
  }
  launch_walk(z, B, wa, any_fast, any_exact_frame, st, ev);
  if (timing) {
    ring_pos_ += 1;
    timings_pending_ = true;
  }
  if (wait && !capturing_) sync();
}

void Engine::run_dcell(const FrameSpec& spec, int n_frames, const double* cell,
                       int cell_n, double hint, bool wait) {
  if (n_frames <= 0) return;
  hipStream_t st = stream();
  const int F = n_frames;
  const long long n = spec.n;
  const long long ntot = n * F;
  const long long ctot = dcell_cap_ * F;
  if (ntot >= INT_MAX / 2 || ctot >= INT_MAX / 2) {
    throw std::runtime_error("batch exceeds 2^30 points or cells");
  }
  // an unread status of an earlier search is reported first
  if (dcell_pending_) sync();
  last_F_ = F;
  last_ntot_ = ntot;
  last_ctot_ = ctot;
  const int kc = mode_ == WalkMode::kExactOnly ? 0 : fast_bucket(spec.k);
  const Sizes z = sizes_of(F, ntot, ctot, std::max(1, spec.k), 512ll << 20);
  {
    size_t bytes[kNumBufs];
    buffer_bytes(z, bytes);
    bytes[5] = sizeof(int) * 12;  // counters: the plan status rides in slot 6
    Buf* own[kNumBufs] = {&frames_,  &fpt_,   &fcb_,      &count_,    &status_, &counters_,
                          &off_,     &tile_off_, &item_off_, &items_,   &cellslot_,
                          &tmp_key_, &cand_,  &scell_,    &byidx_,    &tbox_,   &tile_cell_,
                          &cbox_,    &defer_, &flags_,    &heap_d2_,  &heap_idx_};
    for (int i = 0; i < kNumBufs; ++i) grow(*own[i], bytes[i]);
  }
  Bufs B;
  void** slot = reinterpret_cast<void**>(&B);
  Buf* own[kNumBufs] = {&frames_,  &fpt_,   &fcb_,      &count_,    &status_, &counters_,
                        &off_,     &tile_off_, &item_off_, &items_,   &cellslot_,
                        &tmp_key_, &cand_,  &scell_,    &byidx_,    &tbox_,   &tile_cell_,
                        &cbox_,    &defer_, &flags_,    &heap_d2_,  &heap_idx_};
  for (int i = 0; i < kNumBufs; ++i) slot[i] = own[i]->p;
  FrameDesc d;
  std::memset(&d, 0, sizeof d);
  d.xyz = spec.xyz;
  d.mask = spec.mask;
  d.out_nn = spec.out_nn;
  d.out_d2 = spec.out_d2;
  d.out_reach = spec.out_reach;
  d.n = spec.n;
  d.k = spec.k;
  d.kc = kc;
  k_plan_dcell<<<1, 64, 0, st>>>(cell, cell_n, hint, d, F, n, spec.k, dcell_cap_, B.frames,
                                 B.fpt, B.fcb, B.counters + 6);
  check(hipGetLastError(), "plan launch");
  const bool timing = timing_ && !capturing_;
  if (timing) {
    ev_ = ring_[ring_pos_ % kTimingRing];
    for (int e = 0; e < 10; ++e) {
      if (ev_[e] == nullptr) check(hipEventCreate(&ev_[e]), "hipEventCreate");
    }
  }
  hipEvent_t* ev = timing ? ev_ : nullptr;
  launch_prep(z, B, st, ev);
  bool any_fast[33] = {};
  if (kc) any_fast[kc] = true;
  std::vector<FrameSpec> one(1, spec);
  launch_walk(z, B, walk_args(z, B, dense_min_for(one)), any_fast, kc == 0, st, ev);
  if (timing) {
    ring_pos_ += 1;
    timings_pending_ = true;
  }
  dcell_pending_ = true;
  if (wait) sync();
}

void Engine::check_dcell_status() {
  if (!dcell_pending_) return;
  dcell_pending_ = false;
  int stv = 0;
  check(hipMemcpy(&stv, as<int>(counters_.p) + 6, sizeof(int), hipMemcpyDeviceToHost),
        "DtoH plan status");
  if (stv == 1) throw std::runtime_error("bad cell");
  if (stv == 2) throw std::runtime_error("too many cells");
}

void Engine::dcell_plan(const double* cell, int cell_n, double hint, double* out) {
  hipStream_t st = stream();
  sync();
  void* dbuf = nullptr;
  const size_t bytes = sizeof(FrameDesc) + 2 * sizeof(long long) * 2 + sizeof(int) * 2;
  check(hipMalloc(&dbuf, bytes), "hipMalloc");
  FrameDesc d;
  std::memset(&d, 0, sizeof d);
  FrameDesc got;
  int stv = 0;
  long long fcb[2] = {0, 0};
  try {
    char* b = static_cast<char*>(dbuf);
    FrameDesc* df = reinterpret_cast<FrameDesc*>(b);
    long long* dfpt = reinterpret_cast<long long*>(b + sizeof(FrameDesc));
    long long* dfcb = dfpt + 2;
    int* dst = reinterpret_cast<int*>(dfcb + 2);
    k_plan_dcell<<<1, 64, 0, st>>>(cell, cell_n, hint, d, 1, 0, 0, kMaxCells, df, dfpt, dfcb,
                                   dst);
    check(hipGetLastError(), "plan launch");
    check(hipMemcpyAsync(&got, df, sizeof got, hipMemcpyDeviceToHost, st), "DtoH plan");
    check(hipMemcpyAsync(fcb, dfcb, sizeof fcb, hipMemcpyDeviceToHost, st), "DtoH plan");
    check(hipMemcpyAsync(&stv, dst, sizeof stv, hipMemcpyDeviceToHost, st), "DtoH plan");
    check(hipStreamSynchronize(st), "stream wait");
  } catch (...) {
    (void)hipFree(dbuf);
    throw;
  }
  check(hipFree(dbuf), "hipFree");
  for (int t = 0; t < 9; ++t) {
    out[t] = got.h[t];
    out[9 + t] = got.hinv[t];
  }
  for (int t = 0; t < 3; ++t) {
    out[18 + t] = got.origin[t];
    out[21 + t] = got.widths[t];
  }
  out[24] = got.ortho;
  out[25] = got.cell_min;
  out[26] = got.nx;
  out[27] = got.ny;
  out[28] = got.nz;
  out[29] = static_cast<double>(fcb[1]);
  out[30] = got.max_reach;
  out[31] = stv;
  for (int t = 32; t < 36; ++t) out[t] = 0.0;
}

// ------------------------------------------------ fixed-scratch search
//
// A search that allocates nothing: every buffer of the pipeline is carved
// from one file-scope device array, and the frame descriptors are written
// by a kernel from its by-value arguments, so a call makes no hipMalloc,
// hipFree, host-pinned allocation, event or stream. Each GPU holds its own
// copy of the array until the process ends; the kernels receive its
// per-device address (hipGetSymbolAddress), never the host-side name. All
// calls on one GPU share it: calls that overlap on that GPU race.

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up once per device.
void* scratch_for_current_device() {
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* ptr = nullptr;
    check(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress");
    g_scratch_ptr[dev] = ptr;
  }
  return g_scratch_ptr[dev];
}

// Descriptors and bases of F frames sharing one cell: frame f's points,
// outputs and bases follow frame 0's at strides n (points) and ncell.
__global__ void k_noalloc_desc(FrameDesc base, int F, long long n, long long ncell,
                               long long k, FrameDesc* __restrict__ out,
                               long long* __restrict__ fpt, long long* __restrict__ fcb) {
  const long long f = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (f > F) return;
  fpt[f] = f * n;
  fcb[f] = f * ncell;
  if (f == F) return;
  FrameDesc d = base;
  d.xyz = base.xyz + f * n * 3;
  d.out_nn = base.out_nn + f * n * k;
  d.out_d2 = base.out_d2 != nullptr ? base.out_d2 + f * n * k : nullptr;
  d.pt_base = f * n;
  d.cell_base = f * ncell;
  out[f] = d;
}

Sizes noalloc_sizes(int F, long long n, long long ncell, int k) {
  // the exact walk's heaps (k above 128) take at most 8 MiB of the scratch
  return sizes_of(F, static_cast<long long>(F) * n, static_cast<long long>(F) * ncell, k,
                  8ll << 20);
}

}  // namespace

size_t noalloc_bytes(int n_frames, const FrameSpec& frame0) {
  const Sizes z = noalloc_sizes(n_frames, frame0.n, frame0.plan.ncell, frame0.k);
  Bufs B;
  return carve(z, 0, B);
}

void search_noalloc(const FrameSpec& s, int n_frames, hipStream_t st) {
  const long long n = s.n, ncell = s.plan.ncell;
  const Sizes z = noalloc_sizes(n_frames, n, ncell, s.k);
  if (z.ntot >= INT_MAX / 2 || z.ctot >= INT_MAX / 2) {
    throw std::runtime_error("batch exceeds 2^30 points or cells");
  }
  Bufs B;
  const size_t need = carve(z, 0, B);
  if (need > kNoAllocScratchBytes) {
    throw std::runtime_error("the search needs " + std::to_string(need) +
                             " bytes of scratch; the fixed scratch holds " +
                             std::to_string(kNoAllocScratchBytes));
  }
  carve(z, reinterpret_cast<uintptr_t>(scratch_for_current_device()), B);
  FrameDesc d;
  fill_desc(s, d);
  d.pt_base = 0;
  d.cell_base = 0;
  d.kc = fast_bucket(s.k);
  bool any_fast[33] = {};
  if (d.kc) any_fast[d.kc] = true;
  k_noalloc_desc<<<blocks_for(n_frames + 1, 64), 64, 0, st>>>(d, n_frames, n, ncell, s.k,
                                                              B.frames, B.fpt, B.fcb);
  check(hipGetLastError(), "descriptor launch");
  launch_prep(z, B, st, nullptr);
  launch_walk(z, B, walk_args(z, B, kRowWalk), any_fast, d.kc == 0, st, nullptr);
}

namespace {

// Size bucket of a batch for the tuned table: largest register-heap
// bucket, log2 of the point count, log2 of the mean cell occupancy.
std::string bucket_key(const std::vector<FrameSpec>& frames) {
  long long n = 0, c = 0;
  int kc = 0;
  for (const FrameSpec& f : frames) {
    n += f.n;
    c += f.plan.ncell;
    kc = std::max(kc, fast_bucket(f.k));
  }
  int ln = 0, lo = 0;
  while ((1ll << (ln + 1)) <= n) ++ln;
  const long long occ = c > 0 ? n / c : n;
  while ((1ll << (lo + 1)) <= occ) ++lo;
  return "kc" + std::to_string(kc) + ":n2^" + std::to_string(ln) + ":occ2^" +
         std::to_string(lo);
}

}  // namespace

int Engine::dense_min_for(const std::vector<FrameSpec>& frames) const {
  if (dense_min_override_ != 0) return dense_min_override_;
  const std::string key = bucket_key(frames);
  for (const auto& e : tuned_) {
    if (e.first == key) return e.second;
  }
  return kRowWalk;
}

void Engine::capture(const std::vector<FrameSpec>& frames) {
  reset_graph();
  run(frames, true);  // sizes every buffer
  const size_t need = sizeof(FrameDesc) * frames.size() +
                      2 * sizeof(long long) * (frames.size() + 1);
  if (need > graph_staging_bytes_) {
    if (graph_staging_ != nullptr) check(hipHostFree(graph_staging_), "hipHostFree");
    graph_staging_ = nullptr;
    check(hipHostMalloc(&graph_staging_, need, hipHostMallocDefault),
          "hipHostMalloc");
    graph_staging_bytes_ = need;
  }
  hipStream_t st = stream();
  hipGraph_t graph = nullptr;
  check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal),
        "hipStreamBeginCapture");
  capturing_ = true;
  try {
    run(frames, false);
  } catch (...) {
    capturing_ = false;
    (void)hipStreamEndCapture(st, &graph);
    if (graph != nullptr) (void)hipGraphDestroy(graph);
    throw;
  }
  capturing_ = false;
  check(hipStreamEndCapture(st, &graph), "hipStreamEndCapture");
  const hipError_t e = hipGraphInstantiate(&graph_exec_, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  check(e, "hipGraphInstantiate");
}

void Engine::replay(bool wait) {
  if (graph_exec_ == nullptr) throw std::runtime_error("no captured graph");
  check(hipGraphLaunch(graph_exec_, stream()), "hipGraphLaunch");
  if (wait) sync();
}

void Engine::reset_graph() {
  if (graph_exec_ != nullptr) {
    sync();
    (void)hipGraphExecDestroy(graph_exec_);
    graph_exec_ = nullptr;
  }
}

TileSrc Engine::tile_src(long long ctot, int F) {
  TileSrc t;
  t.fr = as<FrameDesc>(frames_.p);
  t.fcb = as<long long>(fcb_.p);
  t.F = F;
  t.count = as<int>(count_.p);
  t.off = as<int>(off_.p);
  t.tile_off = as<int>(tile_off_.p);
  t.tile_cell = as<int>(tile_cell_.p);
  t.tbox = as<double>(tbox_.p);
  t.ncells = ctot;
  return t;
}

Predictor& Engine::predictor() {
  if (pred_ == nullptr) pred_ = std::make_unique<Predictor>();
  return *pred_;
}

long long Engine::sched_collect() {
  const long long nt = tile_count();
  if (nt == 0) return 0;
  grow(labels_, sizeof(float) * static_cast<size_t>(nt));
  tile_labels(as<float>(labels_.p), nt);
  return predictor().collect(tile_src(last_ctot_, last_F_), nt,
                             as<float>(labels_.p), stream());
}

void Engine::sched_dump(unsigned short* x_collect, unsigned short* x_predict,
                        float* caps, int* tile_cell, double* tile_box, int* count,
                        long long nt_cap, long long nc_cap) {
  const long long nt = tile_count();
  if (nt_cap < nt || nc_cap < last_ctot_) {
    throw std::runtime_error("dump buffers shorter than the search's tiles or cells");
  }
  if (nt == 0) return;
  hipStream_t st = stream();
  const TileSrc src = tile_src(last_ctot_, last_F_);
  const size_t xb = sizeof(unsigned short) * static_cast<size_t>(nt) * kFeat;
  void* dx = nullptr;
  void* dp = nullptr;
  void* dc = nullptr;
  check(hipMalloc(&dx, xb), "hipMalloc");
  check(hipMalloc(&dp, xb + sizeof(float) * static_cast<size_t>(nt)), "hipMalloc");
  try {
    predictor().collect_features(src, nt, static_cast<unsigned short*>(dx), st);
    check(hipMemcpyAsync(x_collect, dx, xb, hipMemcpyDeviceToHost, st), "DtoH features");
    if (predictor().trained() && x_predict != nullptr && caps != nullptr) {
      dc = static_cast<char*>(dp) + xb;
      predictor().predict_caps(src, nt, static_cast<float*>(dc), st,
                               static_cast<unsigned short*>(dp));
      check(hipMemcpyAsync(x_predict, dp, xb, hipMemcpyDeviceToHost, st), "DtoH features");
      check(hipMemcpyAsync(caps, dc, sizeof(float) * static_cast<size_t>(nt),
                           hipMemcpyDeviceToHost, st),
            "DtoH caps");
    }
    check(hipMemcpyAsync(tile_cell, tile_cell_.p, sizeof(int) * static_cast<size_t>(nt),
                         hipMemcpyDeviceToHost, st),
          "DtoH tile cells");
    check(hipMemcpyAsync(tile_box, tbox_.p, sizeof(Box) * static_cast<size_t>(nt),
                         hipMemcpyDeviceToHost, st),
          "DtoH tile boxes");
    check(hipMemcpyAsync(count, count_.p, sizeof(int) * static_cast<size_t>(last_ctot_),
                         hipMemcpyDeviceToHost, st),
          "DtoH counts");
    check(hipStreamSynchronize(st), "stream wait");
  } catch (...) {
    (void)hipFree(dx);
    (void)hipFree(dp);
    throw;
  }
  check(hipFree(dx), "hipFree");
  check(hipFree(dp), "hipFree");
}

void Engine::source_labels(float* out, long long n) {
  sync();
  if (last_F_ == 0) return;
  int ns = 0;
  check(hipMemcpy(&ns, as<int>(off_.p) + last_ctot_, sizeof(int), hipMemcpyDeviceToHost),
        "DtoH source count");
  if (n < ns) throw std::runtime_error("label buffer shorter than the search's sources");
  hipStream_t st = stream();
  k_src_labels<<<blocks_for(last_ntot_, 256), 256, 0, st>>>(
      as<FrameDesc>(frames_.p), as<long long>(fcb_.p), last_F_, as<Cand>(cand_.p),
      as<int>(scell_.p), as<int>(off_.p), last_ctot_, out);
  check(hipGetLastError(), "source label launch");
  sync();
}

long long Engine::tile_count() {
  if (last_F_ == 0) return 0;
  sync();
  int t = 0;
  check(hipMemcpy(&t, as<int>(tile_off_.p) + last_ctot_, sizeof(int),
                  hipMemcpyDeviceToHost),
        "DtoH tile count");
  return t;
}

void Engine::tile_labels(float* out, long long n) {
  const long long nt = tile_count();
  if (n < nt) {
    throw std::runtime_error("label buffer shorter than the search's tile count");
  }
  hipStream_t st = stream();
  k_zero_u32<<<blocks_for(nt, 256), 256, 0, st>>>(reinterpret_cast<unsigned*>(out), nt);
  k_tile_labels<<<blocks_for(last_ntot_, 256), 256, 0, st>>>(
      as<FrameDesc>(frames_.p), as<long long>(fcb_.p), last_F_, as<Cand>(cand_.p),
      as<int>(scell_.p), as<int>(off_.p), as<int>(tile_off_.p), last_ctot_,
      reinterpret_cast<int*>(out));
  check(hipGetLastError(), "tile label launch");
  sync();
}

int Engine::autotune(const std::vector<FrameSpec>& frames, float* best_ms) {
  static const int kCands[] = {kRowWalk, kRowWalkPersist, 1, 16, 32, 64, 128, kNeverDense};
  const int keep_override = dense_min_override_;
  const bool keep_timing = timing_;
  int best = kNeverDense;
  float best_t = 0.0f;
  timing_ = true;
  try {
    for (int cand : kCands) {
      dense_min_override_ = cand;
      run(frames, true);  // warm-up at this split
      float t[3];
      for (float& x : t) {
        run(frames, true);
        x = last_timings().total;
      }
      std::sort(t, t + 3);
      if (best_t == 0.0f || t[1] < best_t) {
        best_t = t[1];
        best = cand;
      }
    }
  } catch (...) {
    dense_min_override_ = keep_override;
    timing_ = keep_timing;
    throw;
  }
  dense_min_override_ = keep_override;
  timing_ = keep_timing;
  const std::string key = bucket_key(frames);
  bool set = false;
  for (auto& e : tuned_) {
    if (e.first == key) {
      e.second = best;
      set = true;
    }
  }
  if (!set) tuned_.emplace_back(key, best);
  if (best_ms) *best_ms = best_t;
  return best;
}

}  // namespace lcamd
