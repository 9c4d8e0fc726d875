// Learned shell scheduler on gfx950: tile features, MLP forward / backward
// on bf16 MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulation) and Adam.
//
// One 256-thread block handles 64 samples, 16 per wave. Per wave:
//   z1 = X W1 + b1      16 x 64, K = 128: 4 K-steps x 4 column tiles
//   a1 = relu(z1)       to LDS as BF16 (the next MFMA's A operand)
//   z2 = a1 W2 + b2     16 x 64, K = 64
//   y  = relu(z2) . w3 + b3   (FP32: per-lane partial sums, 16-lane butterfly)
// Backward (same block): dy = w (y - t) / sum(w), dz2 = dy w3 [z2 > 0],
// da1 = dz2 W2^T (MFMA), dz1 = da1 [z1 > 0]; the weight gradients are
// block-level MFMAs over the 64 samples (K = samples): gW2 = a1^T dz2,
// gW1 = X^T dz1. Each block writes its partial gradient; the Adam kernel sums
// the partials in block order (deterministic), updates the FP32 master
// weights and rewrites the BF16 copies the MFMAs read.
//
// MFMA operand layout used here (16x16xK, wave64): lane l supplies A row
// l % 16 and B column l % 16 with 8 K values each, K slot s of lane group
// l / 16 standing for k = 8 (l / 16) + s within the 32-wide K step (A and B
// use the same assignment, which is all the contraction needs); the result
// D[4 (l / 16) + r][l % 16] lands in accumulator element r.
#include "lc_sched.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#pragma clang fp contract(off)

namespace lcamd {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kXS = kFeat + 8;  // LDS row stride of input tiles (BF16)
constexpr int kHS = kHid + 8;   // LDS row stride of hidden tiles (BF16)
constexpr int kPart = kParams + 1;  // partial gradient + loss per block

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
  }
}

unsigned blocks_for(long long n, int threads) {
  const long long b = (n + threads - 1) / threads;
  return static_cast<unsigned>(b < 1 ? 1 : b);
}

// FP32 -> BF16, round to nearest even (NaN stays NaN).
__host__ __device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned u;
  std::memcpy(&u, &f, sizeof u);
  // both results formed, then selected: no divergent branch per value
  const unsigned nan = (u >> 16) | 0x40u;
  const unsigned rne = (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
  return static_cast<unsigned short>((u & 0x7fffffffu) > 0x7f800000u ? nan : rne);
}

__device__ __forceinline__ f32x4 mma(uint4 a, uint4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(
      __builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0,
      0);
}

// 8 consecutive BF16 values of one row (16-byte aligned).
__device__ __forceinline__ uint4 frag_row(const unsigned short* base, int row,
                                          int stride, int col) {
  return *reinterpret_cast<const uint4*>(base + row * stride + col);
}

// 8 BF16 values of one column, rows row0 .. row0 + 7.
__device__ __forceinline__ uint4 frag_col(const unsigned short* base, int row0,
                                          int stride, int col) {
  unsigned v[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const unsigned lo = base[(row0 + 2 * s) * stride + col];
    const unsigned hi = base[(row0 + 2 * s + 1) * stride + col];
    v[s] = lo | (hi << 16);
  }
  return make_uint4(v[0], v[1], v[2], v[3]);
}

struct Net {
  const unsigned short* w1t;  // [kHid][kFeat]: W1 transposed, BF16
  const unsigned short* w2t;  // [kHid out][kHid in], BF16
  const unsigned short* w2;   // [kHid in][kHid out], BF16
  const float* p;             // FP32 master (biases and w3 read from here)
};

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

// Periodic wrap of v = h + d into [0, n), for a home index 0 <= h < n and a
// neighbourhood offset |d| <= 2 (two steps each way cover n = 1), without
// the integer division of v % n.
__device__ __forceinline__ int wrap(int v, int n) {
  v = v < 0 ? v + n : v;
  v = v < 0 ? v + n : v;
  v = v >= n ? v - n : v;
  v = v >= n ? v - n : v;
  return v;
}

// Per-tile context of the feature map.
struct TileCtx {
  bool valid;
  long long cb;
  int nx, ny, nz, hx, hy, hz, k, fill;
  double cell_min;
  float ext[3];  // box edges / cell_min
  float pos[3];  // box centre inside the home cell, fractional
};

__device__ __forceinline__ TileCtx tile_ctx(const TileSrc& S, long long T, long long ntiles) {
  TileCtx c{};
  c.valid = T < ntiles;
  if (!c.valid) return c;
  const int g = S.tile_cell[T];
  const FrameDesc& d = S.fr[find_frame(S.fcb, S.F, g)];
  c.cb = d.cell_base;
  c.nx = d.nx;
  c.ny = d.ny;
  c.nz = d.nz;
  const int cl = g - static_cast<int>(d.cell_base);
  c.hx = cl % d.nx;
  c.hy = (cl / d.nx) % d.ny;
  c.hz = cl / (d.nx * d.ny);
  c.k = d.k;
  c.cell_min = d.cell_min;
  const int cnt = S.off[g + 1] - S.off[g];
  const int tl = static_cast<int>(T) - S.tile_off[g];
  c.fill = min(16, cnt - 16 * tl);
  const double* b = S.tbox + 8 * T;
  c.ext[0] = static_cast<float>((b[1] - b[0]) / c.cell_min);
  c.ext[1] = static_cast<float>((b[3] - b[2]) / c.cell_min);
  c.ext[2] = static_cast<float>((b[5] - b[4]) / c.cell_min);
  const double mx = 0.5 * (b[0] + b[1]) - d.origin[0];
  const double my = 0.5 * (b[2] + b[3]) - d.origin[1];
  const double mz = 0.5 * (b[4] + b[5]) - d.origin[2];
  const double* hi = d.hinv;  // columns, as the fold uses it
  const double s0 = hi[0] * mx + hi[3] * my + hi[6] * mz;
  const double s1 = hi[1] * mx + hi[4] * my + hi[7] * mz;
  const double s2 = hi[2] * mx + hi[5] * my + hi[8] * mz;
  c.pos[0] = static_cast<float>(s0 * d.nx - c.hx);
  c.pos[1] = static_cast<float>(s1 * d.ny - c.hy);
  c.pos[2] = static_cast<float>(s2 * d.nz - c.hz);
  return c;
}

// Occupancy input j < kOcc of a tile: the home cell's 5x5x5 periodic
// neighbourhood, dz -> dy -> dx.
__device__ __forceinline__ long long occ_index(const TileCtx& c, int j) {
  const int dz = j / 25 - 2, dy = (j / 5) % 5 - 2, dx = j % 5 - 2;
  const int x = wrap(c.hx + dx, c.nx);
  const int y = wrap(c.hy + dy, c.ny);
  const int z = wrap(c.hz + dz, c.nz);
  return c.cb + (static_cast<long long>(z) * c.ny + y) * c.nx + x;
}
// 1 + occ >= 1 is a normal float, for which v_log_f32 is log2f itself
// (log2f only adds the rescaling of denormal inputs).
__device__ __forceinline__ float occ_value(int occ) {
  return __builtin_amdgcn_logf(1.0f + static_cast<float>(occ)) * 0.125f;
}
__device__ __forceinline__ float occ_feature(const TileSrc& S, const TileCtx& c, int j) {
  return occ_value(S.count[occ_index(c, j)]);
}

// The 8 tile inputs kOcc .. kOcc + 7, written at constant indices.
__device__ __forceinline__ void tile_extra(const TileCtx& c, float (&e)[8]) {
  e[0] = log2f(c.ext[0] + 1.0f / 256.0f) * 0.125f;
  e[1] = log2f(c.ext[1] + 1.0f / 256.0f) * 0.125f;
  e[2] = log2f(c.ext[2] + 1.0f / 256.0f) * 0.125f;
  e[3] = c.pos[0];
  e[4] = c.pos[1];
  e[5] = c.pos[2];
  e[6] = static_cast<float>(c.fill) * (1.0f / 16.0f);
  e[7] = log2f(static_cast<float>(c.k)) * 0.125f;
}

// Input j of a tile (see lc_sched.hpp for the list).
__device__ __forceinline__ float tile_feature(const TileSrc& S, const TileCtx& c, int j) {
  if (!c.valid) return 0.0f;
  if (j < kOcc) return occ_feature(S, c, j);
  if (j < kOcc + 8) {
    float e[8];
    tile_extra(c, e);
    float v = 0.0f;
#pragma unroll
    for (int t = 0; t < 8; ++t) v = j == kOcc + t ? e[t] : v;
    return v;
  }
  return 0.0f;
}

// Forward pass of one wave over rows r0 .. r0 + 15 of the block's input
// tile sX; writes relu(z1) as BF16 to sA1 (same rows). z1, z2 and y stay in
// registers for the backward pass: z[nt][r] is row r0 + 4 (lane / 16) + r,
// column 16 nt + lane % 16; y[r] is row r0 + 4 (lane / 16) + r.
__device__ __forceinline__ void fwd_wave(const Net& N, const unsigned short* sX,
                                         unsigned short* sA1, int r0, int lane,
                                         f32x4 (&z1)[4], f32x4 (&z2)[4],
                                         float (&y)[4]) {
  const int lr = lane & 15, lg = lane >> 4;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) z1[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < kFeat / 32; ++ks) {
    const uint4 a = frag_row(sX, r0 + lr, kXS, 32 * ks + 8 * lg);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const uint4 b = frag_row(N.w1t, 16 * nt + lr, kFeat, 32 * ks + 8 * lg);
      z1[nt] = mma(a, b, z1[nt]);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float bias = N.p[kOffB1 + 16 * nt + lr];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      z1[nt][r] = z1[nt][r] + bias;
      sA1[(r0 + 4 * lg + r) * kHS + 16 * nt + lr] = f2bf(fmaxf(z1[nt][r], 0.0f));
    }
  }
  __syncthreads();
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) z2[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < kHid / 32; ++ks) {
    const uint4 a = frag_row(sA1, r0 + lr, kHS, 32 * ks + 8 * lg);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const uint4 b = frag_row(N.w2t, 16 * nt + lr, kHid, 32 * ks + 8 * lg);
      z2[nt] = mma(a, b, z2[nt]);
    }
  }
  float s[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const float bias = N.p[kOffB2 + 16 * nt + lr];
    const float w3 = N.p[kOffW3 + 16 * nt + lr];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      z2[nt][r] = z2[nt][r] + bias;
      s[r] = s[r] + fmaxf(z2[nt][r], 0.0f) * w3;
    }
  }
  const float b3 = N.p[kOffB3];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    float v = s[r];
#pragma unroll
    for (int m = 8; m >= 1; m >>= 1) v = v + __shfl_xor(v, m, 16);
    y[r] = v + b3;
  }
}

// Sum over the four 16-lane groups (result in every lane).
__device__ __forceinline__ float sum_groups(float v) {
  v = v + __shfl_xor(v, 16, 64);
  v = v + __shfl_xor(v, 32, 64);
  return v;
}

// Sample i of a pass: (start + i * stride) mod n.
__device__ __forceinline__ long long sample_of(long long i, long long start,
                                               long long stride, long long n) {
  return (start + i * stride) % n;
}

// Forward + backward of 64 samples per block; writes the block's partial
// gradient and loss (kPart floats).
__global__ __launch_bounds__(256) void k_mlp_train(
    Net N, const unsigned short* __restrict__ X, const float* __restrict__ Tg,
    const float* __restrict__ Wg, long long n, long long nrows,
    long long stride, long long start, float inv_wsum,
    float* __restrict__ partial) {
  __shared__ __align__(16) unsigned short sX[kMlpRows * kXS];
  __shared__ __align__(16) unsigned short sA1[kMlpRows * kHS];
  __shared__ __align__(16) unsigned short sD2[kMlpRows * kHS];
  __shared__ __align__(16) unsigned short sD1[kMlpRows * kHS];
  __shared__ float sT[kMlpRows], sW[kMlpRows];
  __shared__ float sRed[4][3][kHid];
  __shared__ float sScal[4][2];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const long long row0 = static_cast<long long>(blockIdx.x) * kMlpRows;
  for (int e = threadIdx.x; e < kMlpRows * (kFeat / 8); e += blockDim.x) {
    const int rr = e / (kFeat / 8), q = e % (kFeat / 8);
    const long long i = row0 + rr;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (i < nrows) {
      const long long sidx = sample_of(i, start, stride, n);
      v = reinterpret_cast<const uint4*>(X + sidx * kFeat)[q];
    }
    *reinterpret_cast<uint4*>(sX + rr * kXS + 8 * q) = v;
  }
  if (threadIdx.x < kMlpRows) {
    const long long i = row0 + threadIdx.x;
    float t = 0.f, w = 0.f;
    if (i < nrows) {
      const long long sidx = sample_of(i, start, stride, n);
      t = Tg[sidx];
      w = Wg[sidx];
    }
    sT[threadIdx.x] = t;
    sW[threadIdx.x] = w;
  }
  __syncthreads();
  const int r0 = 16 * wv;
  f32x4 z1[4], z2[4];
  float y[4];
  fwd_wave(N, sX, sA1, r0, lane, z1, z2, y);

  // output layer
  float dy[4];
  float gb3 = 0.f, loss = 0.f;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const int row = r0 + 4 * lg + r;
    const float e = y[r] - sT[row];
    dy[r] = sW[row] * e * inv_wsum;
    gb3 = gb3 + dy[r];
    loss = loss + 0.5f * sW[row] * e * e * inv_wsum;
  }
  gb3 = sum_groups(gb3);
  loss = sum_groups(loss);
  if (lane == 0) {
    sScal[wv][0] = gb3;
    sScal[wv][1] = loss;
  }
  f32x4 dz2[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int col = 16 * nt + lr;
    const float w3 = N.p[kOffW3 + col];
    float gw3 = 0.f, gb2 = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float a2 = fmaxf(z2[nt][r], 0.0f);
      gw3 = gw3 + dy[r] * a2;
      dz2[nt][r] = z2[nt][r] > 0.0f ? dy[r] * w3 : 0.0f;
      gb2 = gb2 + dz2[nt][r];
      sD2[(r0 + 4 * lg + r) * kHS + col] = f2bf(dz2[nt][r]);
    }
    gw3 = sum_groups(gw3);
    gb2 = sum_groups(gb2);
    if (lg == 0) {
      sRed[wv][0][col] = gw3;
      sRed[wv][1][col] = gb2;
    }
  }
  __syncthreads();
  // da1 = dz2 W2^T: B[k = out][j = in] = W2[in][out]
  f32x4 da1[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) da1[nt] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int ks = 0; ks < kHid / 32; ++ks) {
    const uint4 a = frag_row(sD2, r0 + lr, kHS, 32 * ks + 8 * lg);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const uint4 b = frag_row(N.w2, 16 * nt + lr, kHid, 32 * ks + 8 * lg);
      da1[nt] = mma(a, b, da1[nt]);
    }
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int col = 16 * nt + lr;
    float gb1 = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float dz1 = z1[nt][r] > 0.0f ? da1[nt][r] : 0.0f;
      gb1 = gb1 + dz1;
      sD1[(r0 + 4 * lg + r) * kHS + col] = f2bf(dz1);
    }
    gb1 = sum_groups(gb1);
    if (lg == 0) sRed[wv][2][col] = gb1;
  }
  __syncthreads();
  float* out = partial + static_cast<long long>(blockIdx.x) * kPart;
  // gW2[i][o] = sum over the block's samples of a1[s][i] dz2[s][o]
  {
    const int mt = wv;
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < kMlpRows / 32; ++ks) {
        const uint4 a = frag_col(sA1, 32 * ks + 8 * lg, kHS, 16 * mt + lr);
        const uint4 b = frag_col(sD2, 32 * ks + 8 * lg, kHS, 16 * nt + lr);
        acc = mma(a, b, acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        out[kOffW2 + (16 * mt + 4 * lg + r) * kHid + 16 * nt + lr] = acc[r];
      }
    }
  }
  // gW1[i][h] = sum over the block's samples of x[s][i] dz1[s][h]: the
  // kFeat / 16 input tiles dealt round-robin to the four waves
#pragma unroll 1
  for (int mt = wv; mt < kFeat / 16; mt += 4) {
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int ks = 0; ks < kMlpRows / 32; ++ks) {
        const uint4 a = frag_col(sX, 32 * ks + 8 * lg, kXS, 16 * mt + lr);
        const uint4 b = frag_col(sD1, 32 * ks + 8 * lg, kHS, 16 * nt + lr);
        acc = mma(a, b, acc);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        out[kOffW1 + (16 * mt + 4 * lg + r) * kHid + 16 * nt + lr] = acc[r];
      }
    }
  }
  if (threadIdx.x < kHid) {
    const int col = threadIdx.x;
    float a = 0.f, b = 0.f, c = 0.f;
    for (int w = 0; w < 4; ++w) {
      a = a + sRed[w][0][col];
      b = b + sRed[w][1][col];
      c = c + sRed[w][2][col];
    }
    out[kOffW3 + col] = a;
    out[kOffB2 + col] = b;
    out[kOffB1 + col] = c;
  }
  if (threadIdx.x == 0) {
    float a = 0.f, b = 0.f;
    for (int w = 0; w < 4; ++w) {
      a = a + sScal[w][0];
      b = b + sScal[w][1];
    }
    out[kOffB3] = a;
    out[kParams] = b;
  }
}

__device__ __forceinline__ void cast_param(int p, float w, unsigned short* w1t,
                                           unsigned short* w2t,
                                           unsigned short* w2) {
  if (p < kOffB1) {
    const int in = p / kHid, h = p % kHid;
    w1t[h * kFeat + in] = f2bf(w);
  } else if (p >= kOffW2 && p < kOffB2) {
    const int q = p - kOffW2;
    const int in = q / kHid, o = q % kHid;
    const unsigned short b = f2bf(w);
    w2t[o * kHid + in] = b;
    w2[in * kHid + o] = b;
  }
}

// Sum the blocks' partials (in block order), then one Adam step on the FP32
// master weights and their BF16 copies. c1 = 1 / (1 - beta1^t), c2 =
// 1 / (1 - beta2^t). Thread kParams adds the summed loss to *loss_acc.
__global__ __launch_bounds__(256) void k_mlp_adam(
    const float* __restrict__ partial, int nblk, float* __restrict__ master,
    float* __restrict__ m, float* __restrict__ v, unsigned short* w1t,
    unsigned short* w2t, unsigned short* w2, float lr, float beta1,
    float beta2, float eps, float c1, float c2, int update,
    float* __restrict__ gout, float* __restrict__ loss_acc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p > kParams) return;
  float g = 0.f;
  for (int b = 0; b < nblk; ++b) g = g + partial[static_cast<long long>(b) * kPart + p];
  if (p == kParams) {
    if (loss_acc) *loss_acc = *loss_acc + g;
    return;
  }
  if (gout) gout[p] = g;
  if (!update) return;
  const float mm = beta1 * m[p] + (1.0f - beta1) * g;
  const float vv = beta2 * v[p] + (1.0f - beta2) * (g * g);
  m[p] = mm;
  v[p] = vv;
  const float mh = mm * c1;
  const float vh = vv * c2;
  const float w = master[p] - lr * mh / (sqrtf(vh) + eps);
  master[p] = w;
  cast_param(p, w, w1t, w2t, w2);
}

__global__ __launch_bounds__(256) void k_mlp_cast(const float* __restrict__ master,
                                                  unsigned short* w1t,
                                                  unsigned short* w2t,
                                                  unsigned short* w2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  cast_param(p, master[p], w1t, w2t, w2);
}

__global__ __launch_bounds__(256) void k_fill_f32(float* p, long long n,
                                                  float v) {
  const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (t < n) p[t] = v;
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ x,
                                                 long long n,
                                                 unsigned short* __restrict__ out) {
  const long long t = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  if (t < n) out[t] = f2bf(x[t]);
}

// y for rows i < nrows of the samples (start, stride as in training);
// also copies each row's target and weight when tout is given.
__global__ __launch_bounds__(256) void k_mlp_forward(
    Net N, const unsigned short* __restrict__ X, const float* __restrict__ Tg,
    const float* __restrict__ Wg, long long n, long long nrows,
    long long stride, long long start, float* __restrict__ yout,
    float* __restrict__ tout, float* __restrict__ wout) {
  __shared__ __align__(16) unsigned short sX[kMlpRows * kXS];
  __shared__ __align__(16) unsigned short sA1[kMlpRows * kHS];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const long long row0 = static_cast<long long>(blockIdx.x) * kMlpRows;
  for (int e = threadIdx.x; e < kMlpRows * (kFeat / 8); e += blockDim.x) {
    const int rr = e / (kFeat / 8), q = e % (kFeat / 8);
    const long long i = row0 + rr;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (i < nrows) {
      v = reinterpret_cast<const uint4*>(X + sample_of(i, start, stride, n) * kFeat)[q];
    }
    *reinterpret_cast<uint4*>(sX + rr * kXS + 8 * q) = v;
  }
  __syncthreads();
  f32x4 z1[4], z2[4];
  float y[4];
  fwd_wave(N, sX, sA1, 16 * wv, lane, z1, z2, y);
  if (lr == 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long long i = row0 + 16 * wv + 4 * lg + r;
      if (i < nrows) {
        yout[i] = y[r];
        if (tout) {
          const long long s = sample_of(i, start, stride, n);
          tout[i] = Tg[s];
          wout[i] = Wg[s];
        }
      }
    }
  }
}

// Learned cap of every tile of a search: features straight into LDS, the
// forward pass, out[T] = cell_min * 2^(y + margin). With a list (list_n
// entries, on the device) row i is tile list[i] instead of tile i.
__global__ __launch_bounds__(256) void k_mlp_predict(Net N, TileSrc S,
                                                     float margin,
                                                     float* __restrict__ out,
                                                     unsigned short* __restrict__ xdump,
                                                     const int* __restrict__ list,
                                                     const int* __restrict__ list_n) {
  __shared__ __align__(16) unsigned short sX[kMlpRows * kXS];
  __shared__ __align__(16) unsigned short sA1[kMlpRows * kHS];
  __shared__ float sCM[kMlpRows];
  const long long ntiles = S.tile_off[S.ncells];
  const long long nrows = list != nullptr ? *list_n : ntiles;
  const long long t0 = static_cast<long long>(blockIdx.x) * kMlpRows;
  if (t0 >= nrows) return;
  // the tile of row i (ntiles, which no tile is, past the last row)
  auto tile_of = [&](long long i) -> long long {
    if (i >= nrows) return ntiles;
    return list != nullptr ? static_cast<long long>(list[i]) : i;
  };
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  {
    // four threads per tile, a quarter of the inputs each
    const int rr = threadIdx.x >> 2, part = threadIdx.x & 3;
    const TileCtx c = tile_ctx(S, tile_of(t0 + rr), ntiles);
    if (part == 0) sCM[rr] = static_cast<float>(c.cell_min);
    // occupancy inputs by (dz, dy) rows of five: thread `part` takes rows
    // part, part + 4, ...; the five dx offsets wrap once per thread, and
    // every count load is unconditional (cell 0 stands in for a row a
    // thread does not have), so all of them are in flight at once
    int wx[5];
#pragma unroll
    for (int d = 0; d < 5; ++d) wx[d] = wrap(c.hx + d - 2, c.nx);
    int occ[7][5];
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const int r = part + 4 * i;
      const bool rv = c.valid && r < 25;
      const int z = wrap(c.hz + r / 5 - 2, c.nz);
      const int y = wrap(c.hy + r % 5 - 2, c.ny);
      const long long base = c.cb + (static_cast<long long>(z) * c.ny + y) * c.nx;
#pragma unroll
      for (int d = 0; d < 5; ++d) occ[i][d] = S.count[rv ? base + wx[d] : 0];
    }
#pragma unroll
    for (int i = 0; i < 7; ++i) {
      const int r = part + 4 * i;
      if (r < 25) {
#pragma unroll
        for (int d = 0; d < 5; ++d) {
          sX[rr * kXS + 5 * r + d] = f2bf(c.valid ? occ_value(occ[i][d]) : 0.0f);
        }
      }
    }
    static_assert(kOcc == 125 && kOcc + 8 <= kFeat, "5x5x5 occupancy, then the tile inputs");
    if (part == 3) {
      float e[8];
      tile_extra(c, e);
#pragma unroll
      for (int t = 0; t < 8; ++t) sX[rr * kXS + kOcc + t] = f2bf(c.valid ? e[t] : 0.0f);
#pragma unroll
      for (int j = kOcc + 8; j < kFeat; ++j) sX[rr * kXS + j] = 0;
    }
  }
  __syncthreads();
  if (xdump != nullptr) {
    // diagnostics: the inputs exactly as the forward pass reads them
    for (int e = threadIdx.x; e < kMlpRows * kFeat; e += blockDim.x) {
      const long long T = tile_of(t0 + e / kFeat);
      if (T < ntiles) xdump[T * kFeat + e % kFeat] = sX[(e / kFeat) * kXS + e % kFeat];
    }
  }
  f32x4 z1[4], z2[4];
  float y[4];
  fwd_wave(N, sX, sA1, 16 * wv, lane, z1, z2, y);
  if (lr == 0) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int row = 16 * wv + 4 * lg + r;
      const long long T = tile_of(t0 + row);
      if (T < ntiles) out[T] = sCM[row] * exp2f(y[r] + margin);
    }
  }
}

// Training samples of the tiles of a finished search: one thread per 8
// inputs of a tile; target log2(label / cell_min), weight 1 (0 for an
// infinite label).
__global__ __launch_bounds__(256) void k_sched_collect(
    TileSrc S, long long ntiles, const float* __restrict__ labels,
    unsigned short* __restrict__ X, float* __restrict__ Tg,
    float* __restrict__ Wg) {
  const long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x;
  const long long T = e / (kFeat / 8);
  const int q = static_cast<int>(e % (kFeat / 8));
  if (T >= ntiles) return;
  const TileCtx c = tile_ctx(S, T, ntiles);
  unsigned v[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    const unsigned lo = f2bf(tile_feature(S, c, 8 * q + 2 * s));
    const unsigned hi = f2bf(tile_feature(S, c, 8 * q + 2 * s + 1));
    v[s] = lo | (hi << 16);
  }
  reinterpret_cast<uint4*>(X + T * kFeat)[q] = make_uint4(v[0], v[1], v[2], v[3]);
  if (q == 0 && labels != nullptr) {
    const float lab = labels[T];
    const bool fin = lab < __builtin_inff();
    Tg[T] = fin ? log2f(lab / static_cast<float>(c.cell_min)) : 0.0f;
    Wg[T] = fin ? 1.0f : 0.0f;
  }
}

}  // namespace

Predictor::~Predictor() {
  Buf* all[] = {&master_, &m_, &v_, &w1t_, &w2t_, &w2_, &partial_, &loss_,
                &gout_, &X_, &T_, &W_, &tmpX_, &tmpT_, &tmpW_, &tmpY_};
  for (Buf* b : all) {
    if (b->p != nullptr) (void)hipFree(b->p);
  }
}

void Predictor::grow(Buf& b, size_t bytes, bool keep) {
  if (bytes <= b.bytes && b.p != nullptr) return;
  void* np = nullptr;
  // kept buffers (the training set) grow geometrically
  const size_t want = std::max<size_t>(keep ? std::max(bytes, b.bytes + b.bytes / 2) : bytes, 256);
  check(hipMalloc(&np, want), "hipMalloc");
  if (keep && b.p != nullptr && b.bytes > 0) {
    check(hipMemcpy(np, b.p, b.bytes, hipMemcpyDeviceToDevice), "DtoD grow");
  }
  if (b.p != nullptr) check(hipFree(b.p), "hipFree");
  b.p = np;
  b.bytes = want;
}

void Predictor::ensure_params() {
  if (params_ready_) return;
  grow(master_, sizeof(float) * kParams);
  grow(m_, sizeof(float) * kParams);
  grow(v_, sizeof(float) * kParams);
  grow(w1t_, sizeof(unsigned short) * kFeat * kHid);
  grow(w2t_, sizeof(unsigned short) * kHid * kHid);
  grow(w2_, sizeof(unsigned short) * kHid * kHid);
  grow(loss_, sizeof(float) * 4);
  grow(gout_, sizeof(float) * kParams);
  params_ready_ = true;
}

void Predictor::upload_params(const float* host, hipStream_t st) {
  ensure_params();
  check(hipMemcpyAsync(master_.p, host, sizeof(float) * kParams,
                       hipMemcpyHostToDevice, st),
        "HtoD params");
  k_mlp_cast<<<blocks_for(kParams, 256), 256, 0, st>>>(
      static_cast<float*>(master_.p), static_cast<unsigned short*>(w1t_.p),
      static_cast<unsigned short*>(w2t_.p), static_cast<unsigned short*>(w2_.p));
  check(hipGetLastError(), "cast launch");
  check(hipStreamSynchronize(st), "stream wait");
}

void Predictor::get_params(float* out) const {
  if (!params_ready_) throw std::runtime_error("the predictor has no parameters");
  check(hipMemcpy(out, master_.p, sizeof(float) * kParams, hipMemcpyDeviceToHost),
        "DtoH params");
}

void Predictor::set_params(const float* in, double margin, hipStream_t st) {
  upload_params(in, st);
  margin_ = margin;
  trained_ = true;
}

void Predictor::step(const unsigned short* X, const float* t, const float* w,
                     long long n, long long nrows, long long stride,
                     long long start, double lr, int t_step, float inv_wsum,
                     hipStream_t st, bool update, float* grad_out) {
  const int nblk = static_cast<int>((nrows + kMlpRows - 1) / kMlpRows);
  grow(partial_, sizeof(float) * kPart * static_cast<size_t>(nblk));
  Net N{static_cast<const unsigned short*>(w1t_.p),
        static_cast<const unsigned short*>(w2t_.p),
        static_cast<const unsigned short*>(w2_.p),
        static_cast<const float*>(master_.p)};
  k_mlp_train<<<nblk, 256, 0, st>>>(N, X, t, w, n, nrows, stride, start,
                                    inv_wsum, static_cast<float*>(partial_.p));
  constexpr double kB1 = 0.9, kB2 = 0.999;
  const double c1 = 1.0 / (1.0 - std::pow(kB1, t_step));
  const double c2 = 1.0 / (1.0 - std::pow(kB2, t_step));
  k_mlp_adam<<<blocks_for(kParams + 1, 256), 256, 0, st>>>(
      static_cast<const float*>(partial_.p), nblk, static_cast<float*>(master_.p),
      static_cast<float*>(m_.p), static_cast<float*>(v_.p),
      static_cast<unsigned short*>(w1t_.p), static_cast<unsigned short*>(w2t_.p),
      static_cast<unsigned short*>(w2_.p), static_cast<float>(lr),
      static_cast<float>(kB1), static_cast<float>(kB2), 1e-8f,
      static_cast<float>(c1), static_cast<float>(c2), update ? 1 : 0, grad_out,
      static_cast<float*>(loss_.p));
  check(hipGetLastError(), "mlp step launch");
}

void Predictor::forward_host(const float* x, long long n, float* y,
                             hipStream_t st) {
  if (!params_ready_) throw std::runtime_error("the predictor has no parameters");
  if (n <= 0) return;
  grow(tmpT_, sizeof(float) * static_cast<size_t>(n * kFeat));
  grow(tmpX_, sizeof(unsigned short) * static_cast<size_t>(n * kFeat));
  grow(tmpY_, sizeof(float) * static_cast<size_t>(n));
  check(hipMemcpyAsync(tmpT_.p, x, sizeof(float) * n * kFeat, hipMemcpyHostToDevice, st),
        "HtoD x");
  k_to_bf16<<<blocks_for(n * kFeat, 256), 256, 0, st>>>(
      static_cast<const float*>(tmpT_.p), n * kFeat, static_cast<unsigned short*>(tmpX_.p));
  Net N{static_cast<const unsigned short*>(w1t_.p),
        static_cast<const unsigned short*>(w2t_.p),
        static_cast<const unsigned short*>(w2_.p),
        static_cast<const float*>(master_.p)};
  k_mlp_forward<<<blocks_for(n, kMlpRows), 256, 0, st>>>(
      N, static_cast<const unsigned short*>(tmpX_.p), nullptr, nullptr, n, n, 1,
      0, static_cast<float*>(tmpY_.p), nullptr, nullptr);
  check(hipGetLastError(), "mlp forward launch");
  check(hipMemcpyAsync(y, tmpY_.p, sizeof(float) * n, hipMemcpyDeviceToHost, st),
        "DtoH y");
  check(hipStreamSynchronize(st), "stream wait");
}

double Predictor::grad_host(const float* x, const float* t, const float* w,
                            long long n, float* grad, hipStream_t st) {
  if (!params_ready_) throw std::runtime_error("the predictor has no parameters");
  if (n <= 0) throw std::runtime_error("no samples");
  grow(tmpY_, sizeof(float) * static_cast<size_t>(n * kFeat));
  grow(tmpX_, sizeof(unsigned short) * static_cast<size_t>(n * kFeat));
  grow(tmpT_, sizeof(float) * static_cast<size_t>(n));
  grow(tmpW_, sizeof(float) * static_cast<size_t>(n));
  check(hipMemcpyAsync(tmpY_.p, x, sizeof(float) * n * kFeat, hipMemcpyHostToDevice, st),
        "HtoD x");
  k_to_bf16<<<blocks_for(n * kFeat, 256), 256, 0, st>>>(
      static_cast<const float*>(tmpY_.p), n * kFeat, static_cast<unsigned short*>(tmpX_.p));
  check(hipMemcpyAsync(tmpT_.p, t, sizeof(float) * n, hipMemcpyHostToDevice, st), "HtoD t");
  check(hipMemcpyAsync(tmpW_.p, w, sizeof(float) * n, hipMemcpyHostToDevice, st), "HtoD w");
  double ws = 0.0;
  for (long long i = 0; i < n; ++i) ws += w[i];
  if (!(ws > 0.0)) throw std::runtime_error("sample weights sum to zero");
  k_fill_f32<<<1, 64, 0, st>>>(static_cast<float*>(loss_.p), 1, 0.0f);
  step(static_cast<const unsigned short*>(tmpX_.p), static_cast<const float*>(tmpT_.p),
       static_cast<const float*>(tmpW_.p), n, n, 1, 0, 0.0, 1,
       static_cast<float>(1.0 / ws), st, false, static_cast<float*>(gout_.p));
  float loss = 0.f;
  check(hipMemcpyAsync(grad, gout_.p, sizeof(float) * kParams, hipMemcpyDeviceToHost, st),
        "DtoH grad");
  check(hipMemcpyAsync(&loss, loss_.p, sizeof(float), hipMemcpyDeviceToHost, st),
        "DtoH loss");
  check(hipStreamSynchronize(st), "stream wait");
  return loss;
}

double Predictor::adam_host(const float* x, const float* t, const float* w,
                            long long n, int steps, double lr, hipStream_t st) {
  if (!params_ready_) throw std::runtime_error("the predictor has no parameters");
  if (n <= 0 || steps <= 0) throw std::runtime_error("no samples or steps");
  grow(tmpY_, sizeof(float) * static_cast<size_t>(n * kFeat));
  grow(tmpX_, sizeof(unsigned short) * static_cast<size_t>(n * kFeat));
  grow(tmpT_, sizeof(float) * static_cast<size_t>(n));
  grow(tmpW_, sizeof(float) * static_cast<size_t>(n));
  check(hipMemcpyAsync(tmpY_.p, x, sizeof(float) * n * kFeat, hipMemcpyHostToDevice, st),
        "HtoD x");
  k_to_bf16<<<blocks_for(n * kFeat, 256), 256, 0, st>>>(
      static_cast<const float*>(tmpY_.p), n * kFeat, static_cast<unsigned short*>(tmpX_.p));
  check(hipMemcpyAsync(tmpT_.p, t, sizeof(float) * n, hipMemcpyHostToDevice, st), "HtoD t");
  check(hipMemcpyAsync(tmpW_.p, w, sizeof(float) * n, hipMemcpyHostToDevice, st), "HtoD w");
  double ws = 0.0;
  for (long long i = 0; i < n; ++i) ws += w[i];
  if (!(ws > 0.0)) throw std::runtime_error("sample weights sum to zero");
  k_fill_f32<<<blocks_for(kParams, 256), 256, 0, st>>>(static_cast<float*>(m_.p), kParams, 0.f);
  k_fill_f32<<<blocks_for(kParams, 256), 256, 0, st>>>(static_cast<float*>(v_.p), kParams, 0.f);
  float loss = 0.f;
  for (int s = 1; s <= steps; ++s) {
    k_fill_f32<<<1, 64, 0, st>>>(static_cast<float*>(loss_.p), 1, 0.0f);
    step(static_cast<const unsigned short*>(tmpX_.p), static_cast<const float*>(tmpT_.p),
         static_cast<const float*>(tmpW_.p), n, n, 1, 0, lr, s,
         static_cast<float>(1.0 / ws), st, true, nullptr);
  }
  check(hipMemcpyAsync(&loss, loss_.p, sizeof(float), hipMemcpyDeviceToHost, st),
        "DtoH loss");
  check(hipStreamSynchronize(st), "stream wait");
  return loss;
}

void Predictor::clear_samples() { n_samples_ = 0; }

long long Predictor::collect(const TileSrc& src, long long ntiles,
                             const float* labels, hipStream_t st) {
  if (ntiles <= 0) return 0;
  const long long n = n_samples_ + ntiles;
  grow(X_, sizeof(unsigned short) * static_cast<size_t>(n * kFeat), true);
  grow(T_, sizeof(float) * static_cast<size_t>(n), true);
  grow(W_, sizeof(float) * static_cast<size_t>(n), true);
  k_sched_collect<<<blocks_for(ntiles * (kFeat / 8), 256), 256, 0, st>>>(
      src, ntiles, labels,
      static_cast<unsigned short*>(X_.p) + n_samples_ * kFeat,
      static_cast<float*>(T_.p) + n_samples_, static_cast<float*>(W_.p) + n_samples_);
  check(hipGetLastError(), "collect launch");
  check(hipStreamSynchronize(st), "stream wait");
  n_samples_ = n;
  return ntiles;
}

void Predictor::collect_features(const TileSrc& src, long long ntiles,
                                 unsigned short* x, hipStream_t st) {
  if (ntiles <= 0) return;
  k_sched_collect<<<blocks_for(ntiles * (kFeat / 8), 256), 256, 0, st>>>(
      src, ntiles, nullptr, x, nullptr, nullptr);
  check(hipGetLastError(), "collect launch");
}

double Predictor::train(int epochs, int batch, double lr, unsigned seed,
                        double quantile, hipStream_t st) {
  if (n_samples_ < kMlpRows) throw std::runtime_error("too few training samples");
  if (epochs < 1 || batch < kMlpRows || batch % kMlpRows != 0) {
    throw std::runtime_error("epochs must be >= 1 and batch a multiple of 64");
  }
  if (!(quantile > 0.0 && quantile < 1.0)) {
    throw std::runtime_error("quantile must lie in (0, 1)");
  }
  ensure_params();
  // He-uniform initialisation from the seed
  std::mt19937_64 gen(seed);
  std::vector<float> p0(kParams, 0.0f);
  auto uni = [&](double a) {
    return static_cast<float>((std::uniform_real_distribution<double>(-a, a))(gen));
  };
  for (int i = 0; i < kFeat * kHid; ++i) p0[kOffW1 + i] = uni(std::sqrt(6.0 / kFeat));
  for (int i = 0; i < kHid * kHid; ++i) p0[kOffW2 + i] = uni(std::sqrt(6.0 / kHid));
  for (int i = 0; i < kHid; ++i) p0[kOffW3 + i] = uni(std::sqrt(6.0 / kHid));
  upload_params(p0.data(), st);
  k_fill_f32<<<blocks_for(kParams, 256), 256, 0, st>>>(static_cast<float*>(m_.p), kParams, 0.f);
  k_fill_f32<<<blocks_for(kParams, 256), 256, 0, st>>>(static_cast<float*>(v_.p), kParams, 0.f);
  const long long n = n_samples_;
  const long long b = std::min<long long>(batch, n - n % kMlpRows);
  const long long nb = n / b;
  double last = 0.0;
  int t_step = 0;
  for (int ep = 0; ep < epochs; ++ep) {
    // a fresh affine permutation of the samples per epoch
    long long stride = 1;
    for (;;) {
      stride = static_cast<long long>(gen() % static_cast<unsigned long long>(n));
      if (stride > 0 && std::gcd(stride, n) == 1) break;
    }
    const long long start = static_cast<long long>(gen() % static_cast<unsigned long long>(n));
    k_fill_f32<<<1, 64, 0, st>>>(static_cast<float*>(loss_.p), 1, 0.0f);
    for (long long bi = 0; bi < nb; ++bi) {
      const long long s0 = (start + (bi * b % n) * stride) % n;
      step(static_cast<const unsigned short*>(X_.p), static_cast<const float*>(T_.p),
           static_cast<const float*>(W_.p), n, b, stride, s0, lr, ++t_step,
           static_cast<float>(1.0 / static_cast<double>(b)), st, true, nullptr);
    }
    float acc = 0.f;
    check(hipMemcpyAsync(&acc, loss_.p, sizeof(float), hipMemcpyDeviceToHost, st),
          "DtoH loss");
    check(hipStreamSynchronize(st), "stream wait");
    last = acc / static_cast<double>(nb);
  }
  // margin: the quantile of (target - prediction) over up to 2^20 samples
  const long long nsub = std::min<long long>(n, 1 << 20);
  const long long sstride = n / nsub;
  grow(tmpY_, sizeof(float) * static_cast<size_t>(nsub));
  grow(tmpT_, sizeof(float) * static_cast<size_t>(nsub));
  grow(tmpW_, sizeof(float) * static_cast<size_t>(nsub));
  Net N{static_cast<const unsigned short*>(w1t_.p),
        static_cast<const unsigned short*>(w2t_.p),
        static_cast<const unsigned short*>(w2_.p),
        static_cast<const float*>(master_.p)};
  k_mlp_forward<<<blocks_for(nsub, kMlpRows), 256, 0, st>>>(
      N, static_cast<const unsigned short*>(X_.p), static_cast<const float*>(T_.p),
      static_cast<const float*>(W_.p), n, nsub, sstride, 0,
      static_cast<float*>(tmpY_.p), static_cast<float*>(tmpT_.p),
      static_cast<float*>(tmpW_.p));
  check(hipGetLastError(), "mlp forward launch");
  std::vector<float> y(nsub), t(nsub), w(nsub);
  check(hipMemcpyAsync(y.data(), tmpY_.p, sizeof(float) * nsub, hipMemcpyDeviceToHost, st), "DtoH");
  check(hipMemcpyAsync(t.data(), tmpT_.p, sizeof(float) * nsub, hipMemcpyDeviceToHost, st), "DtoH");
  check(hipMemcpyAsync(w.data(), tmpW_.p, sizeof(float) * nsub, hipMemcpyDeviceToHost, st), "DtoH");
  check(hipStreamSynchronize(st), "stream wait");
  std::vector<float> res;
  res.reserve(nsub);
  for (long long i = 0; i < nsub; ++i) {
    if (w[i] > 0.f) res.push_back(t[i] - y[i]);
  }
  if (res.empty()) throw std::runtime_error("no finite training labels");
  const size_t qi = std::min(res.size() - 1,
                             static_cast<size_t>(quantile * static_cast<double>(res.size())));
  std::nth_element(res.begin(), res.begin() + static_cast<long long>(qi), res.end());
  margin_ = res[qi];
  trained_ = true;
  return last;
}

void Predictor::predict_caps(const TileSrc& src, long long tcap, float* out,
                             hipStream_t st, unsigned short* xdump, const int* list,
                             const int* list_n) {
  if (!trained_) throw std::runtime_error("the predictor is not trained");
  Net N{static_cast<const unsigned short*>(w1t_.p),
        static_cast<const unsigned short*>(w2t_.p),
        static_cast<const unsigned short*>(w2_.p),
        static_cast<const float*>(master_.p)};
  k_mlp_predict<<<blocks_for(tcap, kMlpRows), 256, 0, st>>>(
      N, src, static_cast<float>(margin_), out, xdump, list, list_n);
  check(hipGetLastError(), "mlp predict launch");
}

}  // namespace lcamd
