// Periodic-cell arithmetic shared by host planning and device kernels.
//
// Every expression follows linkcell 0.3.3's Rust CPU path in source
// order (src/cell.rs, src/knearest.rs). The translation units that include
// this header are compiled with -ffp-contract=off, so each + - * / is one
// correctly rounded IEEE-754 double operation, as rustc emits them for the
// reference (no FMA). That is what makes folded coordinates, lattice
// shifts and squared distances bit-identical to the reference CPU walk.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#pragma clang fp contract(off)

#define LC_HD __host__ __device__ __forceinline__

namespace lcamd {

constexpr long long kMaxCells = 16777216LL;  // src/knearest.rs:20
constexpr int kEmptyJ = 0x7fffffff;          // heap slot not yet filled

// Error codes. Messages follow src/error.rs:27-33 and the GPU waist
// (src/gpu/knearest.cpp:711-756, src/gpu/capi.cpp:477-480).
enum Status : int {
  kOk = 0,
  kZeroK,
  kBadBox,
  kEmpty,
  kBufferSize,
  kOverflow,
  kMaskLen,
  kTooManyCells,
};

// A periodic parallelepiped after Cell::from_vectors (src/cell.rs:97-123).
// h[col*3 + row]: columns a, b, c; hinv uses the same layout as the
// reference's `hinv[col][row]`.
struct CellGeom {
  double h[9];
  double hinv[9];
  double origin[3];
  double widths[3];
  int ortho;
};

// Mesh plan (src/knearest.rs:377-397, :430).
struct MeshPlan {
  int nx, ny, nz;
  long long ncell;
  double cell_min;
  int max_reach;
};

// Rust `x as i32`: saturating, NaN -> 0.
LC_HD int sat_i32(double x) {
  if (x != x) return 0;
  if (x >= 2147483647.0) return 2147483647;
  if (x <= -2147483648.0) return -2147483647 - 1;
  return static_cast<int>(x);
}

// src/cell.rs:335-341 (H stored by columns).
LC_HD void mul_cols(const double* m, double v0, double v1, double v2,
                    double& r0, double& r1, double& r2) {
  r0 = m[0] * v0 + m[3] * v1 + m[6] * v2;
  r1 = m[1] * v0 + m[4] * v1 + m[7] * v2;
  r2 = m[2] * v0 + m[5] * v1 + m[8] * v2;
}

// src/cell.rs:291-298
LC_HD double wrap01(double s) {
  s -= floor(s);
  return s >= 1.0 ? 0.0 : s;
}

// Cell::fractional (src/cell.rs:176-194) then Cell::cartesian
// (src/cell.rs:200-207).
LC_HD void fold_point(const CellGeom& g, double rx, double ry, double rz,
                      double& s0, double& s1, double& s2, double& fx,
                      double& fy, double& fz) {
  if (g.ortho) {
    s0 = wrap01((rx - g.origin[0]) / g.widths[0]);
    s1 = wrap01((ry - g.origin[1]) / g.widths[1]);
    s2 = wrap01((rz - g.origin[2]) / g.widths[2]);
  } else {
    const double d0 = rx - g.origin[0];
    const double d1 = ry - g.origin[1];
    const double d2 = rz - g.origin[2];
    mul_cols(g.hinv, d0, d1, d2, s0, s1, s2);
    s0 = wrap01(s0);
    s1 = wrap01(s1);
    s2 = wrap01(s2);
  }
  double t0, t1, t2;
  mul_cols(g.h, s0, s1, s2, t0, t1, t2);
  fx = t0 + g.origin[0];
  fy = t1 + g.origin[1];
  fz = t2 + g.origin[2];
}

// bin = clamp(s * n as i32, 0, n - 1) (src/knearest.rs:408-412)
LC_HD int bin_axis(double s, int n) {
  int v = sat_i32(s * static_cast<double>(n));
  if (v < 0) v = 0;
  if (v > n - 1) v = n - 1;
  return v;
}

// Cell::lattice_shift (src/cell.rs:227-247)
LC_HD void lattice_shift(const CellGeom& g, int na, int nb, int nc,
                         double& sx, double& sy, double& sz) {
  if (g.ortho) {
    sx = static_cast<double>(na) * g.widths[0];
    sy = static_cast<double>(nb) * g.widths[1];
    sz = static_cast<double>(nc) * g.widths[2];
  } else {
    const double fa = static_cast<double>(na);
    const double fb = static_cast<double>(nb);
    const double fc = static_cast<double>(nc);
    sx = fa * g.h[0] + fb * g.h[3] + fc * g.h[6];
    sy = fa * g.h[1] + fb * g.h[4] + fc * g.h[7];
    sz = fa * g.h[2] + fb * g.h[5] + fc * g.h[8];
  }
}

// Cell::dist2_shifted (src/cell.rs:269-274): (q + shift) - p, summed
// left to right.
LC_HD double dist2_shifted(double px, double py, double pz, double qx,
                           double qy, double qz, double sx, double sy,
                           double sz) {
  const double dx = qx + sx - px;
  const double dy = qy + sy - py;
  const double dz = qz + sz - pz;
  return dx * dx + dy * dy + dz * dz;
}

LC_HD int rem_euclid(int a, int n) {
  int r = a % n;
  return r < 0 ? r + n : r;
}

LC_HD int div_euclid(int a, int n) {
  int q = a / n;
  if (a % n < 0) q -= 1;
  return q;
}

LC_HD double rust_min(double a, double b) { return fmin(a, b); }
LC_HD double rust_max(double a, double b) { return fmax(a, b); }

LC_HD double norm3(double x, double y, double z) {
  return sqrt(x * x + y * y + z * z);
}

// Cell::from_vectors + invert_columns + is_axis_aligned
// (src/cell.rs:97-123, :322-331, :343-389). a, b, c, o are the lc_cell
// rows. Returns kOk or kBadBox.
LC_HD int cell_from_vectors(const double* a, const double* b, const double* c,
                            const double* o, CellGeom& g) {
  const double det = a[0] * (b[1] * c[2] - b[2] * c[1]) -
                     a[1] * (b[0] * c[2] - b[2] * c[0]) +
                     a[2] * (b[0] * c[1] - b[1] * c[0]);
  if (!__builtin_isfinite(det) || fabs(det) < 1e-18) return kBadBox;
  const double invdet = 1.0 / det;
  for (int i = 0; i < 3; ++i) {
    g.h[0 + i] = a[i];
    g.h[3 + i] = b[i];
    g.h[6 + i] = c[i];
    g.origin[i] = o[i];
  }
  g.hinv[0] = (b[1] * c[2] - b[2] * c[1]) * invdet;
  g.hinv[1] = (a[2] * c[1] - a[1] * c[2]) * invdet;
  g.hinv[2] = (a[1] * b[2] - a[2] * b[1]) * invdet;
  g.hinv[3] = (b[2] * c[0] - b[0] * c[2]) * invdet;
  g.hinv[4] = (a[0] * c[2] - a[2] * c[0]) * invdet;
  g.hinv[5] = (a[2] * b[0] - a[0] * b[2]) * invdet;
  g.hinv[6] = (b[0] * c[1] - b[1] * c[0]) * invdet;
  g.hinv[7] = (a[1] * c[0] - a[0] * c[1]) * invdet;
  g.hinv[8] = (a[0] * b[1] - a[1] * b[0]) * invdet;
  // cross products b x c, c x a, a x b (src/cell.rs:343-349)
  const double bc0 = b[1] * c[2] - b[2] * c[1];
  const double bc1 = b[2] * c[0] - b[0] * c[2];
  const double bc2 = b[0] * c[1] - b[1] * c[0];
  const double ca0 = c[1] * a[2] - c[2] * a[1];
  const double ca1 = c[2] * a[0] - c[0] * a[2];
  const double ca2 = c[0] * a[1] - c[1] * a[0];
  const double ab0 = a[1] * b[2] - a[2] * b[1];
  const double ab1 = a[2] * b[0] - a[0] * b[2];
  const double ab2 = a[0] * b[1] - a[1] * b[0];
  const double adet = fabs(det);
  const double wa = adet / norm3(bc0, bc1, bc2);
  const double wb = adet / norm3(ca0, ca1, ca2);
  const double wc = adet / norm3(ab0, ab1, ab2);
  if (!(wa > 0.0 && wb > 0.0 && wc > 0.0)) return kBadBox;
  g.widths[0] = wa;
  g.widths[1] = wb;
  g.widths[2] = wc;
  const double scale =
      rust_max(norm3(a[0], a[1], a[2]) + norm3(b[0], b[1], b[2]) +
                   norm3(c[0], c[1], c[2]),
               1.0);
  const double tol = 1e-12 * scale;
  g.ortho = fabs(a[1]) <= tol && fabs(a[2]) <= tol && fabs(b[0]) <= tol &&
            fabs(b[2]) <= tol && fabs(c[0]) <= tol && fabs(c[1]) <= tol;
  return kOk;
}

// bins_1d (src/knearest.rs:42-48)
LC_HD int bins_1d(double width, double edge, int& out) {
  const double n = rust_max(floor(width / edge), 1.0);
  if (!__builtin_isfinite(n) || n > 1000000.0) return kTooManyCells;
  out = sat_i32(n);
  return kOk;
}

// Mesh plan (src/knearest.rs:377-397, :430). `hint` is the Option the
// search receives: NaN or <= 0 selects 3.0.
LC_HD int plan_mesh(const CellGeom& g, double hint, MeshPlan& p) {
  double edge = hint;
  if (!__builtin_isfinite(edge) || edge <= 0.0) edge = 3.0;
  edge = rust_min(rust_min(rust_min(edge, g.widths[0]), g.widths[1]),
                  g.widths[2]);
  int nx, ny, nz, rc;
  if ((rc = bins_1d(g.widths[0], edge, nx)) != kOk) return rc;
  if ((rc = bins_1d(g.widths[1], edge, ny)) != kOk) return rc;
  if ((rc = bins_1d(g.widths[2], edge, nz)) != kOk) return rc;
  const long long nc = static_cast<long long>(nx) * ny * nz;
  if (!(nc > 0 && nc <= kMaxCells)) return kTooManyCells;
  p.nx = nx;
  p.ny = ny;
  p.nz = nz;
  p.ncell = nc;
  p.cell_min =
      rust_min(rust_min(g.widths[0] / static_cast<double>(nx),
                        g.widths[1] / static_cast<double>(ny)),
               g.widths[2] / static_cast<double>(nz));
  int m = nx > ny ? nx : ny;
  m = m > nz ? m : nz;
  p.max_reach = m / 2 + 1;
  return kOk;
}

// Largest reach at which no two stencil offsets map to the same cell on
// any axis (2r + 1 <= n_a). Up to it, every candidate index is visited at
// most once per source, so the per-index min of KHeap::push
// (src/knearest.rs:107-114) never triggers.
LC_HD int safe_reach(const MeshPlan& p) {
  int r = (p.nx - 1) / 2;
  r = (p.ny - 1) / 2 < r ? (p.ny - 1) / 2 : r;
  r = (p.nz - 1) / 2 < r ? (p.nz - 1) / 2 : r;
  return r;
}

}  // namespace lcamd
