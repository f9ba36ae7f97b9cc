// utilities.h -- LPM-CPP's lattice generators (include/utilities.h:15-330) for
// drivers built against the GPU adapter: the Euler-angle rotation matrix and
// the square, hexagonal, simple-cubic and FCC lattices.  Each generator fills a
// padded block (half-width = the box diagonal) in the reference's order,
// rotates every point by R and keeps those inside the box; the coordinates are
// formed by the same expressions, so an identity rotation reproduces the
// reference's points bit for bit.  The reference rotates with cblas_dgemv; the
// rotation here sums each row left to right.
#pragma once

#include <array>
#include <cmath>
#include <vector>

#include "lpm.h"
#include "unit_cell.h"

namespace lpm_gpu {
namespace detail {

// y = R p (row-major 3x3), each row summed left to right
inline void rotate(const double* R, const double* p, double* y) {
  for (int r = 0; r < 3; ++r) y[r] = R[3 * r] * p[0] + R[3 * r + 1] * p[1] + R[3 * r + 2] * p[2];
}

inline bool in_box(const double* q, const std::array<double, 6>& box, int dims) {
  for (int k = 0; k < dims; ++k)
    if (!(q[k] >= box[2 * k] && q[k] <= box[2 * k + 1])) return false;
  return true;
}

}  // namespace detail

// createRMatrix (utilities.h:15-91): eulerflag 0 is the X1-Y2-Z3 rotation,
// 1 the Kocks convention (Z1-Y2-Z3), 2 the Bunge convention (Z1-X2-Z3); the
// matrix is row-major in static storage, as the reference returns it.
inline double* createRMatrix(int eulerflag, double angles[]) {
  static double R[9];
  const double a = angles[0], b = angles[1], g = angles[2];
  const double c1 = std::cos(a), s1 = std::sin(a), c2 = std::cos(b), s2 = std::sin(b), c3 = std::cos(g),
               s3 = std::sin(g);
  if (eulerflag == 0) {
    const double m[9] = {c2 * c3, -c2 * s3, s2,
                         c1 * s3 + s1 * s2 * c3, c1 * c3 - s1 * s2 * s3, -s1 * c2,
                         s1 * s3 - c1 * s2 * c3, s1 * c3 + c1 * s2 * s3, c1 * c2};
    for (int k = 0; k < 9; ++k) R[k] = m[k];
  } else if (eulerflag == 1) {
    const double m[9] = {-c1 * c2 * c3 - s1 * s3, c3 * s1 - s3 * c2 * c1, s2 * c1,
                         s3 * c1 - c3 * c2 * s1, -c1 * c3 - s1 * c2 * s3, s2 * s1,
                         c3 * s2, s3 * s2, c2};
    for (int k = 0; k < 9; ++k) R[k] = m[k];
  } else if (eulerflag == 2) {
    const double m[9] = {-s1 * c2 * s3 + c1 * c3, -s3 * c1 - c3 * c2 * s1, s1 * s2,
                         c3 * s1 + s3 * c2 * c1, -s1 * s3 + c1 * c2 * c3, -c1 * s2,
                         s2 * s3, s2 * c3, c2};
    for (int k = 0; k < 9; ++k) R[k] = m[k];
  }
  return R;
}

// createPlateSQ2D (utilities.h:93-125): particles at cell centres of a square
// grid over the padded block; z = 0 is kept after rotation
inline std::vector<std::array<double, 3>> createPlateSQ2D(std::array<double, 6> box, UnitCell cell, double R_matrix[]) {
  const double a = std::sqrt(std::pow(box[1] - box[0], 2) + std::pow(box[3] - box[2], 2));
  const double lo = -a, hi = a;
  const double hx = 2 * cell.radius, hy = hx;
  const int nx = (int)std::floor((hi - lo) / hx), ny = (int)std::floor((hi - lo) / hy);
  std::vector<std::array<double, 3>> out;
  double p[3] = {0, 0, 0}, q[3];
  for (int j = 1; j <= ny; j++) {
    p[1] = lo + hy * (j - 1) + cell.radius;
    for (int i = 1; i <= nx; i++) {
      p[0] = lo + hx * (i - 1) + cell.radius;
      detail::rotate(R_matrix, p, q);
      if (detail::in_box(q, box, 2)) out.push_back({q[0], q[1], 0});
    }
  }
  return out;
}

// createPlateHEX2D (utilities.h:127-177): odd rows hold one more particle
// than the even rows, which are shifted by r
inline std::vector<std::array<double, 3>> createPlateHEX2D(std::array<double, 6> box, UnitCell cell, double R_matrix[]) {
  const double a = std::sqrt(std::pow(box[1] - box[0], 2) + std::pow(box[3] - box[2], 2));
  const double lo = -a, hi = a;
  const double hx = 2 * cell.radius, hy = hx * std::sqrt(3.0) / 2.0;
  const int nx = 1 + (int)std::floor((hi - lo) / hx), ny = 1 + (int)std::floor((hi - lo) / hy);
  std::vector<std::array<double, 3>> out;
  double p[3] = {0, 0, 0}, q[3];
  for (int j = 1; j <= ny; j++) {
    p[1] = lo + hy * (j - 1);
    const bool odd = j % 2 == 1;
    for (int i = 1; i <= (odd ? nx : nx - 1); i++) {
      p[0] = odd ? lo + hx * (i - 1) : lo + hx * (2 * i - 1) / 2.0;
      p[2] = 0;
      detail::rotate(R_matrix, p, q);
      if (detail::in_box(q, box, 2)) out.push_back({q[0], q[1], q[2]});
    }
  }
  return out;
}

// createCuboidSC3D (utilities.h:179-220): x starts at box[0] (not at -a, as y
// and z do), layers z outermost, x innermost
inline std::vector<std::array<double, 3>> createCuboidSC3D(std::array<double, 6> box, UnitCell cell, double R_matrix[]) {
  const double a = std::sqrt(std::pow(box[1] - box[0], 2) + std::pow(box[3] - box[2], 2) + std::pow(box[5] - box[4], 2));
  const double lo = -a, hi = a;
  const double h = 2 * cell.radius;
  const int nx = 1 + (int)std::floor((hi - lo) / h), ny = nx, nz = nx;
  std::vector<std::array<double, 3>> out;
  double p[3], q[3];
  for (int k = 1; k <= nz; k++) {
    p[2] = lo + h * (k - 1);
    for (int j = 1; j <= ny; j++) {
      p[1] = lo + h * (j - 1);
      for (int i = 1; i <= nx; i++) {
        p[0] = box[0] + h * (i - 1);
        detail::rotate(R_matrix, p, q);
        if (detail::in_box(q, box, 3)) out.push_back({q[0], q[1], q[2]});
      }
    }
  }
  return out;
}

// createCuboidFCC3D (utilities.h:222-330): layers half a cell apart; in odd
// layers the odd rows start at the block's edge and the even rows half a cell
// in, the other way round in even layers; all points are generated, then
// rotated and kept inside the box
inline std::vector<std::array<double, 3>> createCuboidFCC3D(std::array<double, 6> box, UnitCell cell, double R_matrix[]) {
  const double a = std::sqrt(std::pow(box[1] - box[0], 2) + std::pow(box[3] - box[2], 2) + std::pow(box[5] - box[4], 2));
  const double lo = -a, hi = a;
  const double h = 4 * cell.radius / std::sqrt(2.0);
  const int nx = 1 + (int)std::floor((hi - lo) / h), ny = 1 + (int)std::floor((hi - lo) / (h / 2.0)), nz = ny;
  std::vector<std::array<double, 3>> pts;
  for (int k = 1; k <= nz; k++) {
    const double z = lo + h / 2.0 * (k - 1);
    for (int j = 1; j <= ny; j++) {
      const double y = lo + h / 2.0 * (j - 1);
      const bool full = (k % 2 == 1) == (j % 2 == 1);  // a row starting at the block's edge
      if (full)
        for (int i = 1; i <= nx; i++) pts.push_back({lo + h * (i - 1), y, z});
      else
        for (int i = 1; i <= nx - 1; i++) pts.push_back({lo + h * (i - 1) + h / 2.0, y, z});
    }
  }
  std::vector<std::array<double, 3>> out;
  double q[3];
  for (const auto& p : pts) {
    detail::rotate(R_matrix, p.data(), q);
    if (detail::in_box(q, box, 3)) out.push_back({q[0], q[1], q[2]});
  }
  return out;
}

}  // namespace lpm_gpu

using lpm_gpu::createCuboidFCC3D;
using lpm_gpu::createCuboidSC3D;
using lpm_gpu::createPlateHEX2D;
using lpm_gpu::createPlateSQ2D;
using lpm_gpu::createRMatrix;
