// Element stiffness matrices of the Q1 hexahedron (host, FP64).
//
// The reference integrates each of the 8^l sub-cells of a level-l element
// with a 10x10x10 midpoint rule (local_matrix.c:251-319), evaluating the
// eta coordinate at -y (local_matrix.c:298).  On an axis-aligned brick the
// integrand B^T C B is a polynomial of degree <= 2 in each coordinate, and
// the n-point midpoint rule integrates 1, s, s^2 over an interval of
// half-width h exactly like the 2-point rule at c +- h*t with
// t^2 = 1/3 - 1/(3 n^2) (both give 2h, 0, 2h^3/3 - 2h^3/(3 n^2)).  This file
// uses that 2-point tensor rule: the same matrices to rounding, 125x fewer
// evaluations, which keeps set-up cheap at 8^5 sub-cells.
#include <cmath>
#include <cstring>

#include "t3d_internal.hpp"

namespace t3d {

namespace {

constexpr int kMidpointPoints = 10;  // local_matrix.c:256

void constitutive(double C[6][6], double nu) {
  const double t1 = (1.0 - nu) / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double t2 = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double t3 = 1.0 / (2.0 * (1.0 + nu));
  std::memset(C, 0, sizeof(double) * 36);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) C[i][j] = (i == j) ? t1 : t2;
  C[3][3] = C[4][4] = C[5][5] = t3;
}

// iso-parametric corner signs of nodes n1..n8 (local_matrix.c:45-73)
constexpr int kCorner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1},
                               {-1, 1, -1},  {-1, -1, 1}, {1, -1, 1},
                               {1, 1, 1},    {-1, 1, 1}};

// strain rows (xx, yy, zz, xy, yz, xz) of the brick with half size a
void strain(double B[6][24], const double iso[3], double a) {
  std::memset(B, 0, sizeof(double) * 144);
  const double inv = 1.0 / a;
  for (int n = 0; n < 8; n++) {
    const double s0 = kCorner[n][0], s1 = kCorner[n][1], s2 = kCorner[n][2];
    const double dx = inv * 0.125 * s0 * (1 + s1 * iso[1]) * (1 + s2 * iso[2]);
    const double dy = inv * 0.125 * s1 * (1 + s0 * iso[0]) * (1 + s2 * iso[2]);
    const double dz = inv * 0.125 * s2 * (1 + s0 * iso[0]) * (1 + s1 * iso[1]);
    B[0][3 * n] = dx;
    B[1][3 * n + 1] = dy;
    B[2][3 * n + 2] = dz;
    B[3][3 * n] = dy;
    B[3][3 * n + 1] = dx;
    B[4][3 * n + 1] = dz;
    B[4][3 * n + 2] = dy;
    B[5][3 * n] = dz;
    B[5][3 * n + 2] = dx;
  }
}

}  // namespace

void element_matrices(double nu, int l, double *KE) {
  const int n = 1 << l;
  const double a = 0.5 * n;
  double C[6][6];
  constitutive(C, nu);
  const double t = std::sqrt(1.0 / 3.0 - 1.0 / (3.0 * kMidpointPoints * kMidpointPoints));
  const double hw = 1.0 / n;                 // sub-cell half width (iso units)
  const double w = hw * hw * hw * a * a * a;  // point weight * Jacobian
  std::memset(KE, 0, sizeof(double) * 576 * n * n * n);
  for (int ii = 0; ii < n; ii++)
    for (int kk = 0; kk < n; kk++)
      for (int jj = 0; jj < n; jj++) {
        double *K = KE + 576 * (n * n * ii + n * kk + jj);
        const double cx = -1.0 + (2.0 * ii + 1.0) * hw;
        const double cy = -1.0 + (2.0 * jj + 1.0) * hw;
        const double cz = -1.0 + (2.0 * kk + 1.0) * hw;
        for (int p = 0; p < 8; p++) {
          const double iso[3] = {cx + ((p & 1) ? hw * t : -hw * t),
                                 -(cy + ((p & 2) ? hw * t : -hw * t)),
                                 cz + ((p & 4) ? hw * t : -hw * t)};
          double B[6][24], BtC[24][6];
          strain(B, iso, a);
          for (int i = 0; i < 24; i++)
            for (int j = 0; j < 6; j++) {
              double s = 0.0;
              for (int k = 0; k < 6; k++) s += B[k][i] * C[k][j];
              BtC[i][j] = s;
            }
          for (int i = 0; i < 24; i++)
            for (int j = 0; j < 24; j++) {
              double s = 0.0;
              for (int k = 0; k < 6; k++) s += BtC[i][k] * B[k][j];
              K[24 * i + j] += w * s;
            }
        }
      }
}

}  // namespace t3d

extern "C" void getKEsubspace(MTYPE *KEarray, const MTYPE nu, const int l) {
  t3d::element_matrices(nu, l, KEarray);
}
