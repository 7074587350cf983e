// Near-field Taylor coefficients of the retarded Green function over one
// flat face (Greenfun/@greenret/private/init.m:98-177, deriv 'cart'):
//   g_n = sum_q w_q (r - r0)^n / (n! r)
//   f_0 = -sum_q w_q (x - x_q) / r^3
//   f_n = sum_q w_q (x - x_q) [-(r - r0)^n / (r^3 n!) + (r - r0)^(n-1) / (r^2 (n-1)!)]
// with r = |x - x_q| over the 28 points of rule 18, r0 = |x - pos_j| and the
// weights scaled by the face area; G = (sum_n g_n (ik)^n) e^{ik r0} and
// Gp = (sum_n f_n (ik)^n) e^{ik r0}.
#pragma once

#include "mnp_greenret.hpp"

namespace mnp {

// Observation point x, face j of geo; q18 = rule-18 table [x(28) | y(28) | w(28)].
// g and f (3 x (order+1)) are accumulated into (callers zero them).
__device__ inline void near_taylor(const double x[3], const Geo& geo, int64_t n, int64_t j, int order,
                                   const double* fact, const double* q18, double* g,
                                   double (*f)[MNP_MAX_ORDER + 1]) {
  // No FMA contraction: the quadrature points and distances round as the
  // reference's vectorised expressions do.
#pragma clang fp contract(off)
  double vj[3][3], pj[3];
  for (int c = 0; c < 3; ++c) {
    pj[c] = geo.pos[c * n + j];
    for (int a = 0; a < 3; ++a) vj[a][c] = geo.vtx[(a * 3 + c) * n + j];
  }
  const double r0 = sqrt((pj[0] - x[0]) * (pj[0] - x[0]) + (pj[1] - x[1]) * (pj[1] - x[1]) +
                         (pj[2] - x[2]) * (pj[2] - x[2]));
  const double qa = geo.qarea[j];
  for (int q = 0; q < 28; ++q) {
    const double a = q18[q], b = q18[28 + q], z = 1 - a - b, w = q18[56 + q] * qa;
    double xv[3];
    for (int c = 0; c < 3; ++c) xv[c] = x[c] - (a * vj[0][c] + b * vj[1][c] + z * vj[2][c]);
    const double r = sqrt(xv[0] * xv[0] + xv[1] * xv[1] + xv[2] * xv[2]);
    const double dr = r - r0, ir = 1.0 / r, ir2 = ir * ir, ir3 = ir2 * ir;
    double drp = 1.0;  // dr^o
    for (int o = 0; o <= order; ++o) {
      g[o] += drp * ir / fact[o] * w;
      drp *= dr;
    }
    for (int c = 0; c < 3; ++c) f[c][0] -= xv[c] * (w * ir3);
    double drm = 1.0;  // dr^(o-1)
    for (int o = 1; o <= order; ++o) {
      const double fac = -(drm * dr) * ir3 / fact[o] + drm * ir2 / fact[o - 1];
      for (int c = 0; c < 3; ++c) f[c][o] += xv[c] * (fac * w);
      drm *= dr;
    }
  }
}

// Host: the 28-point rule 18 on the unit triangle (x, y, w).
void rule18(double* x, double* y, double* w);

}  // namespace mnp
