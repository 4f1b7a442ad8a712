// Scaled 2x2 least-squares solve of the LOMR step sizes, usable on the host
// and on the device.
//
// Restates methods::common::SolveLeastSquares2x2 (src/common/sparse_operations.cpp:58-73):
// symmetric scaling by 1/sqrt|diag|, then the minimum-norm solve of
// Eigen's CompleteOrthogonalDecomposition (3.4) specialised to 2x2 -- column
// pivoted Householder QR (pivot on the larger updated column norm, reflector
// v = [1, e], beta = -sign(c0) * ||.||, tau = (beta - c0) / beta, rank =
// #{|r_ii| > max pivot * 2 eps}), and for rank 1 one Householder reflection
// from the right.  Any non-finite input or output yields (0, 0).
//
// rank_eps replaces eps in the rank decision only.  The default is the
// reference's (DBL_EPSILON); FP32-bulk callers whose Gram entries carry FP32
// rounding pass a larger one so that pivots below the data's precision count
// as zero instead of producing large cancelling coefficients.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace spai {

struct Householder2 {
  double essential, tau, beta;
};

__host__ __device__ inline Householder2 MakeHouseholder2(double c0, double t) {
  Householder2 h;
  const double tail_sq = t * t;
  if (tail_sq <= DBL_MIN) {
    h.tau = 0.0;
    h.beta = c0;
    h.essential = 0.0;
  } else {
    double beta = sqrt(c0 * c0 + tail_sq);
    if (c0 >= 0.0) beta = -beta;
    h.essential = t / (c0 - beta);
    h.tau = (beta - c0) / beta;
    h.beta = beta;
  }
  return h;
}

__host__ __device__ inline void ApplyHouseholder2(double e, double tau, double& v0, double& v1) {
  if (tau == 0.0) return;
  double tmp = e * v1;
  tmp += v0;
  v0 -= tau * tmp;
  v1 -= (tau * e) * tmp;
}

// m = [[m00, m01], [m10, m11]], x = (x0, x1); result in (d, g).
__host__ __device__ inline void SolveLeastSquares2x2(double m00, double m01, double m10, double m11, double x0,
                                                     double x1, double& d, double& g,
                                                     double rank_eps = DBL_EPSILON) {
  d = 0.0;
  g = 0.0;
  if (!__builtin_isfinite(m00) || !__builtin_isfinite(m01) || !__builtin_isfinite(m10) || !__builtin_isfinite(m11) || !__builtin_isfinite(x0) || !__builtin_isfinite(x1)) return;
  const double r0 = sqrt(fabs(m00)), r1 = sqrt(fabs(m11));
  const double s0 = r0 > 0.0 ? 1.0 / r0 : 1.0;
  const double s1 = r1 > 0.0 ? 1.0 / r1 : 1.0;
  // scaled matrix, column-major a[col][row]
  double a00 = (s0 * m00) * s0, a10 = (s1 * m10) * s0;
  double a01 = (s0 * m01) * s1, a11 = (s1 * m11) * s1;
  const double b0 = s0 * x0, b1 = s1 * x1;
  const double eps = DBL_EPSILON;
  double nd0 = sqrt(a00 * a00 + a10 * a10), nd1 = sqrt(a01 * a01 + a11 * a11);
  double nu0 = nd0, nu1 = nd1;
  const double maxnorm = nu0 > nu1 ? nu0 : nu1;
  const double helper = (maxnorm * eps) * (maxnorm * eps) / 2.0;
  const double downdate = sqrt(eps);
  int nonzero_pivots = 2;
  bool swapped = false;
  if (nu1 > nu0) {
    double t;
    t = a00; a00 = a01; a01 = t;
    t = a10; a10 = a11; a11 = t;
    t = nu0; nu0 = nu1; nu1 = t;
    t = nd0; nd0 = nd1; nd1 = t;
    swapped = true;
  }
  if (nu0 * nu0 < helper * 2.0) nonzero_pivots = 0;
  const Householder2 h0 = MakeHouseholder2(a00, a10);
  a00 = h0.beta;
  double maxpivot = fabs(h0.beta);
  ApplyHouseholder2(h0.essential, h0.tau, a01, a11);
  if (nu1 != 0.0) {
    double t = fabs(a01) / nu1;
    t = (1.0 + t) * (1.0 - t);
    if (t < 0.0) t = 0.0;
    const double ratio = nu1 / nd1;
    const double t2 = t * ratio * ratio;
    if (t2 <= downdate) {
      nd1 = fabs(a11);
      nu1 = nd1;
    } else {
      nu1 *= sqrt(t);
    }
  }
  if (nonzero_pivots == 2 && nu1 * nu1 < helper * 1.0) nonzero_pivots = 1;
  if (fabs(a11) > maxpivot) maxpivot = fabs(a11);
  const double thr = maxpivot * (rank_eps * 2.0);
  int rank = 0;
  if (nonzero_pivots > 0 && fabs(a00) > thr) ++rank;
  if (nonzero_pivots > 1 && fabs(a11) > thr) ++rank;
  if (rank == 0) return;
  double c0 = b0, c1 = b1;
  ApplyHouseholder2(h0.essential, h0.tau, c0, c1);
  double y0 = 0.0, y1 = 0.0;
  if (rank == 2) {
    double u1 = c1, u0 = c0;
    if (u1 != 0.0) {
      u1 /= a11;
      u0 -= u1 * a01;
    }
    if (u0 != 0.0) u0 /= a00;
    y0 = u0;
    y1 = u1;
  } else {
    const Householder2 z0 = MakeHouseholder2(a00, a01);
    double u0 = c0;
    if (u0 != 0.0) u0 /= z0.beta;
    double e0 = u0, e1 = 0.0;
    ApplyHouseholder2(z0.essential, z0.tau, e0, e1);
    y0 = e0;
    y1 = e1;
  }
  const double sol0 = swapped ? y1 : y0;
  const double sol1 = swapped ? y0 : y1;
  d = s0 * sol0;
  g = s1 * sol1;
  if (!__builtin_isfinite(d) || !__builtin_isfinite(g)) {
    d = 0.0;
    g = 0.0;
  }
}

}  // namespace spai
