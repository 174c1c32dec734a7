// Device functions shared by the force kernels and the adjoint kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace mdx {

// exacterud_sqrt() of sim/erexact.c:25-53: u = -erf(alpha r)/r and
// erd = -u'(r)/r, so that the force on a of a pair is qq erd (r_a - r_b),
// in FP64; the series below rr alpha^2 < 0.01 avoids the cancellation of
// erf(alpha r)/r against its limit
__device__ inline double exact_erud(double rr, double alpha, double &erd) {
  const double alphaq = alpha * alpha, A = alpha * (-2 / 1.772453850905516), B = A * alphaq * 2;
  const double err = alphaq * rr;
  if (err < 0.01) {
    double su = 0, sd = 0, q = 1;
    for (int n = 1;; ) {
      const double t = q / (2 * n + 1);
      sd += t * n;
      su += t;
      if (fabs(q) < 1e-16) break;
      q *= -err / ++n;
    }
    erd = B * sd;
    return A * (1 - su * err);
  }
  const double r = sqrt(rr), y = -erf(alpha * r) / r;
  erd = (y - A * exp(-err)) / rr;
  return y;
}

// cos, sin of 2 pi t in FP64 (t reduced to [0, 1) first)
__device__ inline void cs2pi_d(double t, double &c, double &s) {
  sincospi(2 * (t - floor(t)), &s, &c);
}

}  // namespace mdx
