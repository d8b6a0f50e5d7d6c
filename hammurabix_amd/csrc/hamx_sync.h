// The synchrotron kernels of the CRE grid emissivities on the device,
// replacing gsl_sf_synchrotron_1 / _2 (integrator.cc:321,390):
//   F(x) = x int_x^inf K_5/3(t) dt,   G(x) = x K_2/3(x),   0 < x <= 100.
// Each is exp(P(u) - x) with P a piecewise Chebyshev series of ln f(x) + x in
// u = ln x (tables in hamx_synchrotron.h, written by gen_hamx_synchrotron.py:
// 9 segments of width 1 below u = ln 100, then 12 of width 4); below the
// last segment (x < 1.8e-23) f is its leading term c x^(1/3).
#pragma once

#include <hip/hip_runtime.h>

#include "hamx_synchrotron.h"

namespace hamx {

template <bool POL> // false: F, true: G
__device__ inline double sync_kernel(double x) {
  using namespace sync_tab;
  const double u = log(x);
  const double dtop = kUTop - u;
  const double fine_span = kFineN * kFineW;
  if (dtop >= fine_span + kCoarseN * kCoarseW)
    return exp((POL ? kLnG0 : kLnF0) + u / 3.);
  // Segment indices are clamped in floating point first (fmax drops a NaN),
  // so every table read stays inside the table whatever x is.
  int s, deg;
  double lo, w;
  if (dtop < fine_span) {
    s = static_cast<int>(fmin(fmax(dtop / kFineW, 0.), kFineN - 1.));
    w = kFineW;
    deg = kFineDeg;
    lo = kUTop - (s + 1) * kFineW;
  } else {
    const int sc = static_cast<int>(fmin(fmax((dtop - fine_span) / kCoarseW, 0.), kCoarseN - 1.));
    s = kFineN + sc;
    w = kCoarseW;
    deg = kCoarseDeg;
    lo = kUTop - fine_span - (sc + 1) * kCoarseW;
  }
  const double t = 2. * (u - lo) / w - 1.;
  const double *c = POL ? kG[s] : kF[s];
  // Clenshaw recurrence for sum_k c_k T_k(t)
  double b1 = 0., b2 = 0.;
  for (int k = deg; k >= 1; --k) {
    const double b0 = c[k] + 2. * t * b1 - b2;
    b2 = b1;
    b1 = b0;
  }
  return exp(c[0] + t * b1 - b2 - x);
}

} // namespace hamx
