// The synchrotron kernels of the CRE grid emissivities on the device,
// replacing gsl_sf_synchrotron_1 / _2 (integrator.cc:321,390):
//   F(x) = x int_x^inf K_5/3(t) dt,   G(x) = x K_2/3(x),   0 < x <= 100.
// F = exp(P(u) - x) and G = F R(u), with P(u) = ln F + x and R(u) = G / F
// piecewise Chebyshev series in u = ln x (tables in hamx_synchrotron.h,
// written by gen_hamx_synchrotron.py: 9 segments of width 1 below
// u = ln 100, then 12 of width 4, one Clenshaw loop for both series); below
// the last segment (x < 1.8e-23) F is its leading term 2^(2/3) Gamma(2/3)
// x^(1/3) and R = 1/2.
#pragma once

#include <hip/hip_runtime.h>

#include "hamx_synchrotron.h"

namespace hamx {

// Both tables in one array, [kF | kR] with kSeg x kCoef doubles each (the
// march copies it into LDS once per workgroup).
constexpr int kSyncTab = 2 * sync_tab::kSeg * sync_tab::kCoef;

__device__ inline void load_sync_tab(double *dst) {
  using namespace sync_tab;
  for (int i = threadIdx.x; i < kSeg * kCoef; i += blockDim.x) {
    dst[i] = (&kF[0][0])[i];
    dst[kSeg * kCoef + i] = (&kR[0][0])[i];
  }
}

// F(x) and G(x) with u = ln x given by the caller: the march forms it as a
// per-sample logarithm plus a per-bin constant instead of a logarithm per
// bin.  P changes by less than the error in u times 5/6, so u may carry a few
// ulps of error.  `tab` holds the two tables as load_sync_tab lays them out.
__device__ inline void sync_fg(const double *tab, double u, double x, double &F, double &G) {
  using namespace sync_tab;
  const double dtop = kUTop - u;
  const double fine_span = kFineN * kFineW;
  if (dtop >= fine_span + kCoarseN * kCoarseW) {
    F = exp(kLnF0 + u / 3.);
    G = kR0 * F;
    return;
  }
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
  const double *cp = tab + s * kCoef, *cr = tab + (kSeg + s) * kCoef;
  // Clenshaw recurrences for sum_k c_k T_k(t), both series together
  double p1 = 0., p2 = 0., r1 = 0., r2 = 0.;
  for (int k = deg; k >= 1; --k) {
    const double p0 = cp[k] + 2. * t * p1 - p2;
    const double r0 = cr[k] + 2. * t * r1 - r2;
    p2 = p1;
    p1 = p0;
    r2 = r1;
    r1 = r0;
  }
  F = exp(cp[0] + t * p1 - p2 - x);
  G = F * (cr[0] + t * r1 - r2);
}

} // namespace hamx
