// FP64 building blocks shared by the marches and the random-field kernels:
// a Newton-refined reciprocal and sines / cosines by a Cody-Waite reduction
// and Taylor series, each within a few ulp (the FP64 paths are held to 1e-10
// and 1e-11 relative, not to the last bit of a library routine).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace hamx {

// a b + c for a constant c, as one v_fma_f64 that reads c from a scalar
// register pair (two scalar moves set it up).  For fma(a, b, constant) the
// compiler selects the two-address v_fmac_f64 and first copies the constant
// into the accumulator's vector registers with two v_mov_b32: three VALU
// instructions per polynomial step instead of one.  A translation unit that
// defines HAMX_FMA_K_ASM as 0 before this header keeps the compiler's form
// (the reverse march measured 2% faster with it, the LOS march 1.7% slower).
#ifndef HAMX_FMA_K_ASM
#define HAMX_FMA_K_ASM 1
#endif
__device__ inline double fma_k(double a, double b, double c) {
#if HAMX_FMA_K_ASM
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
#else
  return fma(a, b, c);
#endif
}

// 1/x for normal x: the hardware v_rcp_f64 estimate refined by two Newton
// steps (each squares the relative error), within 1 ulp of the quotient;
// a division by a value that feeds no decision, without the IEEE division's
// scaling and fix-up sequence.
__device__ inline double rcp_f64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// 1/sqrt(x) for a positive normal x: the hardware v_rsq_f64 estimate
// (about 26 correct bits) refined by two Newton steps y += y (1 - x y^2) / 2,
// each of which doubles the correct bits.
__device__ inline double rsqrt_f64(double x) {
  double y = __builtin_amdgcn_rsq(x);
  double e = fma(-(x * y), y, 1.0);
  y = fma(0.5 * y, e, y);
  e = fma(-(x * y), y, 1.0);
  return fma(0.5 * y, e, y);
}

// sin and cos of |x| <= pi/4 (no range reduction): sin x = x + x^3 ps(x^2),
// cos x = 1 - x^2/2 + x^4 pc(x^2), with ps and pc degree-5 polynomials in
// x^2 fitted (Chebyshev interpolation at 60 digits, coefficients rounded to
// double) over |x| <= 0.8, a margin above pi/4.  Evaluated with fused
// multiply-adds, within 0.7 ulp (sin) and 0.85 ulp (cos) of the exact values
// on a 20000-point sweep of [0, 0.8].
__device__ inline void sincos_series(double x, double &s, double &c) {
  const double x2 = x * x;
  double ps = fma_k(x2, 1.59128642249034849e-10, -2.50510573944921275e-08);
  ps = fma_k(x2, ps, 2.75573157385868453e-06);
  ps = fma_k(x2, ps, -1.98412698360435748e-04);
  ps = fma_k(x2, ps, 8.33333333333046745e-03);
  ps = fma_k(x2, ps, -1.66666666666666630e-01);
  s = fma(x * x2, ps, x);
  double pc = fma_k(x2, -1.13793385852981359e-11, 2.08760996923149261e-09);
  pc = fma_k(x2, pc, -2.75573170440356690e-07);
  pc = fma_k(x2, pc, 2.48015872983183947e-05);
  pc = fma_k(x2, pc, -1.38888888888870961e-03);
  pc = fma_k(x2, pc, 4.16666666666666644e-02);
  c = fma(x2, fma(x2, pc, -0.5), 1.0);
}

// ln x for a positive normal x (the callers' arguments: radii, |B_perp|,
// uniform deviates, wave numbers): x = m 2^e with m in [1/sqrt 2, sqrt 2),
// ln m = 2 atanh f with f = (m - 1)/(m + 1), |f| <= 0.1716, as
// 2 f (1 + f^2 p(f^2)) with p a degree-6 fit of the atanh series, plus e ln 2 with
// ln 2 in two parts through fused multiply-adds.  Within 2 ulp of the
// library log (a CPU sweep of 2e7 points over 1e-30..1e30 and near 1), at
// about a third of its instructions.
__device__ inline double log_pos(double x) {
  int e;
  double m = frexp(x, &e); // [1/2, 1)
  if (m < 0.70710678118654752) {
    m *= 2.;
    --e;
  }
  const double f = (m - 1.) * rcp_f64(m + 1.);
  const double s = f * f;
  // (atanh f / f - 1) / f^2 = 1/3 + s/5 + s^2/7 + ... as a degree-6
  // polynomial in s fitted over 0 <= s <= 0.02944 (Chebyshev interpolation)
  double p = fma_k(s, 7.30829271244431145e-02, 7.66585530910538632e-02);
  p = fma_k(s, p, 9.09144472783747354e-02);
  p = fma_k(s, p, 1.11111055651299390e-01);
  p = fma_k(s, p, 1.42857143130016351e-01);
  p = fma_k(s, p, 1.99999999999497219e-01);
  p = fma_k(s, p, 3.33333333333333481e-01);
  const double lnm = 2. * f + (2. * f) * (s * p);
  const double ed = static_cast<double>(e);
  return fma(ed, 6.93147180559945286227e-01, fma(ed, 2.31904681384629955842e-17, lnm));
}

// e^x for |x| < 708: k = rint(x / ln 2), r = x - k ln 2 (ln 2 in two parts,
// fused), e^r = 1 + r + r^2/2 + r^3 q(r) on |r| <= ln 2 / 2 with q a
// degree-8 fit of the Taylor remainder, scaled by 2^k.  Within 1 ulp of the library exp
// (a CPU sweep of 2e7 points over -740..710), at under half its instructions.
__device__ inline double exp_fast(double x) {
  const double k = rint(x * 1.44269504088896338700); // 1 / ln 2
  double r = fma(-k, 6.93147180559945286227e-01, x);
  r = fma(-k, 2.31904681384629955842e-17, r);
  // (e^r - 1 - r - r^2/2) / r^3 = 1/6 + r/24 + ... as a degree-8 polynomial
  // fitted over |r| <= ln 2 / 2 (Chebyshev interpolation)
  double p = 2.50955546771111123e-08;
  p = fma_k(p, r, 2.76138068615382899e-07);
  p = fma_k(p, r, 2.75572800670169807e-06);
  p = fma_k(p, r, 2.48015363886907663e-05);
  p = fma_k(p, r, 1.98412698543386110e-04);
  p = fma_k(p, r, 1.38888889058815424e-03);
  p = fma_k(p, r, 8.33333333333215534e-03);
  p = fma_k(p, r, 4.16666666666513572e-02);
  p = fma_k(p, r, 1.66666666666666657e-01);
  p = fma(p, r, 0.5);
  p = fma(p, r, 1.);
  p = fma(p, r, 1.);
  return ldexp(p, static_cast<int>(k));
}

// The library sincos, out of line: the large-argument reduction it carries
// then takes no registers in the callers, which reach it only beyond 2^19 pi.
__device__ __noinline__ inline void sincos_lib(double x, double &s, double &c) {
  sincos(x, &s, &c);
}

// sin and cos of any x: for |x| < 2^19 pi a Cody-Waite reduction by pi/2,
// whose three parts hold 33, 33 and 53 bits of pi/2 (k P1 and k P2 are
// exact for |k| < 2^20, so r = x - k pi/2 is exact to the third part's
// rounding), then the series on |r| <= pi/4 and the quadrant's signs;
// beyond that the library sincos.
__device__ inline void sincos_fast(double x, double &s, double &c) {
  const double k = rint(x * 0.6366197723675814); // 2 / pi
  if (fabs(k) < 1048576.0) {
    double r = fma(-k, 1.5707963267341256, x);
    r = fma(-k, 6.077100506303966e-11, r);
    r = fma(-k, 2.0222662487959506e-21, r);
    double sr, cr;
    sincos_series(r, sr, cr);
    const int q = static_cast<int>(k) & 3;
    s = (q == 0) ? sr : (q == 1) ? cr : (q == 2) ? -sr : -cr;
    c = (q == 0) ? cr : (q == 1) ? -sr : (q == 2) ? -cr : sr;
  } else {
    sincos_lib(x, s, c);
  }
}

} // namespace hamx
