// Device field models evaluated along the line of sight: the regular
// magnetic field (WMAP-LSA, cartesian, uniform), the regular thermal electron
// density (uniform sphere), the analytic and uniform cosmic-ray electron
// models, and the trilinear vertex-grid gathers of the random fields.
//
// Each function keeps the reference's operation order so that the FP64 path
// reproduces the CPU maps to rounding (see the file:line citations).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hamx.h"
#include "hamx_cgs.h"
#include "hamx_math.h"
#include "hamx_models.h"

namespace hamx {

struct Vec3 {
  double x, y, z;
};

// Hamvec::dotprod (hamvec.h:189-195): the sum starts from +0.
__device__ inline double dot0(const Vec3 &a, const Vec3 &b) {
  double t = 0.;
  t += a.x * b.x;
  t += a.y * b.y;
  t += a.z * b.z;
  return t;
}

// Hamvec::length (hamvec.h:150-156)
__device__ inline double length0(const Vec3 &a) {
  double t = 0.;
  t += a.x * a.x;
  t += a.y * a.y;
  t += a.z * a.z;
  return sqrt(t);
}

// Vertex-grid description used by the trilinear gathers.
struct GridGeo {
  double x_min, x_max, y_min, y_max, z_min, z_max;
  double fx, fy, fz; // (n-1) as double
  double sx, sy, sz; // (n-1) / extent: vertex spacings per unit length
  uint64_t ny, nz;
  uint64_t nyz; // ny * nz (check_box keeps it below 2^32)
};

__host__ inline GridGeo grid_geo(const hamx_box &b) {
  GridGeo g;
  g.x_min = b.x_min;
  g.x_max = b.x_max;
  g.y_min = b.y_min;
  g.y_max = b.y_max;
  g.z_min = b.z_min;
  g.z_max = b.z_max;
  g.fx = static_cast<double>(b.nx - 1);
  g.fy = static_cast<double>(b.ny - 1);
  g.fz = static_cast<double>(b.nz - 1);
  g.sx = g.fx / (b.x_max - b.x_min);
  g.sy = g.fy / (b.y_max - b.y_min);
  g.sz = g.fz / (b.z_max - b.z_min);
  g.ny = b.ny;
  g.nz = b.nz;
  g.nyz = b.ny * b.nz;
  return g;
}

// Cell lookup of Brnd::read_grid (brnd.cc:25-49): strict interior of the
// vertex box, bin index by floor, fractional offsets in [0,1).  Returns false
// on and outside the box faces, where the reference returns a zero field.
struct Cell {
  uint64_t idx; // flat index of corner (xl, yl, zl)
  double xd, yd, zd;
};

// The interior tests are written !(0 < tmp < n-1): the same decision as the
// reference's (tmp <= 0 || tmp >= n-1) for every finite tmp, and a NaN
// position also falls outside instead of reaching an index.
// tmp = (p - min) * ((n-1) / extent) is the reference's
// (n-1) * (p - min) / extent within 2 ulp (a product instead of a division:
// the two differ in a decision only for a position within 2 ulp of a grid
// face); tmp lies in (0, n-1) with n - 1 < 2^32, so the 32-bit truncation is
// the floor, and the cell index needs no 64-bit conversions.
__device__ inline bool grid_cell(const GridGeo &g, const Vec3 &p, Cell &c) {
  double tmp = (p.x - g.x_min) * g.sx;
  if (!(tmp > 0 && tmp < g.fx))
    return false;
  const uint32_t xl = static_cast<uint32_t>(tmp);
  c.xd = tmp - static_cast<double>(xl);
  tmp = (p.y - g.y_min) * g.sy;
  if (!(tmp > 0 && tmp < g.fy))
    return false;
  const uint32_t yl = static_cast<uint32_t>(tmp);
  c.yd = tmp - static_cast<double>(yl);
  tmp = (p.z - g.z_min) * g.sz;
  if (!(tmp > 0 && tmp < g.fz))
    return false;
  const uint32_t zl = static_cast<uint32_t>(tmp);
  c.zd = tmp - static_cast<double>(zl);
  c.idx = static_cast<uint64_t>(xl) * g.nyz +
          (yl * static_cast<uint32_t>(g.nz) + zl);
  return true;
}

// z, then y, then x linear interpolation of one scalar component
// (brnd.cc:55-87, ternd.cc:51-73).
__device__ inline double lerp_cell(const double *f, const Cell &c, uint64_t sy,
                                   uint64_t sx) {
  const uint64_t i00 = c.idx, i01 = c.idx + sy, i10 = c.idx + sx,
                 i11 = c.idx + sx + sy;
  const double i1 = f[i00] * (1. - c.zd) + f[i00 + 1] * c.zd;
  const double i2 = f[i01] * (1. - c.zd) + f[i01 + 1] * c.zd;
  const double j1 = f[i10] * (1. - c.zd) + f[i10 + 1] * c.zd;
  const double j2 = f[i11] * (1. - c.zd) + f[i11 + 1] * c.zd;
  const double w1 = i1 * (1. - c.yd) + i2 * c.yd;
  const double w2 = j1 * (1. - c.yd) + j2 * c.yd;
  return w1 * (1. - c.xd) + w2 * c.xd;
}

// Packed gather: the B components and the TE density of one vertex stored
// together as 4 doubles (32 B); the two z-neighbours of a column are one
// contiguous 64 B record.  Same arithmetic as the separate gathers.
// One x-plane of the cell: the z- then y-interpolated 4 components.
__device__ inline double4 lerp_plane_packed(const double4 *__restrict__ f, uint64_t i0,
                                            uint64_t sy, double zd, double yd) {
  const double4 a0 = f[i0], a1 = f[i0 + 1];
  const double4 b0 = f[i0 + sy], b1 = f[i0 + sy + 1];
  double4 w;
  // a (1 - t) + b t with the second product fused: one rounding fewer (the
  // value moves by at most an ulp) and one FP64 instruction fewer per lerp
  const double z1 = 1. - zd, y1 = 1. - yd;
#define HAMX_LERP2(C)                                                          \
  {                                                                            \
    const double i1 = fma(a1.C, zd, a0.C * z1);                                \
    const double i2 = fma(b1.C, zd, b0.C * z1);                                \
    w.C = fma(i2, yd, i1 * y1);                                                \
  }
  HAMX_LERP2(x)
  HAMX_LERP2(y)
  HAMX_LERP2(z)
  HAMX_LERP2(w)
#undef HAMX_LERP2
  return w;
}

// The two x-planes are interpolated one after the other so that only four
// 32-byte records are live at a time.
__device__ inline void lerp_cell_packed(const double4 *__restrict__ f,
                                        const Cell &c, uint64_t sy, uint64_t sx,
                                        Vec3 &b, double &te) {
  const double4 w1 = lerp_plane_packed(f, c.idx, sy, c.zd, c.yd);
  const double4 w2 = lerp_plane_packed(f, c.idx + sx, sy, c.zd, c.yd);
  const double xd = c.xd, x1 = 1. - xd;
  b.x = fma(w2.x, xd, w1.x * x1);
  b.y = fma(w2.y, xd, w1.y * x1);
  b.z = fma(w2.z, xd, w1.z * x1);
  te = fma(w2.w, xd, w1.w * x1);
}

// ---------------------------------------------------------------- models --

// Regular magnetic field models (breg_lsa.cc:9-26, breg_cart.cc, breg_unif.cc,
// breg_helix.cc:13-30, breg_jf12.cc:8-189, breg_jaffe.cc:11-243).  With
// HEAVY false the JF12 and Jaffe evaluators are compiled out (they return a
// zero field), which keeps their registers out of kernels that never see
// them.
template <bool HEAVY = true>
__device__ inline Vec3 breg_eval(int model, const double *p, const Vec3 &pos) {
  Vec3 b{0., 0., 0.};
  if (model == HAMX_BREG_LSA) {
    const double r = sqrt(pos.x * pos.x + pos.y * pos.y);
    if (r > (20. * cgs::kpc) || r < (3. * cgs::kpc))
      return b;
    const double cos_p = pos.x / r;
    const double sin_p = pos.y / r;
    const double psi = p[1] + p[2] * log_pos(r / (8. * cgs::kpc));
    // tanh(u) = (1 - e) / (1 + e), e = exp(-2|u|): absolute error ~1e-16
    const double u = pos.z / (1. * cgs::kpc);
    const double e = exp_fast(-2. * fabs(u));
    const double chi = p[3] * copysign((1. - e) * rcp_f64(1. + e), u);
    double sps, cps, sch, cch;
    sincos_fast(psi, sps, cps);
    sincos_fast(chi, sch, cch);
    b.x = p[0] * (sps * cch * cos_p - cps * cch * sin_p);
    b.y = p[0] * (sps * cch * sin_p + cps * cch * cos_p);
    b.z = p[0] * sch;
  } else if (model == HAMX_BREG_CART) {
    b.x = p[0];
    b.y = p[1];
    b.z = p[2];
  } else if (model == HAMX_BREG_UNIF) {
    b.x = p[0] * cos(p[2]);
    b.y = p[0] * sin(p[2]);
    b.z = p[1];
  } else if (model == HAMX_BREG_HELIX || model == HAMX_BREG_JF12 || model == HAMX_BREG_JAFFE) {
    double v[3];
    if (model == HAMX_BREG_HELIX)
      breg_helix_eval(p, pos.x, pos.y, pos.z, v);
    else if (HEAVY && model == HAMX_BREG_JF12)
      breg_jf12_eval(p, pos.x, pos.y, pos.z, v);
    else if (HEAVY)
      breg_jaffe_eval(p, pos.x, pos.y, pos.z, v);
    else
      v[0] = v[1] = v[2] = 0.;
    b = Vec3{v[0], v[1], v[2]};
  }
  return b;
}

// TEreg_unif::write_field (tereg_unif.cc:6-12)
__device__ inline double tereg_unif(const double *p, const Vec3 &pos,
                                    const Vec3 &obs) {
  const Vec3 d{pos.x - obs.x, pos.y - obs.y, pos.z - obs.z};
  if (length0(d) > p[1])
    return 0.;
  return p[0];
}

// Lanczos approximation of Gamma(x) (Lanczos 1964; g = 7, n = 9 coefficients
// as tabulated by Godfrey), with Euler's reflection for x < 1/2.  Used only
// when the CRE spectral index varies along the ray; constant indices use the
// host-computed value.
__device__ inline double gamma_lanczos(double x) {
  const double c[9] = {0.99999999999980993,  676.5203681218851,
                       -1259.1392167224028,  771.32342877765313,
                       -176.61502916214059,  12.507343278686905,
                       -0.13857109526572012, 9.9843695780195716e-6,
                       1.5056327351493116e-7};
  bool refl = false;
  double xr = x;
  if (x < 0.5) {
    refl = true;
    xr = 1. - x;
  }
  xr -= 1.;
  double a = c[0];
  const double t = xr + 7.5;
  for (int i = 1; i < 9; ++i)
    a += c[i] / (xr + i);
  const double g = sqrt(cgs::twopi) * pow(t, xr + 0.5) * exp(-t) * a;
  return refl ? cgs::pi / (sin(cgs::pi * x) * g) : g;
}

} // namespace hamx
