// hamx_oracle.cc -- CPU restatement of the hammurabiX LOS integrator and
// random-field generators (FP64, OpenMP).
//
// TEST INFRASTRUCTURE ONLY: the checker for the HIP path, the CPU baseline
// timed by bench.py, and nothing else.  The product never links it.
//
// Every routine follows the reference's arithmetic in the reference's
// operation order; the citations give the file:line it restates.  Parity
// status of each part (what pins it) is listed in oracle/README.md.
#include "hamx_oracle.h"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------- units --
// CGS constants (hamunits.h:8-68)
const double kPi = 3.141592653589793238462643383279502884197;
const double kHalfPi = 1.570796326794896619231321691639751442099;
const double kTwoPi = 6.283185307179586476925286766559005768394;
const double kOneThird = 0.333333333333333333333333333333333333333;
const double kTwoThirds = 0.666666666666666666666666666666666666667;
const double kSqrt2 = 1.4142135623730950488016887242096980785697;
const double kSqrt3 = 1.7320508075688772935274463415058723669428;
const double kKpc = 3.0856775806e+21;
const double kPc = 3.0856775806e+18;
const double kRad = kPi / 180.;
const double kM = 1.e2;
const double kSec = 1.;
const double kEv = 1.60217733e-12;
const double kGeV = 1.e9 * kEv;
const double kC = 2.99792458e+10;
const double kH = 6.626075540e-27;
const double kQe = 4.80320425e-10;
const double kMec2 = 0.51099907e-3 * kGeV;
const double kMec = kMec2 / kC;
const double kKB = 1.380622e-16;
const double kUndef = -1.6375e30;

// ------------------------------------------------------------- vector3 --
struct V3 {
  double e[3];
  double operator[](int i) const { return e[i]; }
  double &operator[](int i) { return e[i]; }
};
V3 v3(double x, double y, double z) {
  V3 v;
  v.e[0] = x;
  v.e[1] = y;
  v.e[2] = z;
  return v;
}
V3 add(const V3 &a, const V3 &b) {
  V3 t = a;
  for (int i = 0; i < 3; ++i)
    t[i] += b[i];
  return t;
}
V3 sub(const V3 &a, const V3 &b) {
  V3 t = a;
  for (int i = 0; i < 3; ++i)
    t[i] -= b[i];
  return t;
}
V3 mul(const V3 &a, double s) {
  V3 t = a;
  for (int i = 0; i < 3; ++i)
    t[i] *= s;
  return t;
}
// hamvec.h:150-165
double length(const V3 &a) {
  double t = 0.;
  for (int i = 0; i < 3; ++i)
    t += a[i] * a[i];
  return std::sqrt(t);
}
// hamvec.h:189-195
double dot(const V3 &a, const V3 &b) {
  double t = 0.;
  for (int i = 0; i < 3; ++i)
    t += a[i] * b[i];
  return t;
}
// hamvec.h:198-206
V3 cross(const V3 &a, const V3 &b) {
  V3 t;
  for (int i = 0; i < 3; ++i)
    t[i] = a[(i + 1) % 3] * b[(i + 2) % 3] - a[(i + 2) % 3] * b[(i + 1) % 3];
  return t;
}

// ------------------------------------------------------------- HEALPix --
// Hampix RING geometry (hamdis.h:223-392)
struct Ring {
  uint64_t nside, npix, ncap;
  int order;
  double fact1, fact2;
  explicit Ring(uint64_t n) {
    nside = n;
    int o = 0;
    for (uint64_t t = n; t > 1; t >>= 1)
      ++o;
    order = o;
    const uint64_t npface = n * n;
    npix = 12 * npface;
    ncap = (npface - n) << 1;
    fact2 = 4. / npix;
    fact1 = (n << 1) * fact2;
  }
  static uint64_t isqrt(uint64_t v) {
    return static_cast<uint64_t>(std::floor(std::sqrt(static_cast<double>(v) + 0.5)));
  }
  // base-pixel layout of the HEALPix faces
  static int64_t jrll(int f) { static const int t[12] = {2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4}; return t[f]; }
  static int64_t jpll(int f) { static const int t[12] = {1, 3, 5, 7, 0, 2, 4, 6, 1, 3, 5, 7}; return t[f]; }
  // Hampix::rpixxyf (hamdis.h:515-570)
  void ring2xyf(uint64_t pix, int64_t &ix, int64_t &iy, int &face) const {
    const int64_t ns = (int64_t)nside, nl2 = 2 * ns;
    int64_t iring, iphi, kshift, nr;
    if (pix < ncap) {
      iring = (1 + (int64_t)isqrt(1 + 2 * pix)) >> 1;
      iphi = (int64_t)pix + 1 - 2 * iring * (iring - 1);
      kshift = 0;
      nr = iring;
      face = (int)((iphi - 1) / nr);
    } else if (pix < npix - ncap) {
      const int64_t ip = (int64_t)(pix - ncap);
      const int64_t tmp = ip / (4 * ns);
      iring = tmp + ns;
      iphi = ip - tmp * 4 * ns + 1;
      kshift = (iring + ns) & 1;
      nr = ns;
      const int64_t ire = tmp + 1, irm = nl2 + 1 - tmp;
      const int64_t ifm = (iphi - (ire >> 1) + ns - 1) / ns;
      const int64_t ifp = (iphi - (irm >> 1) + ns - 1) / ns;
      face = (int)((ifp == ifm) ? (ifp | 4) : ((ifp < ifm) ? ifp : (ifm + 8)));
    } else {
      const int64_t ip = (int64_t)(npix - pix);
      iring = (1 + (int64_t)isqrt(2 * ip - 1)) >> 1;
      iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
      kshift = 0;
      nr = iring;
      iring = 2 * nl2 - iring;
      face = 8 + (int)((iphi - 1) / nr);
    }
    const int64_t irt = iring - jrll(face) * ns + 1;
    int64_t ipt = 2 * iphi - jpll(face) * nr - kshift - 1;
    if (ipt >= nl2)
      ipt -= 8 * ns;
    ix = (ipt - irt) >> 1;
    iy = (-ipt - irt) >> 1;
  }
  // Hampix::xyfrpix (hamdis.h:572-605)
  uint64_t xyf2ring(int64_t ix, int64_t iy, int face) const {
    const int64_t ns = (int64_t)nside, nl4 = 4 * ns;
    const int64_t jr = jrll(face) * ns - ix - iy - 1;
    int64_t nr, n_before, kshift;
    if (jr < ns) {
      nr = jr;
      n_before = 2 * nr * (nr - 1);
      kshift = 0;
    } else if (jr < 3 * ns) {
      nr = ns;
      n_before = (int64_t)ncap + (jr - ns) * nl4;
      kshift = (jr - ns) & 1;
    } else {
      nr = nl4 - jr;
      n_before = (int64_t)npix - 2 * nr * (nr + 1);
      kshift = 0;
    }
    int64_t jp = (jpll(face) * nr + ix - iy + 1 + kshift) / 2;
    if (jp < 1)
      jp += nl4;
    return (uint64_t)(n_before + jp - 1);
  }
  // fillpoint (hamdis.h:302-356) + Hamp::norm (hamp.h:23-28)
  void pix2ang(uint64_t idx, double &theta, double &phi) const {
    double z = 0.0, sth = 0.0;
    bool have_sth = false;
    if (idx < ncap) {
      const uint64_t iring = (1 + isqrt(1 + 2 * idx)) >> 1;
      const uint64_t iphi = (idx + 1) - 2 * iring * (iring - 1);
      const double tmp = (iring * iring) * fact2;
      z = 1.0 - tmp;
      if (z > 0.99) {
        sth = std::sqrt(tmp * (2.0 - tmp));
        have_sth = true;
      }
      phi = (iphi - 0.5) * kHalfPi / iring;
    } else if (idx < (npix - ncap)) {
      const uint64_t nl4 = 4 * nside;
      const uint64_t ip = idx - ncap;
      const uint64_t tmp = ip >> (order + 2);
      const uint64_t iring = tmp + nside;
      const uint64_t iphi = ip - nl4 * tmp + 1;
      const double fodd = ((iring + nside) & 1) ? 1 : 0.5;
      z = (static_cast<double>(2 * nside) - static_cast<double>(iring)) * fact1;
      phi = (iphi - fodd) * kHalfPi * 1.5 * fact1;
    } else {
      const uint64_t ip = npix - idx;
      const uint64_t iring = (1 + isqrt(2 * ip - 1)) >> 1;
      const uint64_t iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
      const double tmp = (iring * iring) * fact2;
      z = tmp - 1.0;
      if (z < -0.99) {
        sth = std::sqrt(tmp * (2.0 - tmp));
        have_sth = true;
      }
      phi = (iphi - 0.5) * kHalfPi / iring;
    }
    theta = have_sth ? std::atan2(sth, z) : std::acos(z);
    theta = std::fmod(theta, kPi);
    theta += (theta < 0) * kPi;
    phi = std::fmod(phi, kTwoPi);
    phi += (phi < 0) * kTwoPi;
  }
  // ring_above (hamdis.h:358-365)
  uint64_t rabove(double z) const {
    const double az = std::fabs(z);
    if (az <= kTwoThirds)
      return static_cast<uint64_t>(nside * (2 - 1.5 * z));
    const uint64_t iring = static_cast<uint64_t>(nside * std::sqrt(3 * (1 - az)));
    return (z > 0) ? iring : 4 * nside - iring - 1;
  }
  // get_ring_info2 (hamdis.h:367-392)
  void rinfo(uint64_t ring, uint64_t &startpix, int64_t &ringpix, double &theta,
             bool &shifted) const {
    const uint64_t northring = (ring > 2 * nside) ? 4 * nside - ring : ring;
    if (northring < nside) {
      const double tmp = northring * northring * fact2;
      const double costheta = 1 - tmp;
      const double sintheta = std::sqrt(tmp * (2 - tmp));
      theta = std::atan2(sintheta, costheta);
      ringpix = 4 * northring;
      shifted = true;
      startpix = 2 * northring * (northring - 1);
    } else {
      theta = std::acos((static_cast<double>(2 * nside) -
                         static_cast<double>(northring)) * fact1);
      ringpix = 4 * nside;
      shifted = ((northring - nside) & 1) == 0;
      startpix = ncap + (northring - nside) * ringpix;
    }
    if (northring != ring) {
      theta = kPi - theta;
      startpix = npix - startpix - ringpix;
    }
  }
  // interpolate (hamdis.h:609-696)
  double interpolate(const double *map, double theta, double phi) const {
    uint64_t pix[4] = {0, 0, 0, 0};
    double wght[4] = {0, 0, 0, 0};
    const double z = std::cos(theta);
    const uint64_t ring1 = rabove(z);
    const uint64_t ring2 = ring1 + 1;
    double theta1 = 0.0, theta2 = 0.0;
    for (int r = 0; r < 2; ++r) {
      const uint64_t ring = r == 0 ? ring1 : ring2;
      const bool ok = r == 0 ? (ring1 > 0) : (ring2 < 4 * nside);
      if (!ok)
        continue;
      uint64_t start;
      int64_t ringpix;
      bool shift;
      double th;
      rinfo(ring, start, ringpix, th, shift);
      if (r == 0)
        theta1 = th;
      else
        theta2 = th;
      const double dphi = 6.283185307179586476925286766559005768394 / ringpix;
      const double tmp = (phi / dphi - 0.5 * shift);
      int64_t i1 = static_cast<int64_t>(tmp) - static_cast<int64_t>(tmp < 0);
      const double weight = (phi - (i1 + 0.5 * shift) * dphi) / dphi;
      int64_t i2 = i1 + 1;
      i1 += ringpix * (i1 < 0);
      i2 -= ringpix * (i2 >= ringpix);
      pix[2 * r] = start + i1;
      pix[2 * r + 1] = start + i2;
      wght[2 * r] = 1.0 - weight;
      wght[2 * r + 1] = weight;
    }
    if (ring1 == 0) {
      const double wtheta = theta / theta2;
      wght[2] *= wtheta;
      wght[3] *= wtheta;
      const double fac = (1.0 - wtheta) * 0.25;
      wght[0] = fac;
      wght[1] = fac;
      wght[2] += fac;
      wght[3] += fac;
      pix[0] = (pix[2] + 2) & 3;
      pix[1] = (pix[3] + 2) & 3;
    } else if (ring2 == 4 * nside) {
      const double wtheta = (theta - theta1) / (kPi - theta1);
      wght[0] *= (1.0 - wtheta);
      wght[1] *= (1.0 - wtheta);
      const double fac = wtheta * 0.25;
      wght[0] += fac;
      wght[1] += fac;
      wght[2] = fac;
      wght[3] = fac;
      pix[2] = ((pix[0] + 2) & 3) + npix - 4;
      pix[3] = ((pix[1] + 2) & 3) + npix - 4;
    } else {
      const double wtheta = (theta - theta1) / (theta2 - theta1);
      wght[0] *= (1.0 - wtheta);
      wght[1] *= (1.0 - wtheta);
      wght[2] *= wtheta;
      wght[3] *= wtheta;
    }
    double wtot = 0.0, res = 0.0;
    for (int i = 0; i < 4; ++i) {
      const double val = map[pix[i]];
      if (val > -1.63749e30) {
        res += val * wght[i];
        wtot += wght[i];
      }
    }
    return (wtot == 0.0) ? kUndef : res / wtot;
  }
};

// accumulate (hamdis.h:698-721)
void accumulate(double *dst, uint64_t nside_dst, const double *src,
                uint64_t nside_src) {
  const Ring rd(nside_dst), rs(nside_src);
  if (rd.npix == rs.npix) {
#pragma omp parallel for
    for (int64_t i = 0; i < (int64_t)rd.npix; ++i)
      dst[i] = dst[i] + src[i];
  } else {
#pragma omp parallel for
    for (int64_t i = 0; i < (int64_t)rd.npix; ++i) {
      double th, ph;
      rd.pix2ang(i, th, ph);
      dst[i] = dst[i] + rs.interpolate(src, th, ph);
    }
  }
}

// ---------------------------------------------------------------- grids --
// Brnd::read_grid / TErnd::read_grid (brnd.cc:25-88, ternd.cc:21-71)
bool cell(const hamx_box &b, const V3 &pos, uint64_t &idx, double &xd,
          double &yd, double &zd) {
  double tmp = (b.nx - 1) * (pos[0] - b.x_min) / (b.x_max - b.x_min);
  if (tmp <= 0 or tmp >= b.nx - 1)
    return false;
  const uint64_t xl = (uint64_t)std::floor(tmp);
  xd = tmp - xl;
  tmp = (b.ny - 1) * (pos[1] - b.y_min) / (b.y_max - b.y_min);
  if (tmp <= 0 or tmp >= b.ny - 1)
    return false;
  const uint64_t yl = (uint64_t)std::floor(tmp);
  yd = tmp - yl;
  tmp = (b.nz - 1) * (pos[2] - b.z_min) / (b.z_max - b.z_min);
  if (tmp <= 0 or tmp >= b.nz - 1)
    return false;
  const uint64_t zl = (uint64_t)std::floor(tmp);
  zd = tmp - zl;
  idx = xl * b.ny * b.nz + yl * b.nz + zl;
  return true;
}

double trilinear(const double *f, const hamx_box &b, uint64_t idx, double xd,
                 double yd, double zd) {
  const uint64_t sy = b.nz, sx = b.ny * b.nz;
  const double i1 = f[idx] * (1. - zd) + f[idx + 1] * zd;
  const double i2 = f[idx + sy] * (1. - zd) + f[idx + sy + 1] * zd;
  const double j1 = f[idx + sx] * (1. - zd) + f[idx + sx + 1] * zd;
  const double j2 = f[idx + sx + sy] * (1. - zd) + f[idx + sx + sy + 1] * zd;
  const double w1 = i1 * (1. - yd) + i2 * yd;
  const double w2 = j1 * (1. - yd) + j2 * yd;
  return w1 * (1. - xd) + w2 * xd;
}

V3 read_grid3(const double *bx, const double *by, const double *bz,
              const hamx_box &b, const V3 &pos) {
  uint64_t idx;
  double xd, yd, zd;
  if (!cell(b, pos, idx, xd, yd, zd))
    return v3(0., 0., 0.);
  return v3(trilinear(bx, b, idx, xd, yd, zd), trilinear(by, b, idx, xd, yd, zd),
            trilinear(bz, b, idx, xd, yd, zd));
}

double read_grid1(const double *f, const hamx_box &b, const V3 &pos) {
  uint64_t idx;
  double xd, yd, zd;
  if (!cell(b, pos, idx, xd, yd, zd))
    return 0.;
  return trilinear(f, b, idx, xd, yd, zd);
}

// --------------------------------------------------------------- models --
// Breg_jf12::write_field (breg_jf12.cc:8-189), restated step by step: disk
// (ring below 5 kpc, 8 spiral regions above, the 8th strength closing the
// flux), toroidal halo, X-field; cylindrical sum, then Cartesian.
V3 breg_jf12(const double *p, const V3 &pos) {
  const double b_arm[7] = {p[0], p[1], p[2], p[3], p[4], p[5], p[6]};
  const double b_ring = p[7], h_disk = p[8], w_disk = p[9];
  const double Bn = p[10], Bs = p[11], rn = p[12], rs = p[13], wh = p[14], z0 = p[15];
  const double B0_X = p[16], Xtheta_const = p[17], rpc_X = p[18], r0_X = p[19];
  const double Rmax = 20 * kKpc, rho_GC = 1. * kKpc;
  const double inc = 11.5, rmin = 5. * kKpc, rcent = 3. * kKpc;
  const double f[8] = {0.130, 0.165, 0.094, 0.122, 0.13, 0.118, 0.084, 0.156};
  const double rc_B[8] = {5.1, 6.3, 7.1, 8.3, 9.8, 11.4, 12.7, 15.5};
  const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
  const double rho = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
  const double phi = std::atan2(pos[1], pos[0]);
  const double z = pos[2];
  if (r > Rmax || rho < rho_GC)
    return v3(1.e-12, 1.e-12, 1.e-12);
  const double B0 = rmin / r;
  const double zprofile = 1. / (1 + std::exp(-2. / w_disk * (std::fabs(z) - h_disk)));
  double disk[3] = {0., 0., 0.};
  if (r > rcent) {
    if (r < rmin) {
      disk[1] = B0 * b_ring * (1 - zprofile);
    } else {
      double bv[8] = {b_arm[0], b_arm[1], b_arm[2], b_arm[3], b_arm[4], b_arm[5], b_arm[6], 0.};
      double b8 = 0.;
      for (int i = 0; i < 7; i++)
        b8 -= f[i] * bv[i] / f[7];
      bv[7] = b8;
      double b_disk = 0.;
      double r_negx = r * std::exp(-1 / std::tan(M_PI / 180. * (90 - inc)) * (phi - M_PI));
      if (r_negx > rc_B[7] * kKpc)
        r_negx = r * std::exp(-1 / std::tan(M_PI / 180. * (90 - inc)) * (phi + M_PI));
      if (r_negx > rc_B[7] * kKpc)
        r_negx = r * std::exp(-1 / std::tan(M_PI / 180. * (90 - inc)) * (phi + 3 * M_PI));
      for (int i = 7; i >= 0; i--)
        if (r_negx < rc_B[i] * kKpc)
          b_disk = bv[i];
      disk[0] = b_disk * B0 * std::sin(M_PI / 180. * inc) * (1 - zprofile);
      disk[1] = b_disk * B0 * std::cos(M_PI / 180. * inc) * (1 - zprofile);
    }
  }
  double b1, rh;
  if (z >= 0) {
    b1 = Bn;
    rh = rn;
  } else {
    b1 = Bs;
    rh = rs;
  }
  const double B_h =
      b1 * (1. - 1. / (1. + std::exp(-2. / wh * (r - rh)))) * std::exp(-(std::fabs(z)) / (z0));
  const double halo[3] = {0., B_h * zprofile, 0.};
  double Xtheta, rp_X, B_X;
  const double r_sign = z < 0 ? -1. : 1.;
  const double rc_X = rpc_X + std::fabs(z) / std::tan(Xtheta_const);
  if (r < rc_X) {
    rp_X = r * rpc_X / rc_X;
    B_X = B0_X * std::pow(rpc_X / rc_X, 2.) * std::exp(-rp_X / r0_X);
    Xtheta = std::atan(std::fabs(z) / (r - rp_X));
    if (z == 0.)
      Xtheta = M_PI / 2.;
  } else {
    Xtheta = Xtheta_const;
    rp_X = r - std::fabs(z) / std::tan(Xtheta);
    B_X = B0_X * rp_X / r * std::exp(-rp_X / r0_X);
  }
  const double xf[3] = {B_X * std::cos(Xtheta) * r_sign, 0., B_X * std::sin(Xtheta)};
  double cyl[3];
  for (int i = 0; i < 3; ++i)
    cyl[i] = disk[i] + halo[i] + xf[i];
  return v3(cyl[0] * std::cos(phi) - cyl[1] * std::sin(phi),
            cyl[0] * std::sin(phi) + cyl[1] * std::cos(phi), cyl[2]);
}

// Breg_jaffe (breg_jaffe.cc:11-243, bfield.h:202-224), restated part by part:
// orientation, radial scaling, height scalings, distance to each arm (or the
// ring / bar), compression, and the sum of write_field.
struct Jaffe {
  const double *p;
  bool quadruple() const { return p[0] != 0.; }
  bool bss() const { return p[1] != 0.; }
  bool ring() const { return p[9] != 0.; }
  bool bar() const { return p[10] != 0.; }
  V3 orientation(const V3 &pos) const {
    const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
    const double r_lim = p[12], bar_lim = p[14] + 0.5 * p[30];
    const double cos_p = std::cos(p[28]), sin_p = std::sin(p[28]);
    V3 t = v3(0., 0., 0.);
    double quad = 1;
    if (r < 0.5 * kKpc)
      return t;
    if (pos[2] > p[3])
      quad = (1 - 2 * static_cast<double>(quadruple()));
    if (ring()) {
      if (r > r_lim) {
        t[0] = (cos_p * (pos[1] / r) - sin_p * (pos[0] / r)) * quad;
        t[1] = (-cos_p * (pos[0] / r) - sin_p * (pos[1] / r)) * quad;
      } else {
        t[0] = (1 - 2 * static_cast<double>(bss())) * pos[1] / r;
        t[1] = (2 * static_cast<double>(bss()) - 1) * pos[0] / r;
      }
    } else if (bar()) {
      const double cos_phi = std::cos(p[16]), sin_phi = std::sin(p[16]);
      const double x = cos_phi * pos[0] - sin_phi * pos[1];
      const double y = sin_phi * pos[0] + cos_phi * pos[1];
      if (r > bar_lim) {
        t[0] = (cos_p * (pos[1] / r) - sin_p * (pos[0] / r)) * quad;
        t[1] = (-cos_p * (pos[0] / r) - sin_p * (pos[1] / r)) * quad;
      } else if (y != 0) {
        const double new_x = std::copysign(1, y);
        const double new_y = -std::copysign(1, y) * (x / y) * p[15] * p[15] / (p[14] * p[14]);
        t[0] = (cos_phi * new_x + sin_phi * new_y) * (1 - 2 * static_cast<double>(bss()));
        t[1] = (-sin_phi * new_x + cos_phi * new_y) * (1 - 2 * static_cast<double>(bss()));
        const double l2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
        if (l2 != 0) {
          const double den = 1. / std::sqrt(l2);
          for (int i = 0; i < 3; ++i)
            t[i] = t[i] * den;
        }
      } else {
        t[0] = (2 * static_cast<double>(bss()) - 1) * std::copysign(1, x) * sin_phi;
        t[1] = (2 * static_cast<double>(bss()) - 1) * std::copysign(1, x) * cos_phi;
      }
    }
    return t;
  }
  double radial_scaling(const V3 &pos) const {
    const double r2 = pos[0] * pos[0] + pos[1] * pos[1];
    const double s1 = 1. - std::exp(-r2 / (p[6] * p[6]));
    const double s2 = std::exp(-r2 / (p[7] * p[7]));
    const double s3 = std::exp(-r2 * r2 / (p[8] * p[8] * p[8] * p[8]));
    return s1 * (s2 + s3);
  }
  double sech2(double z, double z0) const {
    return 1. / (std::cosh(z / z0) * std::cosh(z / z0));
  }
  std::vector<double> dist2arm(const V3 &pos) const {
    std::vector<double> d;
    const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
    const double r_lim = p[12], bar_lim = p[14] + 0.5 * p[30];
    const double cos_p = std::cos(p[28]), sin_p = std::sin(p[28]);
    const double beta_inv = -sin_p / cos_p;
    double theta = std::atan2(pos[1], pos[0]);
    if (theta < 0)
      theta += 2 * M_PI;
    auto arms = [&]() {
      for (int i = 0; i < static_cast<int>(p[17]) && i < 4; ++i) {
        const double d_ang = p[20 + i] - theta;
        const double d_rad = std::fabs(p[18] * std::exp(d_ang * beta_inv) - r);
        const double d_rad_p = std::fabs(p[18] * std::exp((d_ang + 2 * M_PI) * beta_inv) - r);
        const double d_rad_m = std::fabs(p[18] * std::exp((d_ang - 2 * M_PI) * beta_inv) - r);
        d.push_back(std::min(std::min(d_rad, d_rad_p), d_rad_m) * cos_p);
      }
    };
    if (ring()) {
      if (r < r_lim)
        d.push_back(std::fabs(p[12] - r));
      else
        arms();
    } else if (bar()) {
      const double cos_tmp = std::cos(p[16]) * pos[0] / r - std::sin(p[16]) * pos[1] / r;
      const double sin_tmp = std::cos(p[16]) * pos[1] / r + std::sin(p[16]) * pos[0] / r;
      if (r < bar_lim)
        d.push_back(std::fabs(p[14] * p[15] /
                                  std::sqrt(p[14] * p[14] * sin_tmp * sin_tmp +
                                            p[15] * p[15] * cos_tmp * cos_tmp) -
                              r));
      else
        arms();
    }
    return d;
  }
  std::vector<double> arm_compress(const V3 &pos) const {
    const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]) / p[31];
    const double c0 = 1. / p[29] - 1.;
    std::vector<double> a0 = dist2arm(pos);
    const double r_scaling = radial_scaling(pos);
    const double z_scaling = sech2(pos[2], p[19]);
    const double d0_inv = (r_scaling * z_scaling) / p[30];
    const double factor = c0 * r_scaling * z_scaling;
    if (r > 1) {
      const double cdrop = std::pow(r, -p[32]);
      for (auto &a : a0)
        a = factor * cdrop * std::exp(-a * a * cdrop * cdrop * d0_inv * d0_inv);
    } else {
      for (auto &a : a0)
        a = factor * std::exp(-a * a * d0_inv * d0_inv);
    }
    return a0;
  }
  V3 field(const V3 &pos) const {
    const double inner_b = ring() ? p[11] : (bar() ? p[13] : 0.);
    const V3 bhat = orientation(pos);
    const double rs = radial_scaling(pos);
    const double h = p[2] * sech2(pos[2], p[3]) + p[4] * sech2(pos[2], p[5]);
    V3 bt;
    for (int i = 0; i < 3; ++i)
      bt[i] = bhat[i] * rs * h;
    const std::vector<double> arm = arm_compress(pos);
    if (arm.size() == 1) {
      for (int i = 0; i < 3; ++i)
        bt[i] += bhat[i] * arm[0] * inner_b;
    } else {
      for (size_t j = 0; j < arm.size(); ++j)
        for (int i = 0; i < 3; ++i)
          bt[i] += bhat[i] * arm[j] * p[24 + j];
    }
    return bt;
  }
};

// Breg_lsa::write_field (breg_lsa.cc:9-26), Breg_cart, Breg_unif,
// Breg_helix (breg_helix.cc:13-30), Breg_jf12, Breg_jaffe
V3 breg_model(int model, const double *p, const V3 &pos) {
  if (model == HAMX_BREG_LSA) {
    const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
    if (r > (20. * kKpc) or r < (3. * kKpc))
      return v3(0., 0., 0.);
    const double cos_p = pos[0] / r;
    const double sin_p = pos[1] / r;
    const double psi = p[1] + p[2] * std::log(r / (8. * kKpc));
    const double chi = p[3] * std::tanh(pos[2] / (1. * kKpc));
    return v3(p[0] * (std::sin(psi) * std::cos(chi) * cos_p -
                      std::cos(psi) * std::cos(chi) * sin_p),
              p[0] * (std::sin(psi) * std::cos(chi) * sin_p +
                      std::cos(psi) * std::cos(chi) * cos_p),
              p[0] * std::sin(chi));
  }
  if (model == HAMX_BREG_CART)
    return v3(p[0], p[1], p[2]);
  if (model == HAMX_BREG_UNIF)
    return v3(p[0] * std::cos(p[2]), p[0] * std::sin(p[2]), p[1]);
  if (model == HAMX_BREG_HELIX) {
    const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
    const double phi = std::atan2(pos[1], pos[0]) + M_PI / 2.0;
    if ((r > p[3]) && (r < p[4]))
      return v3(p[0] * std::cos(phi), p[1] * std::sin(phi), p[2]);
    return v3(0., 0., 0.);
  }
  if (model == HAMX_BREG_JF12)
    return breg_jf12(p, pos);
  if (model == HAMX_BREG_JAFFE)
    return Jaffe{p}.field(pos);
  return v3(0., 0., 0.);
}

V3 breg_field(int model, const double *p, const hamx_fields *f, const V3 &pos) {
  if (model == HAMX_BREG_GRID)
    return read_grid3(f->breg_x, f->breg_y, f->breg_z, f->breg_box, pos);
  return breg_model(model, p, pos);
}

// CRE_ana / CRE_unif (cre_ana.cc:11-47, cre_unif.cc)
struct Cre {
  int model;
  const double *p;
  V3 obs;
  double flux_idx(const V3 &pos) const {
    if (model == HAMX_CRE_UNIF)
      return -(p[0]);
    const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]) / kKpc;
    const double z = std::fabs(pos[2]) / kKpc;
    return -p[0] + p[1] * r + p[2] * z;
  }
  double spatial_profile(const V3 &pos) const {
    if (model == HAMX_CRE_UNIF)
      return (length(sub(pos, obs)) > p[1]) ? 0 : 1;
    const double R0 = std::sqrt(obs[0] * obs[0] + obs[1] * obs[1]);
    const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
    return std::exp((R0 - r) / p[3]) *
           (1. / (std::cosh(pos[2] / p[4]) * std::cosh(pos[2] / p[4])));
  }
  double flux_norm(const V3 &pos) const {
    const double E0 = model == HAMX_CRE_UNIF ? p[2] : p[5];
    const double j0 = model == HAMX_CRE_UNIF ? p[3] : p[6];
    const double gamma0 = E0 / kMec2;
    const double beta0 = std::sqrt(1. - 1. / gamma0);
    const double unit = 4. * kPi * kMec / (kGeV * kM * kM * kSec * beta0);
    const double norm = j0 * std::pow(gamma0, -flux_idx(obs));
    return norm * unit * spatial_profile(pos);
  }
};

// sync_emissivity_t / _p analytic branch (integrator.cc:333-346,402-415)
double sync_t(const Cre &cre, double freq, const V3 &pos, double bper) {
  const double index = cre.flux_idx(pos);
  const double norm = cre.flux_norm(pos) * 1.73205081 * (kQe * kQe * kQe) *
                      std::fabs(bper) / (4. * kPi * kMec2 * (1. - index));
  const double A = 2. * kPi * freq * kMec / (3. * kQe * std::fabs(bper));
  return norm * (std::pow(A, 0.5 * (index + 1)) *
                 std::tgamma(-0.25 * index + 19. / 12.) *
                 std::tgamma(-0.25 * index - 1. / 12.));
}
double sync_p(const Cre &cre, double freq, const V3 &pos, double bper) {
  const double index = cre.flux_idx(pos);
  const double norm = cre.flux_norm(pos) * 1.73205081 * (kQe * kQe * kQe) *
                      std::fabs(bper) / (16. * kPi * kMec2);
  const double A = 2. * kPi * freq * kMec / (3. * kQe * std::fabs(bper));
  return norm * (std::pow(A, 0.5 * (index + 1)) *
                 std::tgamma(-0.25 * index + 7. / 12.) *
                 std::tgamma(-0.25 * index - 1. / 12.));
}

// TEreg_ymw16 (tereg_ymw16.cc:10-333), restated component by component with
// the parameters named as Param::tereg_ymw16 names them (param.cc:480-641).
struct Ymw16 {
  const double *p;
  double r_warp() const { return p[0]; }
  double r0() const { return p[1]; }
  static double sech_sq(double v) { return std::pow(1. / std::cosh(v), 2); }
  // thick disk
  double thick(double zz, double rr) const {
    const double ad = p[3], bd = p[4], n1 = p[5], h1 = p[6];
    if (zz > 10. * h1)
      return 0.;
    double gd = 1.;
    if (rr > bd)
      gd = sech_sq((rr - bd) / ad);
    return n1 * gd * sech_sq(zz / h1);
  }
  // thin disk
  double thin(double zz, double rr) const {
    const double ad = p[3], bd = p[4], a2 = p[7], b2 = p[8], n2 = p[9], k2 = p[10];
    const double h0 = k2 * (32 * kPc + 1.6e-3 * rr + (4.e-7 / kPc) * rr * rr);
    if (zz > 10. * h0)
      return 0.;
    double gd = 1.;
    if (rr > bd)
      gd = sech_sq((rr - bd) / ad);
    return n2 * gd * sech_sq((rr - b2) / a2) * sech_sq(zz / h0);
  }
  // spiral arms; the loop runs over arms 0..3 only, as the reference's does
  double spiral(double xx, double yy, double zz, double rr) const {
    const double ad = p[3], bd = p[4], b2s = p[11], aa = p[42], ka = p[43];
    const double *narm = p + 12, *warm = p + 17, *rmin = p + 22, *phimin = p + 27;
    const double *tpitch = p + 32, *cpitch = p + 37;
    const double ncn = p[44], thetacn = p[45], wcn = p[46], nsg = p[47], thetasg = p[48],
                 wsg = p[49];
    double scaling = 1.;
    if (rr > bd) {
      if ((rr - bd) > 10. * ad)
        return 0.;
      scaling = sech_sq((rr - bd) / ad);
    }
    const double h0 = ka * (32 * kPc + 1.6e-3 * rr + (4.e-7 / kPc) * std::pow(rr, 2));
    if (zz > 10. * h0)
      return 0.;
    scaling *= sech_sq(zz / h0);
    if ((rr - b2s) > 10. * aa)
      return 0.;
    scaling *= sech_sq((rr - b2s) / aa);
    double theta = std::atan2(yy, xx);
    if (theta < 0)
      theta += 2 * kPi;
    double ne3s = 0.;
    for (int i = 0; i < 4; ++i) {
      double d_phi = theta - phimin[i];
      if (d_phi < 0)
        d_phi += 2. * kPi;
      const double d = std::fabs(rmin[i] * std::exp(d_phi * tpitch[i]) - rr);
      const double d_p = std::fabs(rmin[i] * std::exp((d_phi + 2. * kPi) * tpitch[i]) - rr);
      const double smin = std::min(d, d_p) * cpitch[i];
      if (smin > 10. * warm[i])
        continue;
      if (i != 2) {
        ne3s += narm[i] * scaling * sech_sq(smin / warm[i]);
      } else if (rr > 6 * kKpc && theta * kRad > thetacn) {
        const double ga = (1. - nsg * std::exp(-std::pow((theta * kRad - thetasg) / wsg, 2))) *
                          (1. + ncn) * sech_sq(smin / warm[i]);
        ne3s += narm[i] * scaling * ga;
      } else {
        const double ga = (1. - nsg * std::exp(-std::pow((theta * kRad - thetasg) / wsg, 2))) *
                          (1. + ncn * std::exp(-std::pow((theta * kRad - thetacn) / wcn, 2))) *
                          sech_sq(smin / warm[i]);
        ne3s += narm[i] * scaling * ga;
      }
    }
    return ne3s;
  }
  double galcen(double xx, double yy, double zz) const {
    const double ngc = p[50], agc = p[51], hgc = p[52];
    const double Xgc = 50. * kPc, Ygc = 0., Zgc = -7. * kPc;
    const double R2gc = (xx - Xgc) * (xx - Xgc) + (yy - Ygc) * (yy - Ygc);
    if (R2gc > 10. * agc * agc)
      return 0.;
    const double Ar = std::exp(-R2gc / (agc * agc));
    if (std::fabs(zz - Zgc) > 10. * hgc)
      return 0.;
    return ngc * Ar * sech_sq((zz - Zgc) / hgc);
  }
  double gum(double xx, double yy, double zz) const {
    const double ngn = p[53], wgn = p[54], agn = p[55], kgn = p[56];
    if (yy < 0 or xx > 0)
      return 0.;
    const double lc = 264. * kRad, bc = -4. * kRad, dc = 450. * kPc;
    const double xc = dc * std::cos(bc) * std::sin(lc);
    const double yc = r0() - dc * std::cos(bc) * std::cos(lc);
    const double zc = dc * std::sin(bc);
    const double theta = std::atan2(std::fabs(zz - zc),
                                    std::sqrt((xx - xc) * (xx - xc) + (yy - yc) * (yy - yc)));
    const double tantheta = std::tan(theta);
    const double zp = (agn * kgn) / std::sqrt(1. + kgn * kgn / (tantheta * tantheta));
    const double xyp = zp / tantheta;
    const double xy_dist = std::sqrt(agn * agn - xyp * xyp) * double(agn > xyp);
    const double alpha = std::atan2(kgn * xyp, xy_dist) + theta;
    const double R2 = (xx - xc) * (xx - xc) + (yy - yc) * (yy - yc) + (zz - zc) * (zz - zc);
    const double r2 = zp * zp + xyp * xyp;
    const double D2min = (R2 + r2 - 2. * std::sqrt(R2 * r2)) * std::sin(alpha) * std::sin(alpha);
    if (D2min > 10. * wgn * wgn)
      return 0.;
    return ngn * std::exp(-D2min / (wgn * wgn));
  }
  static double rlb(double xx, double yy, double zz) {
    return std::sqrt(std::pow(((yy - 8.34 * kKpc) * 0.94 - 0.34 * zz), 2) + std::pow(xx, 2));
  }
  double localbubble(double xx, double yy, double zz, double ll, double Rlb) const {
    if (yy < 0)
      return 0.;
    double nel = 0.;
    const double rLB = rlb(xx, yy, zz);
    const double nlb1 = p[58], thetalb1 = p[59], detlb1 = p[60], wlb1 = p[61], hlb1 = p[62];
    const double nlb2 = p[63], thetalb2 = p[64], detlb2 = p[65], wlb2 = p[66], hlb2 = p[67];
    const double dl1 = std::min(std::fabs(ll + 360. - thetalb1), std::fabs(thetalb1 - ll));
    if (dl1 < 10. * detlb1 or (rLB - Rlb) < 10. * wlb1 or zz < 10. * hlb1)
      nel += nlb1 * sech_sq(dl1 / detlb1) * sech_sq((rLB - Rlb) / wlb1) * sech_sq(zz / hlb1);
    const double dl2 = std::min(std::fabs(ll + 360 - thetalb2), std::fabs(thetalb2 - ll));
    if (dl2 < 10. * detlb2 or (rLB - Rlb) < 10. * wlb2 or zz < 10. * hlb2)
      nel += nlb2 * sech_sq(dl2 / detlb2) * sech_sq((rLB - Rlb) / wlb2) * sech_sq(zz / hlb2);
    return nel;
  }
  double nps(double xx, double yy, double zz) const {
    const double nli = p[68], rli = p[69], wli = p[70], detthetali = p[71], thetali = p[72];
    if (yy < 0)
      return 0.;
    const double theta_LI = thetali * kRad;
    const double x_c = -10.156 * kPc, y_c = 8106.207 * kPc, z_c = 10.467 * kPc;
    const double rLI = std::sqrt((xx - x_c) * (xx - x_c) + (yy - y_c) * (yy - y_c) +
                                 (zz - z_c) * (zz - z_c));
    const double theta =
        std::acos(((xx - x_c) * std::cos(theta_LI) + (zz - z_c) * std::sin(theta_LI)) / rLI) /
        kRad;
    if (theta > 10. * detthetali or (rLI - rli) > 10. * wli)
      return 0.;
    return nli * std::exp(-std::pow((rLI - rli) / wli, 2)) *
           std::exp(-std::pow(theta / detthetali, 2));
  }
  double field(const V3 &pos) const {
    V3 g = v3(pos[1], -pos[0], pos[2]);
    const double r_cyl = std::sqrt(g[0] * g[0] + g[1] * g[1]);
    if (r_cyl >= r_warp())
      g.e[2] -= p[2] * (r_cyl - r_warp()) * std::cos(std::atan2(g[1], g[0]));
    if (length(g) > 25 * kKpc)
      return 0.;
    double comp[8] = {0.};
    double w_lb = 0., w_gum = 0., w_loop = 0.;
    const double ec_l = std::atan2(g[0], r0() - g[1]) / kRad;
    comp[1] = thick(g[2], r_cyl);
    comp[2] = thin(g[2], r_cyl);
    comp[3] = spiral(g[0], g[1], g[2], r_cyl);
    comp[4] = galcen(g[0], g[1], g[2]);
    comp[5] = gum(g[0], g[1], g[2]);
    const double Rlb = 110. * kPc;
    comp[6] = localbubble(g[0], g[1], g[2], ec_l, Rlb);
    comp[7] = nps(g[0], g[1], g[2]);
    comp[0] = comp[1] + std::max(comp[2], comp[3]);
    if (rlb(g[0], g[1], g[2]) < Rlb) {
      comp[0] = p[57] * comp[1] + std::max(comp[2], comp[3]);
      if (comp[6] > comp[0])
        w_lb = 1;
    } else if (comp[6] > comp[0] and comp[6] > comp[5]) {
      w_lb = 1;
    }
    if (comp[7] > comp[0])
      w_loop = 1;
    if (comp[5] > comp[0])
      w_gum = 1;
    return (1 - w_lb) * ((1 - w_gum) * ((1 - w_loop) * (comp[0] + comp[4]) + w_loop * comp[7]) +
                         w_gum * comp[5]) +
           w_lb * comp[6];
  }
};

// ------------------------------------------------ CRE grid (CRE_num) --
// The synchrotron kernels the reference takes from GSL
// (gsl_sf_synchrotron_1 / _2, integrator.cc:321,390), restated from their
// definitions F(x) = x int_x^inf K_5/3(t) dt and G(x) = x K_2/3(x) through
// K_nu(t) = int_0^inf exp(-t cosh s) cosh(nu s) ds, so that
//   F(x) = x int_0^inf cosh(5s/3) / cosh(s) exp(-x cosh s) ds,
//   G(x) = x int_0^inf cosh(2s/3) exp(-x cosh s) ds,
// each summed by the trapezoid rule.  The integrands are analytic in a strip
// about the real axis, so the rule converges exponentially; the step shrinks
// as 1/sqrt(1+x) with the width of the peak at s = 0, and the sum stops past
// the peak once a term is below 1e-18 of it.
double sync_kernel(double x, double nu, bool over_cosh) {
  const double h = 0.05 / std::sqrt(1. + x);
  double sum = 0.5 * std::exp(-x);
  for (int k = 1; k < 4000000; ++k) {
    const double s = k * h;
    const double c = std::cosh(s);
    double w = std::cosh(nu * s) * std::exp(-x * c);
    if (over_cosh)
      w /= c;
    sum += w;
    if (x * c > 1. && w < 1e-18 * sum)
      break;
  }
  return x * h * sum;
}
double sync_F(double x) { return sync_kernel(x, 5. / 3., true); }
double sync_G(double x) { return sync_kernel(x, 2. / 3., false); }

struct CreGrid {
  const double *f;
  hamx_box box;
  double E_min, E_fact;
  uint64_t nE;
  // CREfield::read_grid_num (cre.cc:144-199): the trilinear gather of energy
  // bin e, with toolkit::index4d (toolkit.h:38-44) called as
  // index4d(nx, ny, nz, nE, x, y, z, e) = ((x * ny + y) * nz + z) * nE + e
  // (the first extent is unused).
  double read_num(const V3 &pos, uint64_t e) const {
    uint64_t idx;
    double xd, yd, zd;
    if (!cell(box, pos, idx, xd, yd, zd))
      return 0.;
    const uint64_t xl = idx / (box.ny * box.nz), yl = (idx / box.nz) % box.ny, zl = idx % box.nz;
    auto at = [&](uint64_t x, uint64_t y, uint64_t z) {
      return f[x * box.ny * box.nz * nE + y * box.nz * nE + z * nE + e];
    };
    const double i1 = at(xl, yl, zl) * (1. - zd) + at(xl, yl, zl + 1) * zd;
    const double i2 = at(xl, yl + 1, zl) * (1 - zd) + at(xl, yl + 1, zl + 1) * zd;
    const double j1 = at(xl + 1, yl, zl) * (1 - zd) + at(xl + 1, yl, zl + 1) * zd;
    const double j2 = at(xl + 1, yl + 1, zl) * (1 - zd) + at(xl + 1, yl + 1, zl + 1) * zd;
    const double w1 = i1 * (1 - yd) + i2 * yd;
    const double w2 = j1 * (1 - yd) + j2 * yd;
    return w1 * (1 - xd) + w2 * xd;
  }
};

// sync_emissivity_t / _p grid branch (integrator.cc:286-325 and 355-394):
// the midpoint rule over the energy bins, skipping bins whose mean x
// exceeds 100, with the flux over beta = sqrt(1 - mec2/E) averaged at the bin
// ends; `pol` selects G and the J_pol form of x_fact.
double sync_grid(const CreGrid &g, double freq, const V3 &pos, double bper, bool pol) {
  const uint64_t nE = g.nE;
  std::vector<double> KE(nE), x(nE), beta(nE);
  const double x_fact =
      pol ? (2. * kMec * kMec2 * kMec2 * 2. * kPi * freq) / (3. * kQe * std::fabs(bper))
          : 4. * kPi * freq * kMec * kMec2 * kMec2 / (3. * kQe * std::fabs(bper));
  for (uint64_t i = 0; i != nE; ++i) {
    KE[i] = g.E_min * std::exp(i * g.E_fact);
    x[i] = x_fact / (KE[i] * KE[i]);
    beta[i] = std::sqrt(1 - kMec2 / KE[i]);
  }
  double J = 0;
  for (uint64_t i = 0; i + 1 < nE; ++i) {
    const double xv = 0.5 * (x[i + 1] + x[i]);
    if (xv > 100)
      continue;
    const double dE = std::fabs(KE[i + 1] - KE[i]);
    const double flux = 0.5 * (g.read_num(pos, i + 1) / beta[i + 1] + g.read_num(pos, i) / beta[i]);
    J += (pol ? sync_G(xv) : sync_F(xv)) * flux * dE;
  }
  const double fore_factor = 1.73205081 * kQe * kQe * kQe * std::fabs(bper) /
                             (kMec2 * kC * kGeV * kM * kM * kSec);
  return J * fore_factor;
}

// CRE_ana / CRE_unif::write_field (cre_ana.cc:52-63, cre_unif.cc:48-59)
double cre_write_field(const Cre &cre, const V3 &pos, double En) {
  const double E0 = cre.model == HAMX_CRE_UNIF ? cre.p[2] : cre.p[5];
  const double j0 = cre.model == HAMX_CRE_UNIF ? cre.p[3] : cre.p[6];
  const double gamma = En / kMec2;
  const double gamma0 = E0 / kMec2;
  const double beta_ratio = std::sqrt((1. - 1. / gamma) / (1. - 1. / gamma0));
  const double norm = j0 * std::pow(gamma0, -cre.flux_idx(cre.obs));
  return norm * beta_ratio * std::pow(gamma, cre.flux_idx(pos)) * cre.spatial_profile(pos);
}

// ----------------------------------------------------------- integrator --
V3 los_versor(double the, double phi) {
  return v3(std::cos(phi) * std::sin(the), std::sin(phi) * std::sin(the),
            std::cos(the));
}
double sync_ipa(const V3 &b, double the, double phi) {
  const V3 et = v3(std::cos(the) * std::cos(phi), std::cos(the) * std::sin(phi),
                   -std::sin(the));
  const V3 ep = v3(-std::sin(phi), std::cos(phi), 0.);
  return std::atan2(-dot(et, b), -dot(ep, b));
}
double temp_convert(double t, double freq) {
  const double p = kH * freq / (kKB * 2.725);
  return t * (std::exp(p) - 1.) * (std::exp(p) - 1.) / (p * p * std::exp(p));
}

struct Shell {
  uint64_t step;
  double d_start, d_stop, delta;
  std::vector<double> dist;
};
// assemble_shell_ref (integrator.cc:231-244)
Shell shell_ref(const hamx_los_desc *d, uint32_t s) {
  Shell sh;
  sh.d_start = d->radii_shell[s];
  sh.d_stop = d->radii_shell[s + 1];
  sh.delta = d->oc_r_res;
  const double st = std::floor(sh.d_stop / sh.delta - sh.d_start / sh.delta);
  sh.step = st > 0 ? (uint64_t)st : 0;
  sh.delta = (sh.d_stop - sh.d_start) / sh.step;
  sh.dist.reserve(sh.step);
  for (uint64_t i = 0; i < sh.step; ++i)
    sh.dist.push_back(sh.d_start + (i + 0.5) * sh.delta);
  return sh;
}

// radial_integration (integrator.cc:136-228)
void radial(const hamx_los_desc *d, const hamx_fields *f, const Shell &sh,
            double the, double phi, double fd_inner, double o[5]) {
  const V3 obs = v3(d->observer[0], d->observer[1], d->observer[2]);
  const bool do_fd_any = d->do_fd || d->do_sync;
  double inner = 0.;
  if (do_fd_any)
    inner = fd_inner;
  double dm = 0., fd = 0., is = 0., qs = 0., us = 0.;
  const V3 l = los_versor(the, phi);
  double lambda_square = 0, i2bt = 0.;
  const double fd_fore = -(kQe * kQe * kQe) / (2. * kPi * kMec2 * kMec2);
  if (d->do_sync) {
    lambda_square = (kC / d->sync_freq) * (kC / d->sync_freq);
    i2bt = kC * kC / (2. * kKB * d->sync_freq * d->sync_freq);
  }
  Cre cre;
  cre.model = d->cre_model;
  cre.p = d->cre;
  cre.obs = obs;
  CreGrid cg{};
  if (d->cre_model == HAMX_CRE_GRID)
    cg = CreGrid{f->cre, f->cre_box, f->cre_E_min, f->cre_E_fact, f->cre_nE};
  const bool has_brnd = f && f->brnd_x && f->brnd_y && f->brnd_z;
  for (uint64_t k = 0; k < sh.step; ++k) {
    const V3 oc = mul(l, sh.dist[k]);
    const V3 pos = add(oc, obs);
    if (length(pos) < d->gc_r_min)
      continue;
    if (length(pos) > d->gc_r_max)
      continue;
    if (pos[2] < d->gc_z_min)
      continue;
    if (pos[2] > d->gc_z_max)
      continue;
    V3 B = breg_field(d->breg_model, d->breg, f, pos);
    if (has_brnd)
      B = add(B, read_grid3(f->brnd_x, f->brnd_y, f->brnd_z, f->brnd_box, pos));
    const double b_par = dot(B, l);
    const double b_per = length(cross(B, l));
    double te = 0.;
    if (d->tereg_model == HAMX_TEREG_UNIF)
      te = (length(sub(pos, obs)) > d->tereg[1]) ? 0. : d->tereg[0];
    else if (d->tereg_model == HAMX_TEREG_GRID)
      te = read_grid1(f->tereg, f->tereg_box, pos);
    else if (d->tereg_model == HAMX_TEREG_YMW16)
      te = Ymw16{d->tereg}.field(pos);
    if (f && f->ternd)
      te += read_grid1(f->ternd, f->ternd_box, pos);
    te *= double(te > 0.);
    if (d->do_dm)
      dm += te * sh.delta;
    if (do_fd_any)
      fd += te * b_par * fd_fore * sh.delta;
    if (d->do_sync) {
      const bool grid = d->cre_model == HAMX_CRE_GRID;
      const double jt = (grid ? sync_grid(cg, d->sync_freq, pos, b_per, false)
                              : sync_t(cre, d->sync_freq, pos, b_per)) *
                        sh.delta * i2bt;
      const double jp = (grid ? sync_grid(cg, d->sync_freq, pos, b_per, true)
                              : sync_p(cre, d->sync_freq, pos, b_per)) *
                        sh.delta * i2bt;
      is += jt;
      const double qui = (inner + fd) * lambda_square + sync_ipa(B, the, phi);
      qs += std::cos(2. * qui) * jp;
      us += std::sin(2. * qui) * jp;
    }
  }
  o[0] = dm;
  o[1] = fd;
  o[2] = is;
  o[3] = qs;
  o[4] = us;
}

bool pow2(uint64_t n) { return n && !(n & (n - 1)); }

} // namespace

// ================================================================= C API ==
extern "C" {

void orc_pix2ang(uint64_t nside, uint64_t ipix, double *theta, double *phi) {
  Ring(nside).pix2ang(ipix, *theta, *phi);
}

void orc_pix2ang_map(uint64_t nside, double *tp) {
  const Ring r(nside);
  for (uint64_t i = 0; i < r.npix; ++i)
    r.pix2ang(i, tp[2 * i], tp[2 * i + 1]);
}

double orc_interpolate(const double *map, uint64_t nside, double theta,
                       double phi) {
  return Ring(nside).interpolate(map, theta, phi);
}

void orc_accumulate(double *dst, uint64_t nside_dst, const double *src,
                    uint64_t nside_src) {
  accumulate(dst, nside_dst, src, nside_src);
}

void orc_shell_ref(double d_start, double d_stop, double res, uint64_t *step,
                   double *delta) {
  hamx_los_desc d;
  std::memset(&d, 0, sizeof(d));
  d.radii_shell[0] = d_start;
  d.radii_shell[1] = d_stop;
  d.oc_r_res = res;
  const Shell s = shell_ref(&d, 0);
  *step = s.step;
  *delta = s.delta;
}

void orc_los_versor(double theta, double phi, double out[3]) {
  const V3 l = los_versor(theta, phi);
  for (int i = 0; i < 3; ++i)
    out[i] = l[i];
}

double orc_los_parproj(const double b[3], const double l[3]) {
  return dot(v3(b[0], b[1], b[2]), v3(l[0], l[1], l[2]));
}

double orc_los_perproj(const double b[3], const double l[3]) {
  return length(cross(v3(b[0], b[1], b[2]), v3(l[0], l[1], l[2])));
}

double orc_sync_ipa(const double b[3], double theta, double phi) {
  return sync_ipa(v3(b[0], b[1], b[2]), theta, phi);
}

double orc_temp_convert(double t, double freq) { return temp_convert(t, freq); }

// integrator.h:109-118
int orc_check_upper(double value, double limit) { return value > limit; }
int orc_check_lower(double value, double limit) { return value < limit; }

void orc_breg(int model, const double *par, const double pos[3], double out[3]) {
  const V3 b = breg_model(model, par, v3(pos[0], pos[1], pos[2]));
  for (int i = 0; i < 3; ++i)
    out[i] = b[i];
}

double orc_tereg_unif(const double *par, const double obs[3], const double pos[3]) {
  const V3 d = sub(v3(pos[0], pos[1], pos[2]), v3(obs[0], obs[1], obs[2]));
  return (length(d) > par[1]) ? 0. : par[0];
}

void orc_read_grid3(const double *bx, const double *by, const double *bz,
                    const hamx_box *box, const double pos[3], double out[3]) {
  const V3 b = read_grid3(bx, by, bz, *box, v3(pos[0], pos[1], pos[2]));
  for (int i = 0; i < 3; ++i)
    out[i] = b[i];
}

double orc_read_grid1(const double *f, const hamx_box *box, const double pos[3]) {
  return read_grid1(f, *box, v3(pos[0], pos[1], pos[2]));
}

void orc_sync_emissivity(const hamx_los_desc *d, const double pos[3],
                         double b_per, double *j_tot, double *j_pol) {
  Cre cre;
  cre.model = d->cre_model;
  cre.p = d->cre;
  cre.obs = v3(d->observer[0], d->observer[1], d->observer[2]);
  const V3 p = v3(pos[0], pos[1], pos[2]);
  *j_tot = sync_t(cre, d->sync_freq, p, b_per);
  *j_pol = sync_p(cre, d->sync_freq, p, b_per);
}

void orc_sync_kernels(const double *x, uint64_t n, double *F, double *G) {
  for (uint64_t i = 0; i < n; ++i) {
    F[i] = sync_F(x[i]);
    G[i] = sync_G(x[i]);
  }
}

void orc_sync_emissivity_grid(const hamx_los_desc *d, const hamx_fields *f,
                              const double pos[3], double b_per, double *j_tot,
                              double *j_pol) {
  const CreGrid g{f->cre, f->cre_box, f->cre_E_min, f->cre_E_fact, f->cre_nE};
  const V3 p = v3(pos[0], pos[1], pos[2]);
  *j_tot = sync_grid(g, d->sync_freq, p, b_per, false);
  *j_pol = sync_grid(g, d->sync_freq, p, b_per, true);
}

double orc_tereg_ymw16(const double *p, const double pos[3]) {
  return Ymw16{p}.field(v3(pos[0], pos[1], pos[2]));
}

double orc_cre_read_grid_num(const hamx_fields *f, const double pos[3], uint64_t e) {
  const CreGrid g{f->cre, f->cre_box, f->cre_E_min, f->cre_E_fact, f->cre_nE};
  return g.read_num(v3(pos[0], pos[1], pos[2]), e);
}

// CREfield::write_grid (cre.cc:204-225)
int orc_cre_write_grid(const hamx_los_desc *d, const hamx_box *box, double E_min,
                       double E_fact, uint32_t nE, double *out) {
  if (d->cre_model != HAMX_CRE_ANA && d->cre_model != HAMX_CRE_UNIF)
    return 1;
  Cre cre;
  cre.model = d->cre_model;
  cre.p = d->cre;
  cre.obs = v3(d->observer[0], d->observer[1], d->observer[2]);
  const double lx = box->x_max - box->x_min;
  const double ly = box->y_max - box->y_min;
  const double lz = box->z_max - box->z_min;
  for (uint64_t i = 0; i != box->nx; ++i) {
    for (uint64_t j = 0; j != box->ny; ++j) {
      for (uint64_t k = 0; k != box->nz; ++k) {
        const V3 pos = v3(lx * i / (box->nx - 1) + box->x_min, ly * j / (box->ny - 1) + box->y_min,
                          lz * k / (box->nz - 1) + box->z_min);
        for (uint64_t m = 0; m != nE; ++m) {
          const double E = E_min * std::exp(m * E_fact);
          out[((i * box->ny + j) * box->nz + k) * nE + m] = cre_write_field(cre, pos, E);
        }
      }
    }
  }
  return 0;
}

void orc_radial_integration(const hamx_los_desc *d, const hamx_fields *f,
                            uint32_t shell, double theta, double phi,
                            double fd_inner, double out[5]) {
  const Shell sh = shell_ref(d, shell);
  radial(d, f, sh, theta, phi, fd_inner, out);
}

// Hampisk::duplicate (hamsk.h:55-103): the pivot mask at nside `piv_nside`
// brought to nside `nside`.
static std::vector<double> mask_duplicate(const double *piv, uint64_t piv_nside,
                                          uint64_t nside) {
  const Ring rt(nside), rp(piv_nside);
  std::vector<double> out(rt.npix, 0.);
  if (piv_nside == nside) {
    out.assign(piv, piv + rt.npix);
  } else if (piv_nside > nside) {
    const int64_t fact = (int64_t)(piv_nside / nside);
    for (uint64_t m = 0; m < rt.npix; ++m) {
      int64_t x, y;
      int f;
      rt.ring2xyf(m, x, y, f);
      bool hit = false;
      for (int64_t j = fact * y; j < fact * (y + 1) && !hit; ++j)
        for (int64_t i = fact * x; i < fact * (x + 1) && !hit; ++i)
          hit = piv[rp.xyf2ring(i, j, f)] > 0;
      out[m] = hit ? 1. : 0.;
    }
  } else { // upgrade: every child pixel takes its parent's value
    const int64_t fact = (int64_t)(nside / piv_nside);
    for (uint64_t m = 0; m < rp.npix; ++m) {
      int64_t x, y;
      int f;
      rp.ring2xyf(m, x, y, f);
      for (int64_t j = fact * y; j < fact * (y + 1); ++j)
        for (int64_t i = fact * x; i < fact * (x + 1); ++i)
          out[rt.xyf2ring(i, j, f)] = piv[m];
    }
  }
  return out;
}

// Integrator::write_grid (integrator.cc:23-134) for one sync entry.
int orc_write_grid(const hamx_los_desc *d, const hamx_fields *f, hamx_maps *out,
                   int nthreads) {
  if (!d || !out || d->n_shell < 1 || d->n_shell > HAMX_MAX_SHELL)
    return HAMX_ERR_ARG;
  for (uint32_t s = 0; s < d->n_shell; ++s)
    if (!pow2(d->nside_shell[s]))
      return HAMX_ERR_NSIDE;
  if (d->do_sync && d->cre_model == HAMX_CRE_NONE)
    return HAMX_ERR_ARG;
  if (d->cre_model == HAMX_CRE_GRID && (!f || !f->cre || f->cre_nE < 2))
    return HAMX_ERR_ARG;
#ifdef _OPENMP
  if (nthreads > 0)
    omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  // Grid_obs::build_grid (grid_obs.cc:21-43): maps start at zero; the FD map
  // has nside_fd when do_fd, else nside_sync.
  const uint64_t nside_fd = d->do_fd ? d->nside_fd : d->nside_sync;
  std::vector<double> dm_map, is_map, qs_map, us_map, fd_map;
  if (d->do_dm)
    dm_map.assign(Ring(d->nside_dm).npix, 0.);
  if (d->do_sync) {
    const uint64_t n = Ring(d->nside_sync).npix;
    is_map.assign(n, 0.);
    qs_map.assign(n, 0.);
    us_map.assign(n, 0.);
  }
  if (d->do_fd || d->do_sync)
    fd_map.assign(Ring(nside_fd).npix, 0.);
  const Ring rfd(d->do_fd || d->do_sync ? nside_fd : 1);
  for (uint32_t s = 0; s < d->n_shell; ++s) {
    const uint64_t nside = d->nside_shell[s];
    const Ring rs(nside);
    const Shell sh = shell_ref(d, s);
    std::vector<double> t_dm(d->do_dm ? rs.npix : 0), t_is, t_qs, t_us, t_fd;
    if (d->do_sync) {
      t_is.assign(rs.npix, 0.);
      t_qs.assign(rs.npix, 0.);
      t_us.assign(rs.npix, 0.);
    }
    if (d->do_fd || d->do_sync)
      t_fd.assign(rs.npix, 0.);
    std::vector<double> msk; // integrator.cc:36-39
    if (f && f->mask)
      msk = mask_duplicate(f->mask, f->nside_mask, nside);
#pragma omp parallel for schedule(dynamic)
    for (int64_t ipix = 0; ipix < (int64_t)rs.npix; ++ipix) {
      if (!msk.empty() && msk[ipix] != 1.0)
        continue; // masked: every observable of the pixel stays 0 (integrator.cc:63-70)
      double the, phi;
      rs.pix2ang(ipix, the, phi);
      double fd_inner = 0.;
      if (d->do_fd || d->do_sync)
        fd_inner = rfd.interpolate(fd_map.data(), the, phi);
      double o[5];
      radial(d, f, sh, the, phi, fd_inner, o);
      if (d->do_dm)
        t_dm[ipix] = o[0];
      if (d->do_sync) {
        t_is[ipix] = temp_convert(o[2], d->sync_freq);
        t_qs[ipix] = temp_convert(o[3], d->sync_freq);
        t_us[ipix] = temp_convert(o[4], d->sync_freq);
      }
      if (d->do_fd || d->do_sync)
        t_fd[ipix] = o[1];
    }
    if (d->do_dm)
      accumulate(dm_map.data(), d->nside_dm, t_dm.data(), nside);
    if (d->do_sync) {
      accumulate(is_map.data(), d->nside_sync, t_is.data(), nside);
      accumulate(qs_map.data(), d->nside_sync, t_qs.data(), nside);
      accumulate(us_map.data(), d->nside_sync, t_us.data(), nside);
    }
    if (d->do_fd)
      accumulate(fd_map.data(), nside_fd, t_fd.data(), nside);
  }
  auto put = [](double *dst, const std::vector<double> &v) {
    if (dst && !v.empty())
      std::memcpy(dst, v.data(), v.size() * sizeof(double));
  };
  put(out->dm, dm_map);
  put(out->is, is_map);
  put(out->qs, qs_map);
  put(out->us, us_map);
  put(out->fd, fd_map);
  return HAMX_OK;
}

void orc_ring2xyf(uint64_t nside, uint64_t pix, int64_t *ix, int64_t *iy, int *face) {
  Ring(nside).ring2xyf(pix, *ix, *iy, *face);
}

uint64_t orc_xyf2ring(uint64_t nside, int64_t ix, int64_t iy, int face) {
  return Ring(nside).xyf2ring(ix, iy, face);
}

void orc_mask_duplicate(const double *piv, uint64_t piv_nside, uint64_t nside, double *out) {
  const std::vector<double> v = mask_duplicate(piv, piv_nside, nside);
  std::memcpy(out, v.data(), v.size() * sizeof(double));
}

} // extern "C"
