// Regular-field models shared by the device march and the C++ host (grid
// export, host-side field queries): compiled as __host__ __device__ by hipcc
// and as plain C++ by the host compiler.  Positions are galactocentric [cm],
// fields in Gauss; p follows the HAMX_BREG_* parameter layouts of hamx.h.
#pragma once

#include <cmath>

#include "hamx_cgs.h"

#if defined(__HIPCC__)
#define HAMX_HD __host__ __device__
#else
#define HAMX_HD
#endif

namespace hamx {

// Breg_helix::write_field (breg_helix.cc:13-30): (bx cos phi', by sin phi',
// bz) with phi' = atan2(y, x) + pi/2 strictly inside r_min < r < r_max.
HAMX_HD inline void breg_helix_eval(const double *p, double x, double y, double /*z*/,
                                    double b[3]) {
  b[0] = b[1] = b[2] = 0.;
  const double r = sqrt(x * x + y * y);
  if ((r > p[3]) && (r < p[4])) {
    const double phi = atan2(y, x) + cgs::pi / 2.0;
    b[0] = p[0] * cos(phi);
    b[1] = p[1] * sin(phi);
    b[2] = p[2];
  }
}

// Breg_jf12::write_field (breg_jf12.cc:8-189): the Jansson & Farrar (2012)
// disk (a molecular ring below 5 kpc, 8 logarithmic-spiral regions above
// with the 8th strength fixed by zero net flux), toroidal halo and X-field,
// summed in cylindrical components.  Outside r < 20 kpc and |pos| > 1 kpc
// every component is 1e-12 G, as in the reference.
HAMX_HD inline void breg_jf12_eval(const double *p, double x, double y, double z,
                                   double b[3]) {
  const double kpc = cgs::kpc, pi = cgs::pi;
  const double r = sqrt(x * x + y * y);
  const double rho = sqrt(x * x + y * y + z * z);
  const double phi = atan2(y, x);
  if (r > 20. * kpc || rho < 1. * kpc) {
    b[0] = b[1] = b[2] = 1.e-12;
    return;
  }
  const double inc = 11.5;      // spiral pitch, degrees
  const double rmin = 5. * kpc; // the ring / spiral boundary
  const double B0 = rmin / r;   // 1 at 5 kpc
  const double zprofile = 1. / (1 + exp(-2. / p[9] * (fabs(z) - p[8])));
  double d_r = 0., d_p = 0.; // disk, cylindrical (r, phi)
  if (r > 3. * kpc) {
    if (r < rmin) {
      d_p = B0 * p[7] * (1 - zprofile);
    } else {
      // circumference fractions and the radii where the arm boundaries
      // cross the negative x axis
      const double f[8] = {0.130, 0.165, 0.094, 0.122, 0.13, 0.118, 0.084, 0.156};
      const double rc[8] = {5.1, 6.3, 7.1, 8.3, 9.8, 11.4, 12.7, 15.5};
      double bv[8];
      double b8 = 0.;
      for (int i = 0; i < 7; ++i) {
        bv[i] = p[i];
        b8 -= f[i] * bv[i] / f[7];
      }
      bv[7] = b8;
      const double k = -1 / tan(pi / 180. * (90 - inc));
      double r_negx = r * exp(k * (phi - pi));
      if (r_negx > rc[7] * kpc)
        r_negx = r * exp(k * (phi + pi));
      if (r_negx > rc[7] * kpc)
        r_negx = r * exp(k * (phi + 3 * pi));
      double b_disk = 0.;
      for (int i = 7; i >= 0; --i)
        if (r_negx < rc[i] * kpc)
          b_disk = bv[i];
      d_r = b_disk * B0 * sin(pi / 180. * inc) * (1 - zprofile);
      d_p = b_disk * B0 * cos(pi / 180. * inc) * (1 - zprofile);
    }
  }
  // toroidal halo, north / south
  const double b1 = z >= 0 ? p[10] : p[11];
  const double rh = z >= 0 ? p[12] : p[13];
  const double B_h =
      b1 * (1. - 1. / (1. + exp(-2. / p[14] * (r - rh)))) * exp(-(fabs(z)) / (p[15]));
  const double h_p = B_h * zprofile;
  // X-field: varying elevation angle inside rc_X, constant outside
  const double xt0 = p[17], rpc = p[18];
  const double rc_X = rpc + fabs(z) / tan(xt0);
  double xt, rp, B_X;
  if (r < rc_X) {
    rp = r * rpc / rc_X;
    const double q = rpc / rc_X;
    B_X = p[16] * (q * q) * exp(-rp / p[19]);
    xt = atan(fabs(z) / (r - rp));
    if (z == 0.)
      xt = pi / 2.;
  } else {
    xt = xt0;
    rp = r - fabs(z) / tan(xt);
    B_X = p[16] * rp / r * exp(-rp / p[19]);
  }
  const double x_r = B_X * cos(xt) * (z < 0 ? -1. : 1.);
  const double x_z = B_X * sin(xt);
  const double b_r = d_r + 0. + x_r;
  const double b_p = d_p + h_p + 0.;
  const double b_z = 0. + 0. + x_z;
  const double cp = cos(phi), sp = sin(phi);
  b[0] = b_r * cp - b_p * sp;
  b[1] = b_r * sp + b_p * cp;
  b[2] = b_z;
}

// Breg_jaffe::write_field (breg_jaffe.cc:11-243, scalings bfield.h:202-224):
// orientation (spiral arms at pitch arm_pitch, with a molecular ring or an
// elliptical bar inside), radial and vertical scalings, and per-arm (or
// ring / bar) compression with Gaussian cross-sections.  As in the reference,
// a single compression factor (the ring / bar region, or a one-arm model)
// takes the ring / bar amplitude.
HAMX_HD inline void breg_jaffe_eval(const double *p, double x, double y, double z,
                                    double b[3]) {
  const bool quadruple = p[0] != 0., bss = p[1] != 0., ring = p[9] != 0., bar = p[10] != 0.;
  const double r2 = x * x + y * y;
  const double r = sqrt(x * x + y * y);
  const double cos_p = cos(p[28]), sin_p = sin(p[28]);
  const double r_lim = p[12], bar_lim = p[14] + 0.5 * p[30];
  // orientation
  double o[3] = {0., 0., 0.};
  if (!(r < 0.5 * cgs::kpc)) {
    const double qs = (z > p[3]) ? (1 - 2 * static_cast<double>(quadruple)) : 1.;
    const double s_bss = 1 - 2 * static_cast<double>(bss);
    if (ring) {
      if (r > r_lim) {
        o[0] = (cos_p * (y / r) - sin_p * (x / r)) * qs;
        o[1] = (-cos_p * (x / r) - sin_p * (y / r)) * qs;
      } else {
        o[0] = s_bss * y / r;
        o[1] = (2 * static_cast<double>(bss) - 1) * x / r;
      }
    } else if (bar) {
      const double cf = cos(p[16]), sf = sin(p[16]);
      const double xb = cf * x - sf * y, yb = sf * x + cf * y;
      if (r > bar_lim) {
        o[0] = (cos_p * (y / r) - sin_p * (x / r)) * qs;
        o[1] = (-cos_p * (x / r) - sin_p * (y / r)) * qs;
      } else if (yb != 0) {
        const double nx = copysign(1., yb);
        const double ny = -copysign(1., yb) * (xb / yb) * p[15] * p[15] / (p[14] * p[14]);
        o[0] = (cf * nx + sf * ny) * s_bss;
        o[1] = (-sf * nx + cf * ny) * s_bss;
        const double l2 = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
        if (l2 != 0) {
          const double inv = 1. / sqrt(l2);
          o[0] = o[0] * inv;
          o[1] = o[1] * inv;
          o[2] = o[2] * inv;
        }
      } else {
        o[0] = (2 * static_cast<double>(bss) - 1) * copysign(1., xb) * sf;
        o[1] = (2 * static_cast<double>(bss) - 1) * copysign(1., xb) * cf;
      }
    }
  }
  // radial and vertical scalings
  const double s1 = 1. - exp(-r2 / (p[6] * p[6]));
  const double s2 = exp(-r2 / (p[7] * p[7]));
  const double s3 = exp(-r2 * r2 / (p[8] * p[8] * p[8] * p[8]));
  const double rs = s1 * (s2 + s3);
  const double ch_d = cosh(z / p[3]), ch_h = cosh(z / p[5]), ch_a = cosh(z / p[19]);
  const double ds = 1. / (ch_d * ch_d), hs = 1. / (ch_h * ch_h), zs = 1. / (ch_a * ch_a);
  const double amp = p[2] * ds + p[4] * hs;
  for (int c = 0; c < 3; ++c)
    b[c] = o[c] * rs * amp;
  // distances to the arms (or the ring / bar), then their compression
  double dist[4];
  int n = 0;
  const int arm_num = static_cast<int>(p[17]);
  if (ring || bar) {
    const bool inner = ring ? (r < r_lim) : (r < bar_lim);
    if (inner) {
      if (ring) {
        dist[0] = fabs(p[12] - r);
      } else {
        const double cf = cos(p[16]), sf = sin(p[16]);
        const double cos_t = cf * x / r - sf * y / r;
        const double sin_t = cf * y / r + sf * x / r;
        dist[0] = fabs(p[14] * p[15] /
                           sqrt(p[14] * p[14] * sin_t * sin_t + p[15] * p[15] * cos_t * cos_t) -
                       r);
      }
      n = 1;
    } else {
      const double beta_inv = -sin_p / cos_p;
      double theta = atan2(y, x);
      if (theta < 0)
        theta += 2 * cgs::pi;
      for (int i = 0; i < arm_num && i < 4; ++i) {
        const double d_ang = p[20 + i] - theta;
        const double d0 = fabs(p[18] * exp(d_ang * beta_inv) - r);
        const double dp = fabs(p[18] * exp((d_ang + 2 * cgs::pi) * beta_inv) - r);
        const double dm = fabs(p[18] * exp((d_ang - 2 * cgs::pi) * beta_inv) - r);
        const double m = d0 < dp ? d0 : dp;
        dist[n++] = (m < dm ? m : dm) * cos_p;
      }
    }
  }
  if (n == 0)
    return;
  const double rc = sqrt(x * x + y * y) / p[31];
  const double c0 = 1. / p[29] - 1.;
  const double d0_inv = (rs * zs) / p[30];
  const double factor = c0 * rs * zs;
  const double inner_b = ring ? p[11] : (bar ? p[13] : 0.);
  double cdrop = 1.;
  if (rc > 1)
    cdrop = pow(rc, -p[32]);
  for (int i = 0; i < n; ++i) {
    const double a = dist[i];
    const double f = rc > 1 ? factor * cdrop * exp(-a * a * cdrop * cdrop * d0_inv * d0_inv)
                            : factor * exp(-a * a * d0_inv * d0_inv);
    const double amp_i = n == 1 ? inner_b : p[24 + i];
    for (int c = 0; c < 3; ++c)
      b[c] += o[c] * f * amp_i;
  }
}

} // namespace hamx
