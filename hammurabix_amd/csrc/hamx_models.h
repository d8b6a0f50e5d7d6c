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

// ------------------------------------------------------------------ YMW16 --
// TEreg_ymw16::write_field (tereg_ymw16.cc:10-333): the YMW16 thermal
// electron density in its own frame (x, y, z) = (y, -x, z) of ours, with the
// Galactic warp applied to z beyond r_warp and nothing beyond 25 kpc.
// Components: thick disk, thin disk, spiral arms (the reference's arm loop
// stops before the Local arm, index 4), Galactic centre, Gum Nebula, Local
// Bubble and Loop I, combined by the reference's weighting rules.  Parameter
// order (HAMX_TEREG_YMW16 in hamx.h):
//   0 r_warp, 1 r0, 2 gamma_w | 3 ad, 4 bd, 5 n1, 6 h1 | 7 a2, 8 b2, 9 n2, 10 k2 |
//   11 b2s, 12-16 narm, 17-21 warm, 22-26 rmin, 27-31 phimin [rad],
//   32-36 tan(pitch), 37-41 cos(pitch), 42 aa, 43 ka, 44 ncn, 45 thetacn [deg],
//   46 wcn [deg], 47 nsg, 48 thetasg [deg], 49 wsg [deg] |
//   50 ngc, 51 agc, 52 hgc | 53 ngn, 54 wgn, 55 agn, 56 kgn |
//   57 j_lb, 58 nlb1, 59 thetalb1 [deg], 60 detlb1 [deg], 61 wlb1, 62 hlb1,
//   63 nlb2, 64 thetalb2 [deg], 65 detlb2 [deg], 66 wlb2, 67 hlb2 |
//   68 nli, 69 rli, 70 wli, 71 detthetali [deg], 72 thetali [deg];
// lengths in cm, densities in cm^-3.
namespace ymw16 {

HAMX_HD inline double sech2(double v) {
  const double s = 1. / cosh(v);
  return s * s;
}

// thick disk (tereg_ymw16.cc:82-94)
HAMX_HD inline double thick(const double *p, double z, double r) {
  if (z > 10. * p[6])
    return 0.;
  const double gd = r > p[4] ? sech2((r - p[4]) / p[3]) : 1.;
  return p[5] * gd * sech2(z / p[6]);
}

// thin disk (tereg_ymw16.cc:97-115)
HAMX_HD inline double thin(const double *p, double z, double r) {
  const double pc = cgs::pc;
  const double h0 = p[10] * (32 * pc + 1.6e-3 * r + (4.e-7 / pc) * r * r);
  if (z > 10. * h0)
    return 0.;
  const double gd = r > p[4] ? sech2((r - p[4]) / p[3]) : 1.;
  return p[9] * gd * sech2((r - p[8]) / p[7]) * sech2(z / h0);
}

// spiral arms (tereg_ymw16.cc:118-210); theta * rad compared with degrees is
// the reference's expression, kept as written
HAMX_HD inline double spiral(const double *p, double x, double y, double z, double r) {
  const double pc = cgs::pc, rad = cgs::rad, twopi = 2. * cgs::pi;
  double scaling = 1.;
  if (r > p[4]) {
    if ((r - p[4]) > 10. * p[3])
      return 0.;
    scaling = sech2((r - p[4]) / p[3]);
  }
  const double h0 = p[43] * (32 * pc + 1.6e-3 * r + (4.e-7 / pc) * (r * r));
  if (z > 10. * h0)
    return 0.;
  scaling *= sech2(z / h0);
  if ((r - p[11]) > 10. * p[42])
    return 0.;
  scaling *= sech2((r - p[11]) / p[42]);
  double theta = atan2(y, x);
  if (theta < 0)
    theta += twopi;
  double ne = 0.;
  for (int i = 0; i < 4; ++i) {
    double dphi = theta - p[27 + i];
    if (dphi < 0)
      dphi += twopi;
    const double d = fabs(p[22 + i] * exp(dphi * p[32 + i]) - r);
    const double dw = fabs(p[22 + i] * exp((dphi + twopi) * p[32 + i]) - r);
    const double smin = (d < dw ? d : dw) * p[37 + i];
    if (smin > 10. * p[17 + i])
      continue;
    const double arm = sech2(smin / p[17 + i]);
    if (i != 2) {
      ne += p[12 + i] * scaling * arm;
      continue;
    }
    const double th = theta * rad;
    const double sg = (th - p[48]) / p[49];
    const double g_sg = 1. - p[47] * exp(-(sg * sg));
    double g_cn;
    if (r > 6 * cgs::kpc && th > p[45]) {
      g_cn = 1. + p[44];
    } else {
      const double cn = (th - p[45]) / p[46];
      g_cn = 1. + p[44] * exp(-(cn * cn));
    }
    ne += p[12 + i] * scaling * (g_sg * g_cn * arm);
  }
  return ne;
}

// Galactic centre (tereg_ymw16.cc:213-229)
HAMX_HD inline double galcen(const double *p, double x, double y, double z) {
  const double xg = 50. * cgs::pc, zg = -7. * cgs::pc;
  const double R2 = (x - xg) * (x - xg) + (y - 0.) * (y - 0.);
  if (R2 > 10. * p[51] * p[51])
    return 0.;
  const double Ar = exp(-R2 / (p[51] * p[51]));
  if (fabs(z - zg) > 10. * p[52])
    return 0.;
  return p[50] * Ar * sech2((z - zg) / p[52]);
}

// Gum Nebula (tereg_ymw16.cc:232-270): an ellipsoidal shell about
// (l, b, d) = (264 deg, -4 deg, 450 pc) from the Sun
HAMX_HD inline double gum(const double *p, double x, double y, double z) {
  if (y < 0 || x > 0)
    return 0.;
  const double lc = 264. * cgs::rad, bc = -4. * cgs::rad, dc = 450. * cgs::pc;
  const double xc = dc * cos(bc) * sin(lc);
  const double yc = p[1] - dc * cos(bc) * cos(lc);
  const double zc = dc * sin(bc);
  const double theta = atan2(fabs(z - zc), sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc)));
  const double tt = tan(theta);
  const double a = p[55], k = p[56];
  const double zp = (a * k) / sqrt(1. + k * k / (tt * tt));
  const double xyp = zp / tt;
  const double xy_dist = sqrt(a * a - xyp * xyp) * static_cast<double>(a > xyp);
  const double alpha = atan2(k * xyp, xy_dist) + theta;
  const double R2 = (x - xc) * (x - xc) + (y - yc) * (y - yc) + (z - zc) * (z - zc);
  const double r2 = zp * zp + xyp * xyp;
  const double sa = sin(alpha);
  const double D2 = (R2 + r2 - 2. * sqrt(R2 * r2)) * sa * sa;
  if (D2 > 10. * p[54] * p[54])
    return 0.;
  return p[53] * exp(-D2 / (p[54] * p[54]));
}

// the distance from the Local Bubble's axis (tereg_ymw16.cc:46-48,280-282)
HAMX_HD inline double lb_radius(double x, double y, double z) {
  const double u = (y - 8.34 * cgs::kpc) * 0.94 - 0.34 * z;
  return sqrt(u * u + x * x);
}

// Local Bubble (tereg_ymw16.cc:273-307), l in degrees
HAMX_HD inline double localbubble(const double *p, double x, double y, double z, double l,
                                  double Rlb) {
  if (y < 0)
    return 0.;
  const double rLB = lb_radius(x, y, z);
  double ne = 0.;
  for (int c = 0; c < 2; ++c) {
    const double *q = p + 58 + 5 * c; // n, theta, det, w, h
    const double a = fabs(l + 360. - q[1]), b = fabs(q[1] - l);
    const double dl = a < b ? a : b;
    if (dl < 10. * q[2] || (rLB - Rlb) < 10. * q[3] || z < 10. * q[4])
      ne += q[0] * sech2(dl / q[2]) * sech2((rLB - Rlb) / q[3]) * sech2(z / q[4]);
  }
  return ne;
}

// Loop I (tereg_ymw16.cc:310-333)
HAMX_HD inline double loop1(const double *p, double x, double y, double z) {
  if (y < 0)
    return 0.;
  const double th_li = p[72] * cgs::rad;
  const double xc = -10.156 * cgs::pc, yc = 8106.207 * cgs::pc, zc = 10.467 * cgs::pc;
  const double rLI = sqrt((x - xc) * (x - xc) + (y - yc) * (y - yc) + (z - zc) * (z - zc));
  const double theta = acos(((x - xc) * cos(th_li) + (z - zc) * sin(th_li)) / rLI) / cgs::rad;
  if (theta > 10. * p[71] || (rLI - p[69]) > 10. * p[70])
    return 0.;
  const double dr = (rLI - p[69]) / p[70], dt = theta / p[71];
  return p[68] * exp(-(dr * dr)) * exp(-(dt * dt));
}

} // namespace ymw16

HAMX_HD inline double tereg_ymw16_eval(const double *p, double px, double py, double pz) {
  double x = py, y = -px, z = pz;
  const double r = sqrt(x * x + y * y);
  if (r >= p[0]) {
    const double tw = atan2(y, x);
    z -= p[2] * (r - p[0]) * cos(tw);
  }
  if (sqrt(x * x + y * y + z * z) > 25 * cgs::kpc)
    return 0.;
  const double l = atan2(x, p[1] - y) / cgs::rad;
  const double n_thick = ymw16::thick(p, z, r);
  const double n_thin = ymw16::thin(p, z, r);
  const double n_arm = ymw16::spiral(p, x, y, z, r);
  const double n_gc = ymw16::galcen(p, x, y, z);
  const double n_gum = ymw16::gum(p, x, y, z);
  const double Rlb = 110. * cgs::pc;
  const double n_lb = ymw16::localbubble(p, x, y, z, l, Rlb);
  const double n_li = ymw16::loop1(p, x, y, z);
  const double n_disk = n_thin > n_arm ? n_thin : n_arm;
  double n0 = n_thick + n_disk;
  double w_lb = 0., w_gum = 0., w_loop = 0.;
  if (ymw16::lb_radius(x, y, z) < Rlb) {
    n0 = p[57] * n_thick + n_disk;
    if (n_lb > n0)
      w_lb = 1;
  } else if (n_lb > n0 && n_lb > n_gum) {
    w_lb = 1;
  }
  if (n_li > n0)
    w_loop = 1;
  if (n_gum > n0)
    w_gum = 1;
  return (1 - w_lb) * ((1 - w_gum) * ((1 - w_loop) * (n0 + n_gc) + w_loop * n_li) +
                       w_gum * n_gum) +
         w_lb * n_lb;
}

} // namespace hamx
