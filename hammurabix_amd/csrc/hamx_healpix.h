// HEALPix RING geometry on the device: pixel centres and the 4-pixel bilinear
// interpolation the integrator uses for the Faraday-depth cache and for
// accumulating shell maps of different nside.
//
// Follows the reference's Hampix (hamdis.h:242-392 pix2ang / ring helpers,
// hamdis.h:609-696 interpolation), itself the published HEALPix C++
// algorithms (healpix_base.cc pix2loc, ring_above, get_ring_info2,
// get_interpol).  Integer ring arithmetic is done in 64-bit like the
// reference's ham_uint, and every floating expression keeps the reference's
// operation order.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "hamx_cgs.h"

namespace hamx {

struct Healpix {
  uint64_t nside, npix, ncap;
  int order;
  double fact1, fact2;
};

__host__ __device__ inline int ilog2_u64(uint64_t n) {
  int r = 0;
  while (n > 1) {
    n >>= 1;
    ++r;
  }
  return r;
}

__host__ __device__ inline Healpix healpix_make(uint64_t nside) {
  Healpix h;
  h.nside = nside;
  h.order = ilog2_u64(nside);
  const uint64_t npface = nside * nside;
  h.npix = 12 * npface;
  h.ncap = (npface - nside) << 1;
  h.fact2 = 4. / h.npix;
  h.fact1 = (nside << 1) * h.fact2;
  return h;
}

__device__ inline uint64_t hp_isqrt(uint64_t n) {
  return static_cast<uint64_t>(floor(sqrt(static_cast<double>(n) + 0.5)));
}

// RING pix2ang (hamdis.h:302-356) followed by the Hamp angle normalisation
// (hamp.h:23-28).
__device__ inline void hp_pix2ang(const Healpix &h, uint64_t idx, double &theta,
                                  double &phi) {
  double z = 0.0, sth = 0.0;
  bool have_sth = false;
  if (idx < h.ncap) {
    const uint64_t iring = (1 + hp_isqrt(1 + 2 * idx)) >> 1;
    const uint64_t iphi = (idx + 1) - 2 * iring * (iring - 1);
    const double tmp = static_cast<double>(iring * iring) * h.fact2;
    z = 1.0 - tmp;
    if (z > 0.99) {
      sth = sqrt(tmp * (2.0 - tmp));
      have_sth = true;
    }
    phi = (static_cast<double>(iphi) - 0.5) * cgs::halfpi / static_cast<double>(iring);
  } else if (idx < (h.npix - h.ncap)) {
    const uint64_t nl4 = 4 * h.nside;
    const uint64_t ip = idx - h.ncap;
    const uint64_t tmp = ip >> (h.order + 2);
    const uint64_t iring = tmp + h.nside;
    const uint64_t iphi = ip - nl4 * tmp + 1;
    const double fodd = ((iring + h.nside) & 1) ? 1. : 0.5;
    z = (static_cast<double>(2 * h.nside) - static_cast<double>(iring)) * h.fact1;
    phi = (static_cast<double>(iphi) - fodd) * cgs::halfpi * 1.5 * h.fact1;
  } else {
    const uint64_t ip = h.npix - idx;
    const uint64_t iring = (1 + hp_isqrt(2 * ip - 1)) >> 1;
    const uint64_t iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
    const double tmp = static_cast<double>(iring * iring) * h.fact2;
    z = tmp - 1.0;
    if (z < -0.99) {
      sth = sqrt(tmp * (2.0 - tmp));
      have_sth = true;
    }
    phi = (static_cast<double>(iphi) - 0.5) * cgs::halfpi / static_cast<double>(iring);
  }
  theta = have_sth ? atan2(sth, z) : acos(z);
  theta = fmod(theta, cgs::pi);
  theta += (theta < 0) * cgs::pi;
  phi = fmod(phi, cgs::twopi);
  phi += (phi < 0) * cgs::twopi;
}

// ring_above (hamdis.h:358-365)
__device__ inline uint64_t hp_ring_above(const Healpix &h, double z) {
  const double az = fabs(z);
  if (az <= cgs::twothirds)
    return static_cast<uint64_t>(h.nside * (2 - 1.5 * z));
  const uint64_t iring = static_cast<uint64_t>(h.nside * sqrt(3 * (1 - az)));
  return (z > 0) ? iring : 4 * h.nside - iring - 1;
}

// get_ring_info2 (hamdis.h:367-392)
__device__ inline void hp_ring_info(const Healpix &h, uint64_t ring,
                                    uint64_t &startpix, int64_t &ringpix,
                                    double &theta, bool &shifted) {
  const uint64_t northring = (ring > 2 * h.nside) ? 4 * h.nside - ring : ring;
  if (northring < h.nside) {
    const double tmp = static_cast<double>(northring * northring) * h.fact2;
    const double costheta = 1 - tmp;
    const double sintheta = sqrt(tmp * (2 - tmp));
    theta = atan2(sintheta, costheta);
    ringpix = static_cast<int64_t>(4 * northring);
    shifted = true;
    startpix = 2 * northring * (northring - 1);
  } else {
    theta = acos((static_cast<double>(2 * h.nside) - static_cast<double>(northring)) * h.fact1);
    ringpix = static_cast<int64_t>(4 * h.nside);
    shifted = ((northring - h.nside) & 1) == 0;
    startpix = h.ncap + (northring - h.nside) * static_cast<uint64_t>(ringpix);
  }
  if (northring != ring) {
    theta = cgs::pi - theta;
    startpix = h.npix - startpix - static_cast<uint64_t>(ringpix);
  }
}

// Pixel indices and weights of the bilinear interpolation at (theta, phi)
// (hamdis.h:609-682).
__device__ inline void hp_interpol_weights(const Healpix &h, double theta,
                                           double phi, uint64_t pix[4],
                                           double wght[4]) {
  const double z = cos(theta);
  const uint64_t ring1 = hp_ring_above(h, z);
  const uint64_t ring2 = ring1 + 1;
  double theta1 = 0.0, theta2 = 0.0;
  pix[0] = pix[1] = pix[2] = pix[3] = 0;
  wght[0] = wght[1] = wght[2] = wght[3] = 0.0;
  if (ring1 > 0) {
    uint64_t start;
    int64_t ringpix;
    bool shift;
    hp_ring_info(h, ring1, start, ringpix, theta1, shift);
    const double dphi = 6.283185307179586476925286766559005768394 / static_cast<double>(ringpix);
    const double tmp = (phi / dphi - 0.5 * shift);
    int64_t i1 = static_cast<int64_t>(tmp) - static_cast<int64_t>(tmp < 0);
    const double weight = (phi - (static_cast<double>(i1) + 0.5 * shift) * dphi) / dphi;
    int64_t i2 = i1 + 1;
    i1 += ringpix * (i1 < 0);
    i2 -= ringpix * (i2 >= ringpix);
    pix[0] = start + i1;
    pix[1] = start + i2;
    wght[0] = 1.0 - weight;
    wght[1] = weight;
  }
  if (ring2 < (4 * h.nside)) {
    uint64_t start;
    int64_t ringpix;
    bool shift;
    hp_ring_info(h, ring2, start, ringpix, theta2, shift);
    const double dphi = 6.283185307179586476925286766559005768394 / static_cast<double>(ringpix);
    const double tmp = (phi / dphi - 0.5 * shift);
    int64_t i1 = static_cast<int64_t>(tmp) - static_cast<int64_t>(tmp < 0);
    const double weight = (phi - (static_cast<double>(i1) + 0.5 * shift) * dphi) / dphi;
    int64_t i2 = i1 + 1;
    i1 += ringpix * (i1 < 0);
    i2 -= ringpix * (i2 >= ringpix);
    pix[2] = start + i1;
    pix[3] = start + i2;
    wght[2] = 1.0 - weight;
    wght[3] = weight;
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
  } else if (ring2 == (4 * h.nside)) {
    const double wtheta = (theta - theta1) / (cgs::pi - theta1);
    wght[0] *= (1.0 - wtheta);
    wght[1] *= (1.0 - wtheta);
    const double fac = wtheta * 0.25;
    wght[0] += fac;
    wght[1] += fac;
    wght[2] = fac;
    wght[3] = fac;
    pix[2] = ((pix[0] + 2) & 3) + h.npix - 4;
    pix[3] = ((pix[1] + 2) & 3) + h.npix - 4;
  } else {
    const double wtheta = (theta - theta1) / (theta2 - theta1);
    wght[0] *= (1.0 - wtheta);
    wght[1] *= (1.0 - wtheta);
    wght[2] *= wtheta;
    wght[3] *= wtheta;
  }
}

// Interpolated map value with undefined pixels excluded (hamdis.h:683-695).
__device__ inline double hp_interpolate(const Healpix &h, const double *map,
                                        double theta, double phi) {
  uint64_t pix[4];
  double wght[4];
  hp_interpol_weights(h, theta, phi, pix, wght);
  double wtot = 0.0, res = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double val = map[pix[i]];
    if (val > cgs::undef_threshold) {
      res += val * wght[i];
      wtot += wght[i];
    }
  }
  return (wtot == 0.0) ? cgs::undef : res / wtot;
}

} // namespace hamx
