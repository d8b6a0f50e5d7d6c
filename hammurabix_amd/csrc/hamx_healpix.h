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

// ---- (x, y, face) <-> RING index (Hampix::rpixxyf / xyfrpix,
// hamdis.h:515-605; the HEALPix ring2xyf / xyf2ring of Gorski et al. 2005)

// face layout: ring number of each base pixel's south corner / nside and
// its longitude index
__host__ __device__ inline int hp_jrll(int f) { return 2 + (f >> 2); }
__host__ __device__ inline int hp_jpll(int f) {
  return (f < 4) ? 2 * f + 1 : (f < 8 ? 2 * (f - 4) : 2 * (f - 8) + 1);
}

__host__ __device__ inline uint64_t hp_isqrt_hd(uint64_t v) {
  uint64_t r = static_cast<uint64_t>(sqrt(static_cast<double>(v) + 0.5));
  while (r * r > v)
    --r;
  while ((r + 1) * (r + 1) <= v)
    ++r;
  return r;
}

__host__ __device__ inline void hp_ring2xyf(const Healpix &h, uint64_t pix, int64_t &ix,
                                           int64_t &iy, int &face) {
  const int64_t ns = static_cast<int64_t>(h.nside), nl2 = 2 * ns;
  int64_t iring, iphi, kshift, nr;
  if (pix < h.ncap) { // north polar cap
    iring = (1 + static_cast<int64_t>(hp_isqrt_hd(1 + 2 * pix))) >> 1;
    iphi = static_cast<int64_t>(pix) + 1 - 2 * iring * (iring - 1);
    kshift = 0;
    nr = iring;
    face = static_cast<int>((iphi - 1) / nr);
  } else if (pix < h.npix - h.ncap) { // equatorial belt
    const int64_t ip = static_cast<int64_t>(pix - h.ncap);
    const int64_t tmp = ip >> (h.order + 2);
    iring = tmp + ns;
    iphi = ip - tmp * 4 * ns + 1;
    kshift = (iring + ns) & 1;
    nr = ns;
    const int64_t ire = tmp + 1, irm = nl2 + 1 - tmp;
    const int64_t ifm = (iphi - (ire >> 1) + ns - 1) >> h.order;
    const int64_t ifp = (iphi - (irm >> 1) + ns - 1) >> h.order;
    face = static_cast<int>((ifp == ifm) ? (ifp | 4) : ((ifp < ifm) ? ifp : (ifm + 8)));
  } else { // south polar cap
    const int64_t ip = static_cast<int64_t>(h.npix - pix);
    iring = (1 + static_cast<int64_t>(hp_isqrt_hd(2 * ip - 1))) >> 1;
    iphi = 4 * iring + 1 - (ip - 2 * iring * (iring - 1));
    kshift = 0;
    nr = iring;
    iring = 2 * nl2 - iring;
    face = 8 + static_cast<int>((iphi - 1) / nr);
  }
  const int64_t irt = iring - hp_jrll(face) * ns + 1;
  int64_t ipt = 2 * iphi - hp_jpll(face) * nr - kshift - 1;
  if (ipt >= nl2)
    ipt -= 8 * ns;
  ix = (ipt - irt) >> 1;
  iy = (-ipt - irt) >> 1;
}

__host__ __device__ inline uint64_t hp_xyf2ring(const Healpix &h, int64_t ix, int64_t iy,
                                               int face) {
  const int64_t ns = static_cast<int64_t>(h.nside), nl4 = 4 * ns;
  const int64_t jr = hp_jrll(face) * ns - ix - iy - 1;
  int64_t nr, kshift, n_before;
  if (jr < ns) {
    nr = jr;
    n_before = 2 * nr * (nr - 1);
    kshift = 0;
  } else if (jr > 3 * ns) {
    nr = nl4 - jr;
    n_before = static_cast<int64_t>(h.npix) - 2 * (nr + 1) * nr;
    kshift = 0;
  } else {
    nr = ns;
    n_before = static_cast<int64_t>(h.ncap) + (jr - ns) * nl4;
    kshift = (jr - ns) & 1;
  }
  int64_t jp = (hp_jpll(face) * nr + ix - iy + 1 + kshift) / 2;
  if (jp > nl4)
    jp -= nl4;
  else if (jp < 1)
    jp += nl4;
  return static_cast<uint64_t>(n_before + jp - 1);
}

// Even bits of v packed into the low half (the inverse of bit interleaving).
__host__ __device__ inline uint64_t hp_compress_bits(uint64_t v) {
  v &= 0x5555555555555555ull;
  v = (v | (v >> 1)) & 0x3333333333333333ull;
  v = (v | (v >> 2)) & 0x0f0f0f0f0f0f0f0full;
  v = (v | (v >> 4)) & 0x00ff00ff00ff00ffull;
  v = (v | (v >> 8)) & 0x0000ffff0000ffffull;
  v = (v | (v >> 16)) & 0x00000000ffffffffull;
  return v;
}

// RING index of NESTED pixel p (nside a power of two): p = face nside^2 +
// the interleaved bits of (ix, iy) on the face, ix on the even bits.
__host__ __device__ inline uint64_t hp_nest2ring(const Healpix &h, uint64_t p) {
  const uint64_t npface = h.nside * h.nside;
  const uint64_t face = p / npface, q = p - face * npface;
  return hp_xyf2ring(h, static_cast<int64_t>(hp_compress_bits(q)),
                     static_cast<int64_t>(hp_compress_bits(q >> 1)), static_cast<int>(face));
}

// Value of the mask at nside ht.nside, pixel m, from the pivot mask at
// hp.nside (Hampisk::duplicate, hamsk.h:55-103): a copy at the same nside;
// 1 if any of the pixel's sub-pixels is > 0 when downgrading (else 0); the
// parent pixel's value when upgrading.
__host__ __device__ inline double hp_mask_value(const Healpix &ht, uint64_t m,
                                                const double *piv, const Healpix &hp) {
  if (hp.nside == ht.nside)
    return piv[m];
  int64_t x, y;
  int f;
  hp_ring2xyf(ht, m, x, y, f);
  if (hp.nside > ht.nside) {
    const int64_t fact = static_cast<int64_t>(hp.nside / ht.nside);
    for (int64_t j = fact * y; j < fact * (y + 1); ++j)
      for (int64_t i = fact * x; i < fact * (x + 1); ++i)
        if (piv[hp_xyf2ring(hp, i, j, f)] > 0)
          return 1.0;
    return 0.0;
  }
  const int64_t fact = static_cast<int64_t>(ht.nside / hp.nside);
  return piv[hp_xyf2ring(hp, x / fact, y / fact, f)];
}

} // namespace hamx
