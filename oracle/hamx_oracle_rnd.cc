// hamx_oracle_rnd.cc -- CPU restatement of the Gaussian random-field
// generators Brnd_es::write_grid (brnd_es.cc:76-308) and
// TErnd_dft::write_grid (ternd_dft.cc:35-128).
//
// TEST INFRASTRUCTURE ONLY (see hamx_oracle.h).
//
// The reference draws from per-OpenMP-thread GSL taus generators, so its
// realisations depend on the thread count and cannot be reproduced without
// GSL.  Both this restatement and the device path draw from the same
// counter-based stream instead (Philox4x32-10, Salmon et al., SC'11, with
// Box-Muller), so the generator kernels can be checked element-wise; the
// statistics are checked against the target spectrum separately.
#include "hamx_oracle.h"

#include <cmath>
#include <complex>
#include <cstring>
#include <vector>

namespace {

const double kPi = 3.141592653589793238462643383279502884197;
const double kTwoPi = 6.283185307179586476925286766559005768394;
const double kOneThirdES = 0.333333333333333333333333333333333333333;
const double kTwoThirdsES = 0.666666666666666666666666666666666666667;
const double kSqrt2 = 1.4142135623730950488016887242096980785697;
const double kSqrt3ES = 1.7320508075688772935274463415058723669428;
const double kKpc = 3.0856775806e+21;

inline void mulhilo(uint32_t a, uint32_t b, uint32_t &hi, uint32_t &lo) {
  const uint64_t p = static_cast<uint64_t>(a) * b;
  hi = static_cast<uint32_t>(p >> 32);
  lo = static_cast<uint32_t>(p);
}

// Philox4x32 round function and key schedule (Salmon et al. 2011, Random123).
void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
  uint32_t c[4] = {c_in[0], c_in[1], c_in[2], c_in[3]};
  uint32_t k0 = k_in[0], k1 = k_in[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    uint32_t hi0, lo0, hi1, lo1;
    mulhilo(0xD2511F53u, c[0], hi0, lo0);
    mulhilo(0xCD9E8D57u, c[2], hi1, lo1);
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
  for (int i = 0; i < 4; ++i)
    out[i] = c[i];
}

// Pair m of stream `stream`: counter (m_lo, m_hi, stream, 0), key (seed_lo,
// seed_hi); two 53-bit uniforms u1 in (0,1], u2 in [0,1); Box-Muller.
void gauss_pair(uint64_t seed, uint32_t stream, uint64_t m, double &g0,
                double &g1) {
  const uint32_t ctr[4] = {static_cast<uint32_t>(m), static_cast<uint32_t>(m >> 32),
                           stream, 0u};
  const uint32_t key[2] = {static_cast<uint32_t>(seed),
                           static_cast<uint32_t>(seed >> 32)};
  uint32_t x[4];
  philox(ctr, key, x);
  const uint64_t a = (static_cast<uint64_t>(x[1]) << 32) | x[0];
  const uint64_t b = (static_cast<uint64_t>(x[3]) << 32) | x[2];
  const double u1 = (static_cast<double>(a >> 11) + 1.0) * 0x1.0p-53;
  const double u2 = static_cast<double>(b >> 11) * 0x1.0p-53;
  const double r = std::sqrt(-2.0 * std::log(u1));
  const double t = kTwoPi * u2;
  g0 = r * std::cos(t);
  g1 = r * std::sin(t);
}

// Iterative radix-2 complex FFT of one line, exponent sign*2*pi*i*jk/n
// (FFTW convention: sign = +1 is FFTW_BACKWARD), unnormalised.
void fft_line_pow2(std::vector<double> &buf, uint64_t n, int sign) {
  // bit reversal
  for (uint64_t i = 1, j = 0; i < n; ++i) {
    uint64_t bit = n >> 1;
    for (; j & bit; bit >>= 1)
      j ^= bit;
    j ^= bit;
    if (i < j) {
      std::swap(buf[2 * i], buf[2 * j]);
      std::swap(buf[2 * i + 1], buf[2 * j + 1]);
    }
  }
  for (uint64_t len = 2; len <= n; len <<= 1) {
    const uint64_t half = len >> 1;
    for (uint64_t k = 0; k < half; ++k) {
      const double ang = sign * kTwoPi * static_cast<double>(k) / static_cast<double>(len);
      const double wr = std::cos(ang), wi = std::sin(ang);
      for (uint64_t s = 0; s < n; s += len) {
        const uint64_t a = s + k, b = s + k + half;
        const double br = buf[2 * b] * wr - buf[2 * b + 1] * wi;
        const double bi = buf[2 * b] * wi + buf[2 * b + 1] * wr;
        buf[2 * b] = buf[2 * a] - br;
        buf[2 * b + 1] = buf[2 * a + 1] - bi;
        buf[2 * a] += br;
        buf[2 * a + 1] += bi;
      }
    }
  }
}

// Any length: decimation in time by the smallest prime factor p of n,
// X_k = sum_r exp(sign 2 pi i r k / n) FFT_{n/p}(x_{p m + r})_{k mod n/p},
// and the plain DFT for a prime length.  Angles are reduced modulo n in
// integers before the division, so they are exact up to one rounding.
void dft_any(const std::complex<double> *x, uint64_t stride, uint64_t n, int sign,
             std::complex<double> *out) {
  auto root = [&](uint64_t m) {
    const double a = sign * kTwoPi * static_cast<double>(m % n) / static_cast<double>(n);
    return std::complex<double>(std::cos(a), std::sin(a));
  };
  if (n == 1) {
    out[0] = x[0];
    return;
  }
  uint64_t p = 2;
  while (p * p <= n && n % p != 0)
    ++p;
  if (n % p != 0)
    p = n; // prime
  if (p == n) {
    for (uint64_t k = 0; k < n; ++k) {
      std::complex<double> acc(0., 0.);
      for (uint64_t j = 0; j < n; ++j)
        acc += x[j * stride] * root(j * k);
      out[k] = acc;
    }
    return;
  }
  const uint64_t m = n / p;
  std::vector<std::complex<double>> sub(n);
  for (uint64_t r = 0; r < p; ++r)
    dft_any(x + r * stride, stride * p, m, sign, sub.data() + r * m);
  for (uint64_t k = 0; k < n; ++k) {
    std::complex<double> acc = sub[k % m];
    for (uint64_t r = 1; r < p; ++r)
      acc += sub[r * m + k % m] * root(r * k);
    out[k] = acc;
  }
}

void fft_line(std::vector<double> &buf, uint64_t n, int sign) {
  if ((n & (n - 1)) == 0) {
    fft_line_pow2(buf, n, sign);
    return;
  }
  std::vector<std::complex<double>> in(n), out(n);
  for (uint64_t t = 0; t < n; ++t)
    in[t] = std::complex<double>(buf[2 * t], buf[2 * t + 1]);
  dft_any(in.data(), 1, n, sign, out.data());
  for (uint64_t t = 0; t < n; ++t) {
    buf[2 * t] = out[t].real();
    buf[2 * t + 1] = out[t].imag();
  }
}

void fft3d(double *data, uint64_t nx, uint64_t ny, uint64_t nz, int sign) {
  const uint64_t dims[3] = {nx, ny, nz};
  const uint64_t strides[3] = {ny * nz, nz, 1};
  for (int ax = 0; ax < 3; ++ax) {
    const uint64_t n = dims[ax], st = strides[ax];
    if (n < 2)
      continue;
    const uint64_t total = nx * ny * nz;
    const uint64_t lines = total / n;
#pragma omp parallel
    {
      std::vector<double> buf(2 * n);
#pragma omp for
      for (int64_t li = 0; li < (int64_t)lines; ++li) {
        // base offset of line li: enumerate the other two indices
        uint64_t base;
        if (ax == 0)
          base = static_cast<uint64_t>(li); // (j, l) flattened = li
        else if (ax == 1)
          base = (li / nz) * (ny * nz) + (li % nz);
        else
          base = static_cast<uint64_t>(li) * nz;
        for (uint64_t t = 0; t < n; ++t) {
          buf[2 * t] = data[2 * (base + t * st)];
          buf[2 * t + 1] = data[2 * (base + t * st) + 1];
        }
        fft_line(buf, n, sign);
        for (uint64_t t = 0; t < n; ++t) {
          data[2 * (base + t * st)] = buf[2 * t];
          data[2 * (base + t * st) + 1] = buf[2 * t + 1];
        }
      }
    }
  }
}

// toolkit::mean / variance (toolkit.h:48-78), sequential sums
double variance(const double *arr, uint64_t size) {
  double avg = 0;
  for (uint64_t i = 0; i != size; ++i)
    avg += arr[i];
  avg /= size;
  double var = 0.;
  for (uint64_t i = 0; i != size; ++i)
    var += (arr[i] - avg) * (arr[i] - avg);
  var /= size;
  return var;
}

struct V3 {
  double e[3];
};
double dotv(const V3 &a, const V3 &b) {
  double t = 0.;
  for (int i = 0; i < 3; ++i)
    t += a.e[i] * b.e[i];
  return t;
}
double lensq(const V3 &a) {
  double t = 0.;
  for (int i = 0; i < 3; ++i)
    t += a.e[i] * a.e[i];
  return t;
}

// Brnd_es::spectrum (brnd_es.cc:34-49)
double es_spectrum(const hamx_es_desc *d, double k) {
  const double p0 = d->rms * d->rms;
  const double unit = 1. / (4. * kPi * k * k);
  const double band1 = double(k < d->k1);
  const double band2 = double(k > d->k1) * double(k < d->k0);
  const double band3 = double(k > d->k0);
  const double P = band1 * std::pow(d->k0 / d->k1, d->a1) * std::pow(k / d->k1, 6.) +
                   band2 / std::pow(k / d->k0, d->a1) +
                   band3 / std::pow(k / d->k0, d->a0);
  return P * p0 * unit;
}

// TErnd_dft::spectrum (ternd_dft.cc:17-24)
double dft_spectrum(const hamx_dft_desc *d, double k) {
  if (k < d->k0)
    return 0.;
  return d->rms * std::pow(k / d->k0, d->a0) / (4 * kPi * k * k);
}

// energy-profile rescaling (brnd_es.cc:53-59, ternd_dft.cc:26-33)
double profile(const double *obs, double r0, double z0, const double pos[3]) {
  const double r_cyl = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]) - std::fabs(obs[0]);
  const double z = std::fabs(pos[2]) - std::fabs(obs[2]);
  return std::exp(-r_cyl / r0) * std::exp(-z / z0);
}

// wave number of index i on an axis of n points over length len
// (brnd_es.cc:106-122)
double wavenum(uint64_t i, uint64_t n, double len) {
  double k = kKpc * i / len;
  if (i >= (n + 1) / 2)
    k -= kKpc * n / len;
  return k;
}

// regular field for the anisotropy direction (brnd_es.cc:19-22)
V3 breg_at(const hamx_es_desc *d, const double pos[3]) {
  double b[3] = {0., 0., 0.};
  if (d->breg_model == HAMX_BREG_GRID) {
    orc_read_grid3(d->breg_x, d->breg_y, d->breg_z, &d->breg_box, pos, b);
  } else {
    orc_breg(d->breg_model, d->breg, pos, b);
  }
  V3 v;
  for (int i = 0; i < 3; ++i)
    v.e[i] = b[i];
  return v;
}

// ---------------------------------------------------------------- MHD --
// A 53-bit uniform in [0, 1) from the counter-based stream.
double uniform53(uint64_t seed, uint32_t stream, uint64_t m) {
  const uint32_t ctr[4] = {static_cast<uint32_t>(m), static_cast<uint32_t>(m >> 32), stream, 0u};
  const uint32_t key[2] = {static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32)};
  uint32_t x[4];
  philox(ctr, key, x);
  const uint64_t a = (static_cast<uint64_t>(x[1]) << 32) | x[0];
  return static_cast<double>(a >> 11) * 0x1.0p-53;
}

// Brnd_mhd (brnd_mhd.cc:16-262, bfield.h:370-443), with the parameters of
// Param::brnd_mhd.
struct Mhd {
  const hamx_mhd_desc *d;
  double spectrum(double k, double a0, double p0) const {
    const double unit = 0.25 / (kPi * k * k);
    const double band1 = double(k < d->k1);
    const double band2 = double(k > d->k1) * double(k < d->k0);
    const double band3 = double(k > d->k0);
    const double P = band1 * std::pow(d->k0 / d->k1, d->a1) * std::pow(k / d->k1, 6.0) +
                     band2 / std::pow(k / d->k0, d->a1) + band3 / std::pow(k / d->k0, a0);
    return P * p0 * unit;
  }
  static double dynamo(double beta, double cosa) {
    return 1 + 0.25 * beta * beta + beta * (1. - 2. * cosa * cosa);
  }
  static double h_s(double beta, double cosa) {
    const double sqrtD = std::sqrt(dynamo(beta, cosa));
    const double dpp = 1 + sqrtD + 0.5 * beta;
    const double dmm = 1 - sqrtD - 0.5 * beta;
    const double dmp = 1 - sqrtD + 0.5 * beta;
    const double demon = 1 / (dmp * (cosa * cosa + dpp * dpp * (1 - cosa * cosa) / (dmm * dmm)));
    if (std::isfinite(demon))
      return 2. * cosa * cosa * demon * (demon > 0);
    return 0;
  }
  static double h_f(double beta, double cosa) {
    if (cosa == 0)
      return 0;
    const double sqrtD = std::sqrt(dynamo(beta, cosa));
    const double dpp = 1 + sqrtD + 0.5 * beta;
    const double dpm = 1 + sqrtD - 0.5 * beta;
    const double dmp = 1 - sqrtD + 0.5 * beta;
    const double demon = 1 / (dpp * (cosa * cosa + dmp * dmp * (1 - cosa * cosa) / (dpm * dpm)));
    if (std::isfinite(demon))
      return 2. * cosa * cosa * demon;
    return 0;
  }
  static double F(double ma, double cosa) {
    return std::exp(-1. * std::pow(ma, -1.33333333) * std::fabs(cosa) /
                    std::pow(1 - cosa * cosa, 0.33333333));
  }
};

V3 versor(const V3 &v) {
  V3 t = {{0., 0., 0.}};
  const double l2 = lensq(v);
  if (l2 != 0) {
    const double den = 1. / std::sqrt(l2);
    for (int c = 0; c < 3; ++c)
      t.e[c] = v.e[c] * den;
  }
  return t;
}
V3 crossv(const V3 &a, const V3 &b) {
  V3 t;
  for (int i = 0; i < 3; ++i)
    t.e[i] = a.e[(i + 1) % 3] * b.e[(i + 2) % 3] - a.e[(i + 2) % 3] * b.e[(i + 1) % 3];
  return t;
}
V3 normalised_unless_small(const V3 &t) {
  if (lensq(t) < 1e-12)
    return t;
  const double l = std::sqrt(lensq(t));
  V3 r;
  for (int c = 0; c < 3; ++c)
    r.e[c] = t.e[c] / l;
  return r;
}

} // namespace

extern "C" {

void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  philox(ctr, key, out);
}

void orc_gauss_pairs(uint64_t seed, uint32_t stream, uint64_t first_pair,
                     uint64_t count, double *out) {
  for (uint64_t m = 0; m < count; ++m)
    gauss_pair(seed, stream, first_pair + m, out[2 * m], out[2 * m + 1]);
}

void orc_fft3d(double *data, uint64_t nx, uint64_t ny, uint64_t nz, int sign) {
  fft3d(data, nx, ny, nz, sign);
}

double orc_variance(const double *arr, uint64_t n) { return variance(arr, n); }

double orc_es_spectrum(const hamx_es_desc *d, double k) { return es_spectrum(d, k); }
double orc_dft_spectrum(const hamx_dft_desc *d, double k) {
  return dft_spectrum(d, k);
}

// Brnd_es::write_grid (brnd_es.cc:76-308)
// Brnd_jf12::spatial_profile (brnd_jf12.cc:54-114)
static double jf12_profile(const double *q, const double pos[3]) {
  const double b_arms[8] = {q[0], q[1], q[2], q[3], q[4], q[5], q[6], q[7]};
  const double b0_int = q[8], z0_spiral = q[9], b0_halo = q[10], r0_halo = q[11],
               z0_halo = q[12];
  const double Rmax = 20. * kKpc, rho_GC = 1. * kKpc;
  const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
  const double rho = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
  const double z = pos[2];
  const double phi = std::atan2(pos[1], pos[0]);
  const double rc_B[8] = {5.1, 6.3, 7.1, 8.3, 9.8, 11.4, 12.7, 15.5};
  const double inc = 11.5;
  const double muG = 1e-6;
  double scaling_disk = 0.0 * muG;
  double scaling_halo = 0.0 * muG;
  if (r > Rmax || rho < rho_GC)
    return 0.0;
  if (r < 5. * kKpc) {
    scaling_disk = b0_int;
  } else {
    double r_negx = r * std::exp(-1 / std::tan(M_PI / 180. * (90 - inc)) * (phi - M_PI));
    if (r_negx > rc_B[7] * kKpc)
      r_negx = r * std::exp(-1 / std::tan(M_PI / 180. * (90 - inc)) * (phi + M_PI));
    if (r_negx > rc_B[7] * kKpc)
      r_negx = r * std::exp(-1 / std::tan(M_PI / 180. * (90 - inc)) * (phi + 3 * M_PI));
    for (int i = 7; i >= 0; i--)
      if (r_negx < rc_B[i] * kKpc)
        scaling_disk = b_arms[i] * (5. * kKpc) / r;
  }
  scaling_disk = scaling_disk * std::exp(-0.5 * z * z / (z0_spiral * z0_spiral));
  scaling_halo = b0_halo * std::exp(-std::fabs(r / r0_halo)) * std::exp(-std::fabs(z / z0_halo));
  return (scaling_disk * scaling_disk + scaling_halo * scaling_halo) / (muG * muG);
}

// Brnd_es::write_grid (brnd_es.cc:76-308); with jf12, Brnd_jf12::write_grid
// (brnd_jf12.cc:127-385): its energy profile and its literal constants.
static int brnd_generate(const hamx_es_desc *d, uint64_t seed, double *bx, double *by, double *bz,
                  bool jf12) {
  const double kOneThird = jf12 ? 0.33333333 : kOneThirdES;
  const double kTwoThirds = jf12 ? 0.66666667 : kTwoThirdsES;
  const double kSqrt3 = jf12 ? 1.73205081 : kSqrt3ES;
  const uint64_t nx = d->box.nx, ny = d->box.ny, nz = d->box.nz;
  const uint64_t N = nx * ny * nz;
  std::vector<double> c0(2 * N, 0.), c1(2 * N, 0.);
  const double lx = d->box.x_max - d->box.x_min;
  const double ly = d->box.y_max - d->box.y_min;
  const double lz = d->box.z_max - d->box.z_min;
  // STEP I: spectrum draw (brnd_es.cc:92-140)
  const double dk3 = kKpc * kKpc * kKpc / (lx * ly * lz);
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)nx; ++i) {
    const double kx = wavenum(i, nx, lx);
    for (uint64_t j = 0; j < ny; ++j) {
      const double ky = wavenum(j, ny, ly);
      for (uint64_t l = 0; l < nz; ++l) {
        const uint64_t idx = i * ny * nz + j * nz + l;
        if (idx == 0)
          continue;
        const double kz = wavenum(l, nz, lz);
        const double ks = std::sqrt(kx * kx + ky * ky + kz * kz);
        const double sigma = std::sqrt(kOneThird * es_spectrum(d, ks) * dk3);
        double g0, g1, g2, g3;
        gauss_pair(seed, 0, 2 * idx, g0, g1);
        gauss_pair(seed, 0, 2 * idx + 1, g2, g3);
        c0[2 * idx] = sigma * g0;
        c0[2 * idx + 1] = sigma * g1;
        c1[2 * idx] = sigma * g2;
        c1[2 * idx + 1] = sigma * g3;
      }
    }
  }
  fft3d(c0.data(), nx, ny, nz, +1);
  fft3d(c1.data(), nx, ny, nz, +1);
  // STEP II: energy-profile rescale and anisotropy (brnd_es.cc:157-205)
  const double b_var_invsq = 1. / std::sqrt(3. * variance(c0.data(), N));
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)nx; ++i) {
    double pos[3] = {i * lx / (nx - 1) + d->box.x_min, 0, 0};
    for (uint64_t j = 0; j < ny; ++j) {
      pos[1] = j * ly / (ny - 1) + d->box.y_min;
      for (uint64_t l = 0; l < nz; ++l) {
        pos[2] = l * lz / (nz - 1) + d->box.z_min;
        const double prof =
            jf12 ? jf12_profile(d->jf12, pos) : profile(d->observer, d->r0, d->z0, pos);
        const double ratio = std::sqrt(prof) * d->rms * b_var_invsq;
        const uint64_t idx = i * ny * nz + j * nz + l;
        V3 b_re = {{c0[2 * idx] * ratio, c0[2 * idx + 1] * ratio,
                    c1[2 * idx + 1] * ratio}};
        const V3 braw = breg_at(d, pos);
        // Hamvec::versor (hamvec.h:176-186)
        V3 H = {{0., 0., 0.}};
        const double l2 = lensq(braw);
        if (l2 != 0) {
          const double den = 1. / std::sqrt(l2);
          for (int c = 0; c < 3; ++c)
            H.e[c] = braw.e[c] * den;
        }
        const double rho = d->rho;
        const double rho2 = rho * rho;
        const double rhonorm = 1. / std::sqrt(kOneThird * rho2 + kTwoThirds / rho2);
        if (lensq(H) < 1e-10)
          continue;
        const double hb = dotv(H, b_re);
        V3 par, perp;
        for (int c = 0; c < 3; ++c) {
          par.e[c] = H.e[c] * hb;
          perp.e[c] = b_re.e[c] - par.e[c];
        }
        for (int c = 0; c < 3; ++c)
          b_re.e[c] = (par.e[c] * rho + perp.e[c] / rho) * rhonorm;
        c0[2 * idx] = b_re.e[0];
        c0[2 * idx + 1] = b_re.e[1];
        c1[2 * idx] = b_re.e[1];
        c1[2 * idx + 1] = b_re.e[2];
      }
    }
  }
  fft3d(c0.data(), nx, ny, nz, -1);
  fft3d(c1.data(), nx, ny, nz, -1);
  // STEP III: Gram-Schmidt in Fourier space (brnd_es.cc:211-259)
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)nx; ++i) {
    const uint64_t is = i == 0 ? 0 : nx - i;
    V3 k;
    k.e[0] = wavenum(i, nx, lx);
    for (uint64_t j = 0; j < ny; ++j) {
      const uint64_t js = j == 0 ? 0 : ny - j;
      k.e[1] = wavenum(j, ny, ly);
      for (uint64_t l = 0; l < nz; ++l) {
        const uint64_t ls = l == 0 ? 0 : nz - l;
        k.e[2] = wavenum(l, nz, lz);
        const uint64_t idx = i * ny * nz + j * nz + l;
        const uint64_t sym = is * ny * nz + js * nz + ls;
        V3 b = {{0.5 * (c0[2 * idx] + c0[2 * sym]), 0.5 * (c1[2 * idx] + c1[2 * sym]),
                 0.5 * (c1[2 * idx + 1] + c1[2 * sym + 1])}};
        // gramschmidt (brnd_es.cc:63-74)
        V3 fr = b;
        if (!(lensq(k) == 0 or lensq(b) == 0)) {
          const double inv = 1. / lensq(k);
          for (int c = 0; c < 3; ++c)
            fr.e[c] = kSqrt3 * (b.e[c] - k.e[c] * dotv(k, b) * inv);
        }
        bx[idx] = kSqrt2 * fr.e[0];
        by[idx] = kSqrt2 * fr.e[1];
        bz[idx] = kSqrt2 * fr.e[2];
      }
    }
  }
  // reassemble (brnd_es.cc:261-270)
#pragma omp parallel for
  for (int64_t idx = 0; idx < (int64_t)N; ++idx) {
    c0[2 * idx] = bx[idx];
    c0[2 * idx + 1] = by[idx];
    c1[2 * idx] = c0[2 * idx + 1];
    c1[2 * idx + 1] = bz[idx];
  }
  fft3d(c0.data(), nx, ny, nz, +1);
  fft3d(c1.data(), nx, ny, nz, +1);
  // Hermitian extraction /N (brnd_es.cc:274-307)
  const double inv_n = 1. / N;
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)nx; ++i) {
    const uint64_t is = i == 0 ? 0 : nx - i;
    for (uint64_t j = 0; j < ny; ++j) {
      const uint64_t js = j == 0 ? 0 : ny - j;
      for (uint64_t l = 0; l < nz; ++l) {
        const uint64_t ls = l == 0 ? 0 : nz - l;
        const uint64_t idx = i * ny * nz + j * nz + l;
        const uint64_t sym = is * ny * nz + js * nz + ls;
        bx[idx] = 0.5 * (c0[2 * idx] + c0[2 * sym]) * inv_n;
        by[idx] = 0.5 * (c1[2 * idx] + c1[2 * sym]) * inv_n;
        bz[idx] = 0.5 * (c1[2 * idx + 1] + c1[2 * sym + 1]) * inv_n;
      }
    }
  }
  return 0;
}

int orc_brnd_mhd(const hamx_mhd_desc *d, uint64_t seed, double *bx, double *by, double *bz) {
  const uint64_t nx = d->box.nx, ny = d->box.ny, nz = d->box.nz;
  const uint64_t N = nx * ny * nz;
  std::vector<double> c0(2 * N, 0.), c1(2 * N, 0.);
  const double lx = d->box.x_max - d->box.x_min;
  const double ly = d->box.y_max - d->box.y_min;
  const double lz = d->box.z_max - d->box.z_min;
  hamx_es_desc e;
  std::memset(&e, 0, sizeof(e));
  e.breg_model = d->breg_model;
  std::memcpy(e.breg, d->breg, sizeof(e.breg));
  e.breg_x = d->breg_x;
  e.breg_y = d->breg_y;
  e.breg_z = d->breg_z;
  e.breg_box = d->breg_box;
  const V3 B = breg_at(&e, d->observer);
  const V3 bv = versor(B);
  const Mhd m{d};
  const double dk3 = kKpc * kKpc * kKpc / (lx * ly * lz);
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)nx; ++i) {
    for (uint64_t j = 0; j < ny; ++j) {
      for (uint64_t l = 0; l < nz; ++l) {
        const uint64_t idx = i * ny * nz + j * nz + l;
        if (idx == 0)
          continue;
        const V3 k = {{wavenum(i, nx, lx), wavenum(j, ny, ly), wavenum(l, nz, lz)}};
        const double ks = std::sqrt(lensq(k));
        const V3 kv = versor(k);
        V3 ep = normalised_unless_small(crossv(kv, bv));
        V3 em = normalised_unless_small(crossv(crossv(kv, bv), kv));
        double g0, g1, g2, g3;
        gauss_pair(seed, 3, 2 * idx, g0, g1);
        V3 b;
        if (lensq(ep) > 1e-6) {
          const double ang = dotv(bv, kv);
          const double Pa = m.spectrum(ks, d->aa0, d->pa0) * Mhd::F(d->ma, ang) * dk3;
          const double Pf = m.spectrum(ks, d->af0, d->pf0) * Mhd::h_f(d->beta, ang) * dk3;
          const double Ps = m.spectrum(ks, d->as0, d->ps0) * Mhd::F(d->ma, ang) *
                            Mhd::h_s(d->beta, ang) * dk3;
          gauss_pair(seed, 3, 2 * idx + 1, g2, g3);
          const double Ap = g0 * std::sqrt(2. * Pa);
          const double Am = g1 * std::sqrt(2. * Pf) + g2 * std::sqrt(2. * Ps);
          for (int c = 0; c < 3; ++c)
            b.e[c] = ep.e[c] * Ap + em.e[c] * Am;
        } else {
          const double Pf = m.spectrum(ks, d->af0, d->pf0) * Mhd::h_f(d->beta, 1) * dk3;
          if (i == 0 and j == 0) {
            ep.e[0] = k.e[0];
            em.e[1] = k.e[1];
          } else {
            const double kp = std::sqrt(k.e[0] * k.e[0] + k.e[1] * k.e[1]);
            ep = V3{{k.e[0] * k.e[2] / (ks * kp), k.e[1] * k.e[2] / (ks * kp), -kp / ks}};
            em = V3{{-k.e[1] / kp, k.e[0] / kp, 0.}};
          }
          const double Af = g0 * std::sqrt(2. * Pf);
          const double share = uniform53(seed, 4, idx);
          for (int c = 0; c < 3; ++c)
            b.e[c] = ep.e[c] * Af * share + em.e[c] * Af * (1. - share);
        }
        c0[2 * idx] = b.e[0];
        c0[2 * idx + 1] = b.e[1];
        c1[2 * idx] = b.e[1];
        c1[2 * idx + 1] = b.e[2];
      }
    }
  }
  fft3d(c0.data(), nx, ny, nz, +1);
  fft3d(c1.data(), nx, ny, nz, +1);
  // real parts, symmetrised (brnd_mhd.cc:139-167)
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)nx; ++i) {
    const uint64_t is = i == 0 ? 0 : nx - i;
    for (uint64_t j = 0; j < ny; ++j) {
      const uint64_t js = j == 0 ? 0 : ny - j;
      for (uint64_t l = 0; l < nz; ++l) {
        const uint64_t ls = l == 0 ? 0 : nz - l;
        const uint64_t idx = i * ny * nz + j * nz + l;
        const uint64_t sym = is * ny * nz + js * nz + ls;
        bx[idx] = 0.5 * (c0[2 * idx] + c0[2 * sym]);
        by[idx] = 0.5 * (c1[2 * idx] + c1[2 * sym]);
        bz[idx] = 0.5 * (c1[2 * idx + 1] + c1[2 * sym + 1]);
      }
    }
  }
  return 0;
}

// h_f, h_s and F_a (= F_s) of Brnd_mhd
void orc_mhd_terms(double beta, double ma, double cosa, double out[3]) {
  out[0] = Mhd::h_f(beta, cosa);
  out[1] = Mhd::h_s(beta, cosa);
  out[2] = Mhd::F(ma, cosa);
}

double orc_mhd_spectrum(const hamx_mhd_desc *d, double k, double a0, double p0) {
  return Mhd{d}.spectrum(k, a0, p0);
}

int orc_brnd_es(const hamx_es_desc *d, uint64_t seed, double *bx, double *by, double *bz) {
  return brnd_generate(d, seed, bx, by, bz, false);
}

int orc_brnd_jf12(const hamx_es_desc *d, uint64_t seed, double *bx, double *by, double *bz) {
  return brnd_generate(d, seed, bx, by, bz, true);
}

double orc_jf12_profile(const double *q, const double pos[3]) { return jf12_profile(q, pos); }

// TErnd_dft::write_grid (ternd_dft.cc:35-128)
int orc_ternd_dft(const hamx_dft_desc *d, uint64_t seed, double *te) {
  const uint64_t nx = d->box.nx, ny = d->box.ny, nz = d->box.nz;
  const uint64_t N = nx * ny * nz;
  std::vector<double> tk(2 * N, 0.);
  const double lx = d->box.x_max - d->box.x_min;
  const double ly = d->box.y_max - d->box.y_min;
  const double lz = d->box.z_max - d->box.z_min;
  const double dk3 = kKpc * kKpc * kKpc / (lx * ly * lz);
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)nx; ++i) {
    const double kx = wavenum(i, nx, lx);
    for (uint64_t j = 0; j < ny; ++j) {
      const double ky = wavenum(j, ny, ly);
      for (uint64_t l = 0; l < nz; ++l) {
        const uint64_t idx = i * ny * nz + j * nz + l;
        if (idx == 0)
          continue;
        const double kz = wavenum(l, nz, lz);
        const double ks = std::sqrt(kx * kx + ky * ky + kz * kz);
        const double sigma = std::sqrt(dft_spectrum(d, ks) * dk3);
        double g0, g1;
        gauss_pair(seed, 1, idx, g0, g1);
        tk[2 * idx] = sigma * g0;
        tk[2 * idx + 1] = sigma * g1;
      }
    }
  }
  fft3d(tk.data(), nx, ny, nz, +1);
  const double te_var_invsq = 1. / std::sqrt(variance(tk.data(), N));
#pragma omp parallel for
  for (int64_t i = 0; i < (int64_t)nx; ++i) {
    double pos[3] = {i * lx / (nx - 1) + d->box.x_min, 0, 0};
    for (uint64_t j = 0; j < ny; ++j) {
      pos[1] = j * ly / (ny - 1) + d->box.y_min;
      for (uint64_t l = 0; l < nz; ++l) {
        pos[2] = l * lz / (nz - 1) + d->box.z_min;
        const double ratio =
            std::sqrt(profile(d->observer, d->r0, d->z0, pos)) * d->rms * te_var_invsq;
        const uint64_t idx = i * ny * nz + j * nz + l;
        te[idx] = tk[2 * idx] * ratio;
      }
    }
  }
  return 0;
}

} // extern "C"
