// Gaussian random fields on gfx950: Brnd_es::write_grid (brnd_es.cc:76-308)
// and TErnd_dft::write_grid (ternd_dft.cc:35-128).
//
// Pipeline (every step a grid-stride elementwise kernel over the N vertices,
// or the LDS Stockham FFT of hamx_fft.hip):
//   ES : draw c0,c1 -> FFT(+) -> variance -> rescale+anisotropy -> FFT(-)
//        -> Gram-Schmidt (b = sqrt2 * GS(Re part)) -> reassemble -> FFT(+)
//        -> Hermitian extraction / N
//   DFT: draw te_k -> FFT(+) -> variance -> rescale
// The variance stays on the device (two-stage reduction), so the pipeline is
// one stream of launches with no host round trip.
//
// Gaussian variates: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)
// keyed by the seed, counter = (pair index, stream), two 53-bit uniforms per
// call and the Box-Muller transform.  A realisation is a function of the
// seed alone (the reference's per-thread GSL taus streams depend on the
// OpenMP thread count, brnd_es.cc:82-104).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "hamx.h"
#include "hamx_cgs.h"
#include "hamx_fields.h"
#include "hamx_internal.h"
#include "hamx_math.h"

namespace hamx {

// ------------------------------------------------------------------ Philox --
__device__ inline void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r > 0) {
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    // one 32 x 32 -> 64-bit product per multiplier (one v_mad_u64_u32
    // instead of separate low and high multiplies)
    const uint64_t p0 = static_cast<uint64_t>(0xD2511F53u) * c[0];
    const uint64_t p1 = static_cast<uint64_t>(0xCD9E8D57u) * c[2];
    const uint32_t lo0 = static_cast<uint32_t>(p0), hi0 = static_cast<uint32_t>(p0 >> 32);
    const uint32_t lo1 = static_cast<uint32_t>(p1), hi1 = static_cast<uint32_t>(p1 >> 32);
    const uint32_t n0 = hi1 ^ c[1] ^ k0;
    const uint32_t n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
}

__device__ inline void gauss_pair(uint64_t seed, uint32_t stream, uint64_t m,
                                  double &g0, double &g1) {
  uint32_t c[4] = {static_cast<uint32_t>(m), static_cast<uint32_t>(m >> 32), stream, 0u};
  philox4x32_10(c, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
  const uint64_t a = (static_cast<uint64_t>(c[1]) << 32) | c[0];
  const uint64_t b = (static_cast<uint64_t>(c[3]) << 32) | c[2];
  const double u1 = (static_cast<double>(a >> 11) + 1.0) * 0x1.0p-53;
  const double u2 = static_cast<double>(b >> 11) * 0x1.0p-53;
  const double r = sqrt(-2.0 * log_pos(u1));
  double s, co;
  sincos_fast(cgs::twopi * u2, s, co);
  g0 = r * co;
  g1 = r * s;
}

// A 53-bit uniform in [0, 1) from the same counter-based stream.
__device__ inline double philox_uniform(uint64_t seed, uint32_t stream, uint64_t m) {
  uint32_t c[4] = {static_cast<uint32_t>(m), static_cast<uint32_t>(m >> 32), stream, 0u};
  philox4x32_10(c, static_cast<uint32_t>(seed), static_cast<uint32_t>(seed >> 32));
  const uint64_t a = (static_cast<uint64_t>(c[1]) << 32) | c[0];
  return static_cast<double>(a >> 11) * 0x1.0p-53;
}

__global__ void k_gauss_pairs(double *out, uint64_t seed, uint32_t stream,
                              uint64_t first, uint64_t count) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t m = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       m < count; m += stride) {
    double g0, g1;
    gauss_pair(seed, stream, first + m, g0, g1);
    out[2 * m] = g0;
    out[2 * m + 1] = g1;
  }
}

// ------------------------------------------------------------------ common --
struct BoxArgs {
  uint64_t nx, ny, nz, n;
  double lx, ly, lz;
  double x_min, y_min, z_min;
  double dk3; // kpc^3 / (lx ly lz)
};

static BoxArgs box_args(const hamx_box &b) {
  BoxArgs a;
  a.nx = b.nx;
  a.ny = b.ny;
  a.nz = b.nz;
  a.n = b.nx * b.ny * b.nz;
  a.lx = b.x_max - b.x_min;
  a.ly = b.y_max - b.y_min;
  a.lz = b.z_max - b.z_min;
  a.x_min = b.x_min;
  a.y_min = b.y_min;
  a.z_min = b.z_min;
  a.dk3 = cgs::kpc * cgs::kpc * cgs::kpc / (a.lx * a.ly * a.lz);
  return a;
}

// (i, j, l) of the flat vertex index idx (x slowest, z fastest).  Integer
// division is a software routine on the GPU, several times longer in 64 bits
// than in 32: boxes below 2^32 vertices (all but the largest) split the index
// with 32-bit divisions.
__device__ inline void box_ijl(const BoxArgs &a, uint64_t idx, uint64_t &i, uint64_t &j,
                               uint64_t &l) {
  if (a.n <= 0xffffffffull) { // uniform
    const uint32_t id = static_cast<uint32_t>(idx);
    const uint32_t nz = static_cast<uint32_t>(a.nz), ny = static_cast<uint32_t>(a.ny);
    const uint32_t il = id / nz;
    const uint32_t i32 = il / ny;
    l = id - il * nz;
    j = il - i32 * ny;
    i = i32;
  } else {
    i = idx / (a.ny * a.nz);
    j = (idx / a.nz) % a.ny;
    l = idx % a.nz;
  }
}

// Wave number of index i on an axis of n points over length len
// (brnd_es.cc:106-122: periodic spacing n/L).
__device__ inline double wavenum(uint64_t i, uint64_t n, double len) {
  double k = cgs::kpc * static_cast<double>(i) / len;
  if (i >= (n + 1) / 2)
    k -= cgs::kpc * static_cast<double>(n) / len;
  return k;
}

// Vertex position (brnd_es.cc:165-173: vertex spacing L/(n-1)).
__device__ inline Vec3 vertex_pos(const BoxArgs &a, uint64_t i, uint64_t j,
                                  uint64_t l) {
  return Vec3{static_cast<double>(i) * a.lx / static_cast<double>(a.nx - 1) + a.x_min,
              static_cast<double>(j) * a.ly / static_cast<double>(a.ny - 1) + a.y_min,
              static_cast<double>(l) * a.lz / static_cast<double>(a.nz - 1) + a.z_min};
}

// Energy profile (brnd_es.cc:53-59, ternd_dft.cc:26-33).
__device__ inline double profile(const Vec3 &obs, double r0, double z0,
                                 const Vec3 &pos) {
  const double r_cyl = sqrt(pos.x * pos.x + pos.y * pos.y) - fabs(obs.x);
  const double z = fabs(pos.z) - fabs(obs.z);
  return exp_fast(-r_cyl / r0 - z / z0); // one exponential: the product's within a few ulp
}

__device__ inline uint64_t sym_index(const BoxArgs &a, uint64_t idx) {
  uint64_t i, j, l;
  box_ijl(a, idx, i, j, l);
  const uint64_t is = i == 0 ? 0 : a.nx - i;
  const uint64_t js = j == 0 ? 0 : a.ny - j;
  const uint64_t ls = l == 0 ? 0 : a.nz - l;
  return is * a.ny * a.nz + js * a.nz + ls;
}

// ---------------------------------------------------- variance (two stage) --
// toolkit::variance over the first n doubles (toolkit.h:48-78): mean, then
// the mean squared deviation.  Stage 1: per-block partial sums; stage 2: one
// block combines the partials and writes the result to out[slot].
constexpr int kRedBlocks = 1024;

__device__ inline double block_sum(double v) {
  __shared__ double sm[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o, 64);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0)
    sm[w] = v;
  __syncthreads();
  double t = 0.;
  if (threadIdx.x == 0)
    t = sm[0] + sm[1] + sm[2] + sm[3];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(256) void k_partial(const double *x, uint64_t n,
                                                 const double *mean,
                                                 double *partials) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  double acc = 0.;
  const double m = mean ? mean[0] : 0.;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < n; i += stride) {
    const double v = x[i];
    acc += mean ? (v - m) * (v - m) : v;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0)
    partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_final(const double *partials, int np,
                                               uint64_t n, double *out) {
  double acc = 0.;
  for (int i = threadIdx.x; i < np; i += blockDim.x)
    acc += partials[i];
  const double t = block_sum(acc);
  if (threadIdx.x == 0)
    out[0] = t / static_cast<double>(n);
}

// scratch layout: [0] mean, [1] variance, [16..16+kRedBlocks) partials
static int variance_dev(hamx_ctx *ctx, const double *x, uint64_t n,
                        double *scratch) {
  double *partials = scratch + 16;
  hipLaunchKernelGGL(k_partial, dim3(kRedBlocks), dim3(256), 0, ctx->stream, x, n,
                     static_cast<const double *>(nullptr), partials);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, ctx->stream, partials,
                     kRedBlocks, n, scratch);
  hipLaunchKernelGGL(k_partial, dim3(kRedBlocks), dim3(256), 0, ctx->stream, x, n,
                     static_cast<const double *>(scratch), partials);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, ctx->stream, partials,
                     kRedBlocks, n, scratch + 1);
  HAMX_HIP(ctx, hipGetLastError());
  return HAMX_OK;
}

// ---------------------------------------------------------------- Brnd_es --
struct EsArgs {
  BoxArgs b;
  Vec3 obs;
  double rms, k0, k1, a0, a1, rho, r0, z0;
  double c01; // (k0 / k1)^a1, the low band's constant factor
  int breg_model;
  int jf12; // Brnd_jf12: its energy profile and literal constants
  double jp[13];
  double third, twothirds, sqrt3; // 1/3, 2/3, sqrt(3) as the generator's source writes them
  double breg[HAMX_BREG_NPARAM];
  const double *bgx, *bgy, *bgz;
  GridGeo g_breg;
};

// Brnd_es::spectrum (brnd_es.cc:34-49)
__device__ inline double es_spectrum(const EsArgs &a, double k) {
  const double p0 = a.rms * a.rms;
  const double unit = 1. / (4. * cgs::pi * k * k);
  // the reference's sum of band-masked terms, with each power evaluated only
  // where its band is on (a masked-off term adds an exact zero)
  double P = 0.;
  if (k < a.k1) {
    const double x = k / a.k1, x3 = x * x * x;
    P += a.c01 * (x3 * x3); // (k0 / k1)^a1 (k / k1)^6
  }
  // 1 / (k / k0)^a as exp(-a ln(k / k0)): within a few ulp of the power
  if (k > a.k1 && k < a.k0)
    P += exp_fast(-a.a1 * log_pos(k / a.k0));
  if (k > a.k0)
    P += exp_fast(-a.a0 * log_pos(k / a.k0));
  return P * p0 * unit;
}

// Brnd_jf12::spatial_profile (brnd_jf12.cc:54-114): the squared disk (JF12
// spiral regions as in the regular field, b0_int inside 5 kpc, Gaussian in
// z) and halo amplitudes, in muG^2; 0 outside r < 20 kpc, |pos| > 1 kpc.
__device__ inline double jf12_profile(const double *p, const Vec3 &pos) {
  const double kpc = cgs::kpc;
  const double r = sqrt(pos.x * pos.x + pos.y * pos.y);
  const double rho = sqrt(pos.x * pos.x + pos.y * pos.y + pos.z * pos.z);
  const double z = pos.z;
  const double phi = atan2(pos.y, pos.x);
  if (r > 20. * kpc || rho < 1. * kpc)
    return 0.0;
  const double rc[8] = {5.1, 6.3, 7.1, 8.3, 9.8, 11.4, 12.7, 15.5};
  const double inc = 11.5;
  double disk = 0.0 * cgs::mugauss;
  if (r < 5. * kpc) {
    disk = p[8];
  } else {
    const double k = -1 / tan(cgs::pi / 180. * (90 - inc));
    double r_negx = r * exp(k * (phi - cgs::pi));
    if (r_negx > rc[7] * kpc)
      r_negx = r * exp(k * (phi + cgs::pi));
    if (r_negx > rc[7] * kpc)
      r_negx = r * exp(k * (phi + 3 * cgs::pi));
    for (int i = 7; i >= 0; --i)
      if (r_negx < rc[i] * kpc)
        disk = p[i] * (5. * kpc) / r;
  }
  disk = disk * exp(-0.5 * z * z / (p[9] * p[9]));
  const double halo = p[10] * exp(-fabs(r / p[11])) * exp(-fabs(z / p[12]));
  return (disk * disk + halo * halo) / (cgs::mugauss * cgs::mugauss);
}

// STEP I (brnd_es.cc:92-140)
__global__ __launch_bounds__(256) void k_es_draw(double2 *c0, double2 *c1,
                                                 EsArgs a, uint64_t seed) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < a.b.n; idx += stride) {
    if (idx == 0) {
      c0[0] = make_double2(0., 0.);
      c1[0] = make_double2(0., 0.);
      continue;
    }
    uint64_t i, j, l;
    box_ijl(a.b, idx, i, j, l);
    const double kx = wavenum(i, a.b.nx, a.b.lx);
    const double ky = wavenum(j, a.b.ny, a.b.ly);
    const double kz = wavenum(l, a.b.nz, a.b.lz);
    const double ks = sqrt(kx * kx + ky * ky + kz * kz);
    const double sigma = sqrt(a.third * es_spectrum(a, ks) * a.b.dk3);
    double g0, g1, g2, g3;
    gauss_pair(seed, 0, 2 * idx, g0, g1);
    gauss_pair(seed, 0, 2 * idx + 1, g2, g3);
    c0[idx] = make_double2(sigma * g0, sigma * g1);
    c1[idx] = make_double2(sigma * g2, sigma * g3);
  }
}

// STEP II (brnd_es.cc:157-205).  HEAVY compiles in the JF12 energy profile,
// the JF12 / Jaffe regular fields and the imported regular-field grid; the
// light kernel (LSA, cartesian, uniform, helix fields and the ES profile)
// then keeps few enough registers for several waves per SIMD.
template <bool HEAVY>
__global__ __launch_bounds__(256) void k_es_rescale(double2 *c0, double2 *c1,
                                                    EsArgs a, const double *var) {
  const double b_var_invsq = 1. / sqrt(3. * var[0]);
  const double rho = a.rho;
  const double rho2 = rho * rho;
  const double rhonorm = 1. / sqrt(a.third * rho2 + a.twothirds / rho2);
  const double inv_rho = 1. / rho;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < a.b.n; idx += stride) {
    uint64_t i, j, l;
    box_ijl(a.b, idx, i, j, l);
    const Vec3 pos = vertex_pos(a.b, i, j, l);
    const double prof =
        (HEAVY && a.jf12) ? jf12_profile(a.jp, pos) : profile(a.obs, a.r0, a.z0, pos);
    const double ratio = sqrt(prof) * a.rms * b_var_invsq;
    const double2 v0 = c0[idx], v1 = c1[idx];
    Vec3 b{v0.x * ratio, v0.y * ratio, v1.y * ratio};
    // anisotropy direction: versor of the regular field (brnd_es.cc:19-22)
    Vec3 braw;
    if (HEAVY && a.breg_model == HAMX_BREG_GRID) {
      Cell c;
      if (grid_cell(a.g_breg, pos, c)) {
        const uint64_t sy = a.g_breg.nz, sx = a.g_breg.ny * a.g_breg.nz;
        braw = Vec3{lerp_cell(a.bgx, c, sy, sx), lerp_cell(a.bgy, c, sy, sx),
                    lerp_cell(a.bgz, c, sy, sx)};
      } else {
        braw = Vec3{0., 0., 0.};
      }
    } else {
      braw = breg_eval<HEAVY>(a.breg_model, a.breg, pos);
    }
    Vec3 H{0., 0., 0.};
    double l2 = 0.;
    l2 += braw.x * braw.x;
    l2 += braw.y * braw.y;
    l2 += braw.z * braw.z;
    if (l2 != 0) {
      const double den = 1. / sqrt(l2);
      H = Vec3{braw.x * den, braw.y * den, braw.z * den};
    }
    double h2 = 0.;
    h2 += H.x * H.x;
    h2 += H.y * H.y;
    h2 += H.z * H.z;
    if (h2 < 1e-10)
      continue; // zero regular field: the raw FFT values stay
    const double hb = dot0(H, b);
    const Vec3 par{H.x * hb, H.y * hb, H.z * hb};
    const Vec3 perp{b.x - par.x, b.y - par.y, b.z - par.z};
    b = Vec3{(par.x * rho + perp.x * inv_rho) * rhonorm,
             (par.y * rho + perp.y * inv_rho) * rhonorm,
             (par.z * rho + perp.z * inv_rho) * rhonorm};
    c0[idx] = make_double2(b.x, b.y);
    c1[idx] = make_double2(b.y, b.z);
  }
}

// STEP III Gram-Schmidt (brnd_es.cc:211-259) of the symmetrised vector b at
// index idx: the part of b normal to the wave vector, times sqrt(3).
__device__ inline Vec3 es_gs_at(const EsArgs &a, uint64_t idx, const Vec3 &b) {
  uint64_t i, j, l;
  box_ijl(a.b, idx, i, j, l);
  const Vec3 k{wavenum(i, a.b.nx, a.b.lx), wavenum(j, a.b.ny, a.b.ly),
               wavenum(l, a.b.nz, a.b.lz)};
  double k2 = 0.;
  k2 += k.x * k.x;
  k2 += k.y * k.y;
  k2 += k.z * k.z;
  double b2 = 0.;
  b2 += b.x * b.x;
  b2 += b.y * b.y;
  b2 += b.z * b.z;
  if (k2 == 0 || b2 == 0)
    return b;
  const double inv = 1. / k2;
  const double kb = dot0(k, b);
  return Vec3{a.sqrt3 * (b.x - k.x * kb * inv), a.sqrt3 * (b.y - k.y * kb * inv),
              a.sqrt3 * (b.z - k.z * kb * inv)};
}

// STEP III and the reassembly c0 = bx + i by, c1 = by + i bz with
// b = sqrt(2) * the Gram-Schmidt result (brnd_es.cc:211-270), in place.  The
// symmetrised b of index idx reads idx and its mirror sym; the mirror's b is
// the same sum, so the thread of the lower index of each pair reads the two
// elements once and writes both results, each with its own wave vector.
__global__ __launch_bounds__(256) void k_es_gs_pairs(double2 *c0, double2 *c1, EsArgs a) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < a.b.n; idx += stride) {
    const uint64_t sym = sym_index(a.b, idx);
    if (sym < idx)
      continue; // the pair belongs to the thread of sym
    const double2 p0 = c0[idx], q0 = c0[sym], p1 = c1[idx], q1 = c1[sym];
    const Vec3 b{0.5 * (p0.x + q0.x), 0.5 * (p1.x + q1.x), 0.5 * (p1.y + q1.y)};
    const Vec3 fi = es_gs_at(a, idx, b);
    c0[idx] = make_double2(cgs::sqrtwo * fi.x, cgs::sqrtwo * fi.y);
    c1[idx] = make_double2(cgs::sqrtwo * fi.y, cgs::sqrtwo * fi.z);
    if (sym != idx) {
      const Vec3 fs = es_gs_at(a, sym, b);
      c0[sym] = make_double2(cgs::sqrtwo * fs.x, cgs::sqrtwo * fs.y);
      c1[sym] = make_double2(cgs::sqrtwo * fs.y, cgs::sqrtwo * fs.z);
    }
  }
}

// Hermitian extraction of the real field (brnd_es.cc:274-307 with
// scale = 1/N; brnd_mhd.cc:139-167 with scale = 1): the real part of the
// symmetrised pair, which is the same sum at index and mirror, so the
// thread of the lower index of each mirror pair reads the two elements once
// and writes both.
__global__ __launch_bounds__(256) void k_es_final(const double2 *c0, const double2 *c1,
                                                  double *bx, double *by, double *bz,
                                                  BoxArgs b, double inv_n) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < b.n; idx += stride) {
    const uint64_t sym = sym_index(b, idx);
    if (sym < idx)
      continue; // the pair belongs to the thread of sym
    const double2 p0 = c0[idx], q0 = c0[sym], p1 = c1[idx], q1 = c1[sym];
    const double vx = 0.5 * (p0.x + q0.x) * inv_n;
    const double vy = 0.5 * (p1.x + q1.x) * inv_n;
    const double vz = 0.5 * (p1.y + q1.y) * inv_n;
    bx[idx] = vx;
    by[idx] = vy;
    bz[idx] = vz;
    if (sym != idx) {
      bx[sym] = vx;
      by[sym] = vy;
      bz[sym] = vz;
    }
  }
}

// -------------------------------------------------------------- TErnd_dft --
struct DftArgs {
  BoxArgs b;
  Vec3 obs;
  double rms, k0, a0, r0, z0;
};

// TErnd_dft::spectrum (ternd_dft.cc:17-24)
__device__ inline double dft_spectrum(const DftArgs &a, double k) {
  if (k < a.k0)
    return 0.;
  return a.rms * exp(a.a0 * log(k / a.k0)) / (4 * cgs::pi * k * k); // (k / k0)^a0
}

__global__ __launch_bounds__(256) void k_dft_draw(double2 *tk, DftArgs a,
                                                  uint64_t seed) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < a.b.n; idx += stride) {
    if (idx == 0) {
      tk[0] = make_double2(0., 0.);
      continue;
    }
    uint64_t i, j, l;
    box_ijl(a.b, idx, i, j, l);
    const double kx = wavenum(i, a.b.nx, a.b.lx);
    const double ky = wavenum(j, a.b.ny, a.b.ly);
    const double kz = wavenum(l, a.b.nz, a.b.lz);
    const double ks = sqrt(kx * kx + ky * ky + kz * kz);
    const double sigma = sqrt(dft_spectrum(a, ks) * a.b.dk3);
    double g0, g1;
    gauss_pair(seed, 1, idx, g0, g1);
    tk[idx] = make_double2(sigma * g0, sigma * g1);
  }
}

__global__ __launch_bounds__(256) void k_dft_rescale(const double2 *tk, double *te,
                                                     DftArgs a, const double *var) {
  const double te_var_invsq = 1. / sqrt(var[0]);
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < a.b.n; idx += stride) {
    uint64_t i, j, l;
    box_ijl(a.b, idx, i, j, l);
    const Vec3 pos = vertex_pos(a.b, i, j, l);
    const double ratio = sqrt(profile(a.obs, a.r0, a.z0, pos)) * a.rms * te_var_invsq;
    te[idx] = tk[idx].x * ratio;
  }
}

// ------------------------------------------------------------------- host --
static unsigned ew_blocks(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  if (b > 65536)
    b = 65536;
  return static_cast<unsigned>(b ? b : 1);
}

static int check_gen_box(hamx_ctx *ctx, const hamx_box &b) {
  if (b.nx < 2 || b.ny < 2 || b.nz < 2)
    return set_err(ctx, HAMX_ERR_ARG, "random-field grid needs at least 2 vertices per axis");
  if (!fft_length_ok(b.nx) || !fft_length_ok(b.ny) || !fft_length_ok(b.nz))
    return set_err(ctx, HAMX_ERR_UNSUPPORTED,
                   "random-field grid %llux%llux%llu: FFT axes must be 2^a 3^b 5^c 7^d "
                   "up to 8192 or any length up to 4096",
                   static_cast<unsigned long long>(b.nx), static_cast<unsigned long long>(b.ny),
                   static_cast<unsigned long long>(b.nz));
  if (!(b.x_max > b.x_min) || !(b.y_max > b.y_min) || !(b.z_max > b.z_min))
    return set_err(ctx, HAMX_ERR_ARG, "random-field box limits are not increasing");
  return HAMX_OK;
}

} // namespace hamx

using namespace hamx;

extern "C" int hamx_gauss_pairs(hamx_ctx *ctx, uint64_t seed, uint32_t stream,
                                uint64_t first_pair, uint64_t count, double *out,
                                uint32_t flags) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!out && count)
    return set_err(ctx, HAMX_ERR_ARG, "hamx_gauss_pairs: null output");
  if (count == 0)
    return HAMX_OK;
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  double *d = out;
  const bool out_dev = (flags & HAMX_OUT_DEVICE) != 0;
  if (!out_dev) {
    void *w;
    HAMX_TRY(ws_get(ctx, kSlotAux, count * 2 * sizeof(double), &w));
    d = static_cast<double *>(w);
  }
  hipLaunchKernelGGL(k_gauss_pairs, dim3(ew_blocks(count)), dim3(256), 0, ctx->stream,
                     d, seed, stream, first_pair, count);
  HAMX_HIP(ctx, hipGetLastError());
  if (!out_dev)
    HAMX_HIP(ctx, hipMemcpyAsync(out, d, count * 2 * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}

namespace hamx {
int brnd_generate(hamx_ctx *ctx, const hamx_es_desc *d, uint64_t seed, double *bx,
                  double *by, double *bz, uint32_t flags, bool jf12) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!d || !bx || !by || !bz)
    return set_err(ctx, HAMX_ERR_ARG, "%s: null descriptor or output",
                   jf12 ? "hamx_brnd_jf12" : "hamx_brnd_es");
  HAMX_TRY(check_gen_box(ctx, d->box));
  if (d->breg_model < HAMX_BREG_NONE || d->breg_model > HAMX_BREG_JAFFE)
    return set_err(ctx, HAMX_ERR_UNSUPPORTED, "unsupported breg model %d", d->breg_model);
  if (!(d->rho > 0.))
    return set_err(ctx, HAMX_ERR_ARG, "anisotropy ratio rho must be positive");
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  const bool in_dev = (flags & HAMX_IN_DEVICE) != 0;
  const bool out_dev = (flags & HAMX_OUT_DEVICE) != 0;
  EsArgs a;
  std::memset(&a, 0, sizeof(a));
  a.b = box_args(d->box);
  a.obs = Vec3{d->observer[0], d->observer[1], d->observer[2]};
  a.rms = d->rms;
  a.k0 = d->k0;
  a.k1 = d->k1;
  a.a0 = d->a0;
  a.a1 = d->a1;
  a.c01 = std::pow(d->k0 / d->k1, d->a1);
  a.rho = d->rho;
  a.r0 = d->r0;
  a.z0 = d->z0;
  a.breg_model = d->breg_model;
  std::memcpy(a.breg, d->breg, sizeof(a.breg));
  a.jf12 = jf12 ? 1 : 0;
  std::memcpy(a.jp, d->jf12, sizeof(a.jp));
  // the generators' own literals (brnd_es.cc:132,190,71; brnd_jf12.cc:214,240,119)
  a.third = jf12 ? 0.33333333 : cgs::onethird;
  a.twothirds = jf12 ? 0.66666667 : cgs::twothirds;
  a.sqrt3 = jf12 ? 1.73205081 : cgs::sqrthree;
  if (d->breg_model == HAMX_BREG_GRID) {
    if (!(d->breg_x && d->breg_y && d->breg_z))
      return set_err(ctx, HAMX_ERR_ARG, "breg grid model without breg grid arrays");
    if (d->breg_box.nx < 2 || d->breg_box.ny < 2 || d->breg_box.nz < 2)
      return set_err(ctx, HAMX_ERR_ARG, "breg grid needs at least 2 vertices per axis");
    const size_t nb = d->breg_box.nx * d->breg_box.ny * d->breg_box.nz * sizeof(double);
    const void *p;
    HAMX_TRY(stage_in(ctx, kSlotBregX, d->breg_x, nb, in_dev, &p));
    a.bgx = static_cast<const double *>(p);
    HAMX_TRY(stage_in(ctx, kSlotBregY, d->breg_y, nb, in_dev, &p));
    a.bgy = static_cast<const double *>(p);
    HAMX_TRY(stage_in(ctx, kSlotBregZ, d->breg_z, nb, in_dev, &p));
    a.bgz = static_cast<const double *>(p);
    a.g_breg = grid_geo(d->breg_box);
  }
  const uint64_t n = a.b.n;
  void *w;
  HAMX_TRY(ws_get(ctx, kSlotRndC0, n * sizeof(double2), &w));
  double2 *c0 = static_cast<double2 *>(w);
  HAMX_TRY(ws_get(ctx, kSlotRndC1, n * sizeof(double2), &w));
  double2 *c1 = static_cast<double2 *>(w);
  HAMX_TRY(ws_get(ctx, kSlotReduce, (16 + kRedBlocks) * sizeof(double), &w));
  double *red = static_cast<double *>(w);
  double *ox = bx, *oy = by, *oz = bz;
  if (!out_dev) {
    HAMX_TRY(ws_get(ctx, kSlotRndOut0, n * sizeof(double), &w));
    ox = static_cast<double *>(w);
    HAMX_TRY(ws_get(ctx, kSlotRndOut1, n * sizeof(double), &w));
    oy = static_cast<double *>(w);
    HAMX_TRY(ws_get(ctx, kSlotRndOut2, n * sizeof(double), &w));
    oz = static_cast<double *>(w);
  }
  const unsigned blocks = ew_blocks(n);
  hipLaunchKernelGGL(k_es_draw, dim3(blocks), dim3(256), 0, ctx->stream, c0, c1, a, seed);
  HAMX_HIP(ctx, hipGetLastError());
  HAMX_TRY(fft3d_dev2(ctx, c0, c1, a.b.nx, a.b.ny, a.b.nz, +1));
  HAMX_TRY(variance_dev(ctx, reinterpret_cast<const double *>(c0), n, red));
  const bool heavy = a.jf12 || a.breg_model == HAMX_BREG_GRID ||
                     a.breg_model == HAMX_BREG_JF12 || a.breg_model == HAMX_BREG_JAFFE;
  if (heavy)
    hipLaunchKernelGGL(k_es_rescale<true>, dim3(blocks), dim3(256), 0, ctx->stream, c0, c1, a,
                       static_cast<const double *>(red + 1));
  else
    hipLaunchKernelGGL(k_es_rescale<false>, dim3(blocks), dim3(256), 0, ctx->stream, c0, c1,
                       a, static_cast<const double *>(red + 1));
  HAMX_HIP(ctx, hipGetLastError());
  HAMX_TRY(fft3d_dev2(ctx, c0, c1, a.b.nx, a.b.ny, a.b.nz, -1));
  hipLaunchKernelGGL(k_es_gs_pairs, dim3(blocks), dim3(256), 0, ctx->stream, c0, c1, a);
  HAMX_HIP(ctx, hipGetLastError());
  HAMX_TRY(fft3d_dev2(ctx, c0, c1, a.b.nx, a.b.ny, a.b.nz, +1));
  hipLaunchKernelGGL(k_es_final, dim3(blocks), dim3(256), 0, ctx->stream, c0, c1, ox,
                     oy, oz, a.b, 1. / static_cast<double>(n));
  HAMX_HIP(ctx, hipGetLastError());
  if (!out_dev) {
    HAMX_HIP(ctx, hipMemcpyAsync(bx, ox, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HAMX_HIP(ctx, hipMemcpyAsync(by, oy, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HAMX_HIP(ctx, hipMemcpyAsync(bz, oz, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}
// -------------------------------------------------------------- Brnd_mhd --
// Brnd_mhd::write_grid (brnd_mhd.cc:16-167): Alfven, fast and slow modes
// about the regular field at the observer, drawn in k-space and brought back
// by one backward transform per complex array; no energy-profile rescale.
struct MhdArgs {
  BoxArgs b;
  double pa0, pf0, ps0, k0, k1, aa0, af0, as0, a1, beta, ma;
};

// Brnd_mhd::spectrum_a / _f / _s (brnd_mhd.cc:211-262): one form, the mode
// choosing its index a0 and norm p0.
__device__ inline double mhd_spectrum(const MhdArgs &a, double k, double a0, double p0) {
  const double unit = 0.25 / (cgs::pi * k * k);
  const double band1 = static_cast<double>(k < a.k1);
  const double band2 = static_cast<double>(k > a.k1) * static_cast<double>(k < a.k0);
  const double band3 = static_cast<double>(k > a.k0);
  const double P = band1 * pow(a.k0 / a.k1, a.a1) * pow(k / a.k1, 6.0) +
                   band2 / pow(k / a.k0, a.a1) + band3 / pow(k / a.k0, a0);
  return P * p0 * unit;
}

// dynamo, h_s, h_f, F_a = F_s (bfield.h:388-412, brnd_mhd.cc:181-209)
__device__ inline double mhd_dynamo(double beta, double c) {
  return 1 + 0.25 * beta * beta + beta * (1. - 2. * c * c);
}
__device__ inline double mhd_h_s(double beta, double c) {
  const double sq = sqrt(mhd_dynamo(beta, c));
  const double dpp = 1 + sq + 0.5 * beta;
  const double dmm = 1 - sq - 0.5 * beta;
  const double dmp = 1 - sq + 0.5 * beta;
  const double den = 1 / (dmp * (c * c + dpp * dpp * (1 - c * c) / (dmm * dmm)));
  return isfinite(den) ? 2. * c * c * den * static_cast<double>(den > 0) : 0.;
}
__device__ inline double mhd_h_f(double beta, double c) {
  if (c == 0)
    return 0;
  const double sq = sqrt(mhd_dynamo(beta, c));
  const double dpp = 1 + sq + 0.5 * beta;
  const double dpm = 1 + sq - 0.5 * beta;
  const double dmp = 1 - sq + 0.5 * beta;
  const double den = 1 / (dpp * (c * c + dmp * dmp * (1 - c * c) / (dpm * dpm)));
  return isfinite(den) ? 2. * c * c * den : 0.;
}
__device__ inline double mhd_aniso(double ma, double c) {
  return exp(-1. * pow(ma, -1.33333333) * fabs(c) / pow(1 - c * c, 0.33333333));
}

// Hamvec::versor and crossprod (hamvec.h:176-205)
__device__ inline Vec3 versor3(const Vec3 &v) {
  double l2 = 0.;
  l2 += v.x * v.x;
  l2 += v.y * v.y;
  l2 += v.z * v.z;
  if (l2 == 0)
    return Vec3{0., 0., 0.};
  const double den = 1. / sqrt(l2);
  return Vec3{v.x * den, v.y * den, v.z * den};
}
__device__ inline Vec3 cross3(const Vec3 &a, const Vec3 &b) {
  return Vec3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// Brnd_mhd::e_plus / e_minus (brnd_mhd.cc:169-179): normalised unless
// their squared length is below 1e-12
__device__ inline Vec3 unit_or_small(const Vec3 &t) {
  double l2 = 0.;
  l2 += t.x * t.x;
  l2 += t.y * t.y;
  l2 += t.z * t.z;
  if (l2 < 1e-12)
    return t;
  const double l = sqrt(l2);
  return Vec3{t.x / l, t.y / l, t.z / l};
}

// The regular field at the observer (breg->read_field(observer)).
__global__ void k_mhd_bobs(EsArgs a, double *out) {
  if (blockIdx.x != 0 || threadIdx.x != 0)
    return;
  Vec3 b{0., 0., 0.};
  if (a.breg_model == HAMX_BREG_GRID) {
    Cell c;
    if (grid_cell(a.g_breg, a.obs, c)) {
      const uint64_t sy = a.g_breg.nz, sx = a.g_breg.ny * a.g_breg.nz;
      b = Vec3{lerp_cell(a.bgx, c, sy, sx), lerp_cell(a.bgy, c, sy, sx),
               lerp_cell(a.bgz, c, sy, sx)};
    }
  } else {
    b = breg_eval(a.breg_model, a.breg, a.obs);
  }
  out[0] = b.x;
  out[1] = b.y;
  out[2] = b.z;
}

// Philox streams of the MHD draws: Gaussians (2 pairs per mode) and the
// fast/slow power share of the modes along the field.
constexpr uint32_t kMhdGaussStream = 3, kMhdShareStream = 4;

__global__ __launch_bounds__(256) void k_mhd_draw(double2 *c0, double2 *c1, MhdArgs a,
                                                  uint64_t seed, const double *bobs) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  const Vec3 B{bobs[0], bobs[1], bobs[2]};
  const Vec3 bh = versor3(B);
  for (uint64_t idx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       idx < a.b.n; idx += stride) {
    if (idx == 0) {
      c0[0] = double2{0., 0.};
      c1[0] = double2{0., 0.};
      continue;
    }
    uint64_t i, j, l;
    box_ijl(a.b, idx, i, j, l);
    const Vec3 k{wavenum(i, a.b.nx, a.b.lx), wavenum(j, a.b.ny, a.b.ly),
                 wavenum(l, a.b.nz, a.b.lz)};
    const double ks = length0(k);
    const Vec3 kh = versor3(k);
    const Vec3 kxb = cross3(kh, bh);
    Vec3 ep = unit_or_small(kxb);
    Vec3 em = unit_or_small(cross3(kxb, kh));
    const double dk3 = a.b.dk3;
    double g0, g1, g2, g3;
    gauss_pair(seed, kMhdGaussStream, 2 * idx, g0, g1);
    Vec3 bk;
    if (dot0(ep, ep) > 1e-6) {
      const double ang = dot0(bh, kh);
      const double Pa = mhd_spectrum(a, ks, a.aa0, a.pa0) * mhd_aniso(a.ma, ang) * dk3;
      const double Pf = mhd_spectrum(a, ks, a.af0, a.pf0) * mhd_h_f(a.beta, ang) * dk3;
      const double Ps = mhd_spectrum(a, ks, a.as0, a.ps0) * mhd_aniso(a.ma, ang) *
                        mhd_h_s(a.beta, ang) * dk3;
      gauss_pair(seed, kMhdGaussStream, 2 * idx + 1, g2, g3);
      const double Ap = g0 * sqrt(2. * Pa);
      const double Am = g1 * sqrt(2. * Pf) + g2 * sqrt(2. * Ps);
      bk = Vec3{ep.x * Ap + em.x * Am, ep.y * Ap + em.y * Am, ep.z * Ap + em.z * Am};
    } else {
      const double Pf = mhd_spectrum(a, ks, a.af0, a.pf0) * mhd_h_f(a.beta, 1) * dk3;
      if (i == 0 && j == 0) {
        ep.x = k.x;
        em.y = k.y;
      } else {
        const double kp = sqrt(k.x * k.x + k.y * k.y);
        ep = Vec3{k.x * k.z / (ks * kp), k.y * k.z / (ks * kp), -kp / ks};
        em = Vec3{-k.y / kp, k.x / kp, 0.};
      }
      const double Af = g0 * sqrt(2. * Pf);
      const double share = philox_uniform(seed, kMhdShareStream, idx);
      bk = Vec3{ep.x * Af * share + em.x * Af * (1. - share),
                ep.y * Af * share + em.y * Af * (1. - share),
                ep.z * Af * share + em.z * Af * (1. - share)};
    }
    c0[idx] = double2{bk.x, bk.y};
    c1[idx] = double2{bk.y, bk.z};
  }
}

int brnd_mhd(hamx_ctx *ctx, const hamx_mhd_desc *d, uint64_t seed, double *bx, double *by,
             double *bz, uint32_t flags) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!d || !bx || !by || !bz)
    return set_err(ctx, HAMX_ERR_ARG, "hamx_brnd_mhd: null descriptor or output");
  HAMX_TRY(check_gen_box(ctx, d->box));
  if (d->breg_model < HAMX_BREG_NONE || d->breg_model > HAMX_BREG_JAFFE)
    return set_err(ctx, HAMX_ERR_UNSUPPORTED, "unsupported breg model %d", d->breg_model);
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  const bool in_dev = (flags & HAMX_IN_DEVICE) != 0;
  const bool out_dev = (flags & HAMX_OUT_DEVICE) != 0;
  EsArgs e;
  std::memset(&e, 0, sizeof(e));
  e.obs = Vec3{d->observer[0], d->observer[1], d->observer[2]};
  e.breg_model = d->breg_model;
  std::memcpy(e.breg, d->breg, sizeof(e.breg));
  if (d->breg_model == HAMX_BREG_GRID) {
    if (!(d->breg_x && d->breg_y && d->breg_z))
      return set_err(ctx, HAMX_ERR_ARG, "breg grid model without breg grid arrays");
    if (d->breg_box.nx < 2 || d->breg_box.ny < 2 || d->breg_box.nz < 2)
      return set_err(ctx, HAMX_ERR_ARG, "breg grid needs at least 2 vertices per axis");
    const size_t nb = d->breg_box.nx * d->breg_box.ny * d->breg_box.nz * sizeof(double);
    const void *p;
    HAMX_TRY(stage_in(ctx, kSlotBregX, d->breg_x, nb, in_dev, &p));
    e.bgx = static_cast<const double *>(p);
    HAMX_TRY(stage_in(ctx, kSlotBregY, d->breg_y, nb, in_dev, &p));
    e.bgy = static_cast<const double *>(p);
    HAMX_TRY(stage_in(ctx, kSlotBregZ, d->breg_z, nb, in_dev, &p));
    e.bgz = static_cast<const double *>(p);
    e.g_breg = grid_geo(d->breg_box);
  }
  MhdArgs a;
  std::memset(&a, 0, sizeof(a));
  a.b = box_args(d->box);
  a.pa0 = d->pa0;
  a.pf0 = d->pf0;
  a.ps0 = d->ps0;
  a.k0 = d->k0;
  a.k1 = d->k1;
  a.aa0 = d->aa0;
  a.af0 = d->af0;
  a.as0 = d->as0;
  a.a1 = d->a1;
  a.beta = d->beta;
  a.ma = d->ma;
  const uint64_t n = a.b.n;
  void *w;
  HAMX_TRY(ws_get(ctx, kSlotRndC0, n * sizeof(double2), &w));
  double2 *c0 = static_cast<double2 *>(w);
  HAMX_TRY(ws_get(ctx, kSlotRndC1, n * sizeof(double2), &w));
  double2 *c1 = static_cast<double2 *>(w);
  HAMX_TRY(ws_get(ctx, kSlotReduce, (16 + kRedBlocks) * sizeof(double), &w));
  double *bobs = static_cast<double *>(w);
  double *ox = bx, *oy = by, *oz = bz;
  if (!out_dev) {
    HAMX_TRY(ws_get(ctx, kSlotRndOut0, n * sizeof(double), &w));
    ox = static_cast<double *>(w);
    HAMX_TRY(ws_get(ctx, kSlotRndOut1, n * sizeof(double), &w));
    oy = static_cast<double *>(w);
    HAMX_TRY(ws_get(ctx, kSlotRndOut2, n * sizeof(double), &w));
    oz = static_cast<double *>(w);
  }
  hipLaunchKernelGGL(k_mhd_bobs, dim3(1), dim3(64), 0, ctx->stream, e, bobs);
  HAMX_HIP(ctx, hipGetLastError());
  const unsigned blocks = ew_blocks(n);
  hipLaunchKernelGGL(k_mhd_draw, dim3(blocks), dim3(256), 0, ctx->stream, c0, c1, a, seed,
                     static_cast<const double *>(bobs));
  HAMX_HIP(ctx, hipGetLastError());
  HAMX_TRY(fft3d_dev2(ctx, c0, c1, a.b.nx, a.b.ny, a.b.nz, +1));
  hipLaunchKernelGGL(k_es_final, dim3(blocks), dim3(256), 0, ctx->stream, c0, c1, ox, oy, oz,
                     a.b, 1.);
  HAMX_HIP(ctx, hipGetLastError());
  if (!out_dev) {
    HAMX_HIP(ctx, hipMemcpyAsync(bx, ox, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HAMX_HIP(ctx, hipMemcpyAsync(by, oy, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HAMX_HIP(ctx, hipMemcpyAsync(bz, oz, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}
} // namespace hamx

extern "C" int hamx_brnd_es(hamx_ctx *ctx, const hamx_es_desc *d, uint64_t seed,
                            double *bx, double *by, double *bz, uint32_t flags) {
  return brnd_generate(ctx, d, seed, bx, by, bz, flags, false);
}

extern "C" int hamx_brnd_jf12(hamx_ctx *ctx, const hamx_es_desc *d, uint64_t seed,
                              double *bx, double *by, double *bz, uint32_t flags) {
  return brnd_generate(ctx, d, seed, bx, by, bz, flags, true);
}

extern "C" int hamx_ternd_dft(hamx_ctx *ctx, const hamx_dft_desc *d, uint64_t seed,
                              double *te, uint32_t flags) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!d || !te)
    return set_err(ctx, HAMX_ERR_ARG, "hamx_ternd_dft: null descriptor or output");
  HAMX_TRY(check_gen_box(ctx, d->box));
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  const bool out_dev = (flags & HAMX_OUT_DEVICE) != 0;
  DftArgs a;
  std::memset(&a, 0, sizeof(a));
  a.b = box_args(d->box);
  a.obs = Vec3{d->observer[0], d->observer[1], d->observer[2]};
  a.rms = d->rms;
  a.k0 = d->k0;
  a.a0 = d->a0;
  a.r0 = d->r0;
  a.z0 = d->z0;
  const uint64_t n = a.b.n;
  void *w;
  HAMX_TRY(ws_get(ctx, kSlotRndC0, n * sizeof(double2), &w));
  double2 *tk = static_cast<double2 *>(w);
  HAMX_TRY(ws_get(ctx, kSlotReduce, (16 + kRedBlocks) * sizeof(double), &w));
  double *red = static_cast<double *>(w);
  double *ot = te;
  if (!out_dev) {
    HAMX_TRY(ws_get(ctx, kSlotRndOut0, n * sizeof(double), &w));
    ot = static_cast<double *>(w);
  }
  const unsigned blocks = ew_blocks(n);
  hipLaunchKernelGGL(k_dft_draw, dim3(blocks), dim3(256), 0, ctx->stream, tk, a, seed);
  HAMX_HIP(ctx, hipGetLastError());
  HAMX_TRY(fft3d_dev(ctx, tk, a.b.nx, a.b.ny, a.b.nz, +1));
  HAMX_TRY(variance_dev(ctx, reinterpret_cast<const double *>(tk), n, red));
  hipLaunchKernelGGL(k_dft_rescale, dim3(blocks), dim3(256), 0, ctx->stream, tk, ot, a,
                     static_cast<const double *>(red + 1));
  HAMX_HIP(ctx, hipGetLastError());
  if (!out_dev)
    HAMX_HIP(ctx, hipMemcpyAsync(te, ot, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}

extern "C" int hamx_brnd_mhd(hamx_ctx *ctx, const hamx_mhd_desc *d, uint64_t seed,
                             double *bx, double *by, double *bz, uint32_t flags) {
  return brnd_mhd(ctx, d, seed, bx, by, bz, flags);
}
