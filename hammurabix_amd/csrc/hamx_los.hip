// Line-of-sight integrator for gfx950: Integrator::write_grid /
// radial_integration (integrator.cc:23-279) as HIP kernels.
//
// Per shell s (nside_s, step_s samples at dist_k = d0 + (k + 1/2) delta):
//   k_los_wave / k_los_thread : march every ray of the shell, producing the
//       shell maps DM, FD, I, Q, U (I/Q/U temp_convert-ed, integrator.cc:99-115)
//   k_accumulate               : add the shell maps into the output maps,
//       directly or by HEALPix interpolation at the output pixel centres
//       (hamdis.h:698-721)
// FD of the inner shells reaches shell s through the accumulated FD map,
// interpolated at the ray's direction (integrator.cc:78-86).
//
// Two march schedules share the per-sample arithmetic:
//   wave   : one 64-lane wavefront per ray, lane j evaluating sample
//            base + j; the running Faraday depth is carried across chunks
//            and formed inside a chunk by an inclusive wave scan, so every
//            sample's polarisation phase sees the FD of all samples up to
//            and including itself (integrator.cc:209,221).  Consecutive
//            lanes sit 1 step apart on the ray, so the 8-corner gathers of a
//            wave fall in a handful of grid cells (coalesced, L2-resident).
//   thread : one thread per ray, samples summed in the reference's order.
// Both skip, per ray, the step ranges that provably fail the galactocentric
// radius / height limits (a conservative analytic interval, then the exact
// per-sample checks of integrator.cc:179-186): skipped samples contribute
// nothing in the reference either.
//
// The translation unit is compiled with -ffp-contract=off and keeps the
// reference's operation order, so the FP64 maps agree with the CPU
// restatement to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "hamx.h"
#include "hamx_cgs.h"
#include "hamx_fields.h"
#include "hamx_healpix.h"
#include "hamx_internal.h"
#include "hamx_march.h"
#include "hamx_sync.h"

namespace hamx {

// ------------------------------------------------------------- per sample --

struct SampleOut {
  double te;   // clamped density
  double fdc;  // te * B_par * fd_fore * delta
  double jt;   // total emissivity * delta * i2bt
  double jp;   // polarised emissivity * delta * i2bt
  double ipa;  // intrinsic polarisation angle
};

// CRE_ana / CRE_unif spatial profile (cre_ana.cc:11-19, cre_unif.cc)
__device__ inline double cre_profile(const Phys &ph, const Vec3 &pos) {
  if (ph.cre_model == HAMX_CRE_ANA) {
    const double r = sqrt(pos.x * pos.x + pos.y * pos.y);
    const double ch = cosh(pos.z / ph.cre[4]);
    return exp((ph.R0 - r) / ph.cre[3]) * (1. / (ch * ch));
  }
  const Vec3 d{pos.x - ph.obs.x, pos.y - ph.obs.y, pos.z - ph.obs.z};
  return (length0(d) > ph.cre[1]) ? 0. : 1.;
}

// CRE_ana::flux_idx (cre_ana.cc:22-28)
__device__ inline double cre_index(const Phys &ph, const Vec3 &pos) {
  const double r = sqrt(pos.x * pos.x + pos.y * pos.y) / cgs::kpc;
  const double z = fabs(pos.z) / cgs::kpc;
  return -ph.cre[0] + ph.cre[1] * r + ph.cre[2] * z;
}

// CREfield::read_grid_num (cre.cc:144-199) of energy bin e in a cell found by
// grid_cell on the CRE box: the same z, y, x interpolation order over the
// [x][y][z][E] layout (z neighbours nE apart).
__device__ inline double lerp_cell_e(const double *f, const Cell &c, uint64_t nE,
                                     uint64_t sy, uint64_t sx, uint64_t e) {
  const uint64_t i00 = c.idx * nE + e, i01 = i00 + sy, i10 = i00 + sx, i11 = i00 + sx + sy;
  const double i1 = f[i00] * (1. - c.zd) + f[i00 + nE] * c.zd;
  const double i2 = f[i01] * (1 - c.zd) + f[i01 + nE] * c.zd;
  const double j1 = f[i10] * (1 - c.zd) + f[i10 + nE] * c.zd;
  const double j2 = f[i11] * (1 - c.zd) + f[i11 + nE] * c.zd;
  const double w1 = i1 * (1 - c.yd) + i2 * c.yd;
  const double w2 = j1 * (1 - c.yd) + j2 * c.yd;
  return w1 * (1 - c.xd) + w2 * c.xd;
}

// sync_emissivity_t and _p with the CRE grid (integrator.cc:286-325 and
// 355-394) for one sample: the midpoint rule over the energy bins of
// F(x) resp. G(x) times the bin-mean flux / beta, bins whose mean x exceeds
// 100 skipped, then the fore factor.  Each bin's gather is shared by J_tot
// and J_pol and reused by the next bin.  Outside the grid every gathered
// flux is 0 and so are both sums.
// The kernels of one bin come from one evaluation: u = ln(x_fact) +
// ln((1/E_{i+1}^2 + 1/E_i^2) / 2) (the second term tabulated per bin) in
// place of ln(xv), and G at the J_pol mean xvp, which may differ from the
// J_tot mean xvt in the last bits (the two x_fact products round
// differently), as G(xvt) (1 + xvt - xvp), exact to first order.
__device__ inline void cre_grid_emissivity(const Phys &ph, const Grids &gr, const Vec3 &pos,
                                           double b_per, const double *stab, double &jt,
                                           double &jp) {
  jt = 0.;
  jp = 0.;
  Cell c;
  if (!grid_cell(gr.g_cre, pos, c))
    return;
  const double ab = fabs(b_per);
  const double den = ph.three_qe * ab;
  const double xft = ph.xft_num / den;
  const double xfp = ph.xfp_num / den;
  const uint64_t nE = gr.cre_nE;
  const uint64_t sy = gr.g_cre.nz * nE, sx = gr.g_cre.ny * sy;
  const double *ke2 = gr.cre_tab, *beta = ke2 + nE, *dE = beta + nE, *lnc = dE + nE;
  const double lxt = log(xft);
  double xt0 = xft / ke2[0], xp0 = xfp / ke2[0];
  double g0 = 0.;
  bool have0 = false;
  double Jt = 0., Jp = 0.;
  for (uint64_t i = 0; i + 1 < nE; ++i) {
    const double xt1 = xft / ke2[i + 1], xp1 = xfp / ke2[i + 1];
    const double xvt = 0.5 * (xt1 + xt0), xvp = 0.5 * (xp1 + xp0);
    const bool on_t = !(xvt > 100), on_p = !(xvp > 100);
    if (on_t || on_p) {
      if (!have0)
        g0 = lerp_cell_e(gr.cre, c, nE, sy, sx, i) / beta[i];
      const double g1 = lerp_cell_e(gr.cre, c, nE, sy, sx, i + 1) / beta[i + 1];
      const double flux = 0.5 * (g1 + g0);
      double F, G;
      sync_fg(stab, lxt + lnc[i], xvt, F, G);
      if (on_t)
        Jt += F * flux * dE[i];
      if (on_p)
        Jp += G * (1. + (xvt - xvp)) * flux * dE[i];
      g0 = g1;
      have0 = true;
    } else {
      have0 = false;
    }
    xt0 = xt1;
    xp0 = xp1;
  }
  const double fore = ph.fore_num * ab / ph.fore_den;
  jt = Jt * fore;
  jp = Jp * fore;
}

template <bool PACKED, bool CREG, bool HEAVY>
__device__ inline bool eval_sample(const Phys &ph, const Grids &gr,
                                   const ShellArgs &sa, const RayGeo &r,
                                   uint64_t k, SampleOut &s, const double *stab) {
  const double dist =
      sa.d_start + (static_cast<double>(static_cast<uint32_t>(k)) + 0.5) * sa.delta;
  const Vec3 pos{r.l.x * dist + ph.obs.x, r.l.y * dist + ph.obs.y,
                 r.l.z * dist + ph.obs.z};
  // limit checks (integrator.cc:179-186)
  const double len = length0(pos);
  if (len < ph.gc_r_min)
    return false;
  if (len > ph.gc_r_max)
    return false;
  if (pos.z < ph.gc_z_min)
    return false;
  if (pos.z > ph.gc_z_max)
    return false;
  // regular + random magnetic field (integrator.cc:188-190)
  Vec3 B;
  if (ph.breg_model == HAMX_BREG_GRID)
    B = gather3(gr.breg_x, gr.breg_y, gr.breg_z, gr.g_breg, pos);
  else
    B = breg_eval<HEAVY>(ph.breg_model, ph.breg, pos);
  // regular + random thermal electrons (integrator.cc:197-201)
  double te = 0.;
  if (ph.tereg_model == HAMX_TEREG_UNIF)
    te = tereg_unif(ph.tereg, pos, ph.obs);
  else if (HEAVY && ph.tereg_model == HAMX_TEREG_YMW16)
    te = tereg_ymw16_eval(ph.tereg, pos.x, pos.y, pos.z);
  else if (ph.tereg_model == HAMX_TEREG_GRID)
    te = gather1(gr.tereg, gr.g_tereg, pos);
  if (PACKED) {
    Cell c;
    Vec3 br{0., 0., 0.};
    double tr = 0.;
    if (grid_cell(gr.g_packed, pos, c))
      lerp_cell_packed(gr.packed, c, gr.g_packed.nz,
                       gr.g_packed.ny * gr.g_packed.nz, br, tr);
    B.x += br.x;
    B.y += br.y;
    B.z += br.z;
    te += tr;
  } else {
    if (gr.brnd_x) {
      const Vec3 br = gather3(gr.brnd_x, gr.brnd_y, gr.brnd_z, gr.g_brnd, pos);
      B.x += br.x;
      B.y += br.y;
      B.z += br.z;
    }
    if (gr.ternd)
      te += gather1(gr.ternd, gr.g_ternd, pos);
  }
  te *= static_cast<double>(te > 0.);
  const double b_par = dot0(B, r.l);
  s.te = te;
  s.fdc = 0.;
  if (ph.do_fd || ph.do_sync)
    s.fdc = te * b_par * ph.fd_fore * sa.delta;
  if (ph.do_sync) {
    // los_perproj (integrator.h:39-42)
    const Vec3 cr{B.y * r.l.z - B.z * r.l.y, B.z * r.l.x - B.x * r.l.z,
                  B.x * r.l.y - B.y * r.l.x};
    const double b_per = length0(cr);
    if (CREG) {
      double jt, jp;
      cre_grid_emissivity(ph, gr, pos, b_per, stab, jt, jp);
      s.jt = jt * sa.delta * ph.i2bt;
      s.jp = jp * sa.delta * ph.i2bt;
      s.ipa = atan2(-dot0(r.eth, B), -dot0(r.eph, B));
      return true;
    }
    // analytic emissivities (integrator.cc:333-346, 402-415)
    double index, g1, g2, g3;
    if (ph.idx_const) {
      index = ph.idx;
      g1 = ph.g_t1;
      g2 = ph.g_t2;
      g3 = ph.g_p1;
    } else {
      index = cre_index(ph, pos);
      g1 = gamma_lanczos(-0.25 * index + 19. / 12.);
      g2 = gamma_lanczos(-0.25 * index - 1. / 12.);
      g3 = gamma_lanczos(-0.25 * index + 7. / 12.);
    }
    const double flux_norm = ph.cre_nu * cre_profile(ph, pos);
    const double abs_bper = fabs(b_per);
    const double norm_t = flux_norm * 1.73205081 * ph.qe3 * abs_bper /
                          (ph.fourpi_mec2 * (1. - index));
    const double norm_p =
        flux_norm * 1.73205081 * ph.qe3 * abs_bper / ph.sixteenpi_mec2;
    const double A = ph.A_num / (ph.three_qe * abs_bper);
    const double pw = pow(A, 0.5 * (index + 1));
    const double jt = norm_t * (pw * g1 * g2);
    const double jp = norm_p * (pw * g3 * g2);
    s.jt = jt * sa.delta * ph.i2bt;
    s.jp = jp * sa.delta * ph.i2bt;
    // sync_ipa (integrator.cc:257-265)
    s.ipa = atan2(-dot0(r.eth, B), -dot0(r.eph, B));
  }
  return true;
}

// Shell-map epilogue: temp_convert of I/Q/U and the stores
// (integrator.cc:99-115).
__device__ inline void store_ray(const Phys &ph, const ShellArgs &sa,
                                 uint64_t ipix, double dm, double fd, double I,
                                 double Q, double U) {
  if (ph.do_dm)
    sa.t_dm[ipix] = dm;
  if (ph.do_sync) {
    sa.t_is[ipix] = I * ph.tc_em1 * ph.tc_em1 / ph.tc_den;
    sa.t_qs[ipix] = Q * ph.tc_em1 * ph.tc_em1 / ph.tc_den;
    sa.t_us[ipix] = U * ph.tc_em1 * ph.tc_em1 / ph.tc_den;
  }
  if (ph.do_fd || ph.do_sync)
    sa.t_fd[ipix] = fd;
}

// Masked-out pixel (integrator.cc:63-70): every observable of the shell is
// zero.  True when the caller should return.
__device__ inline bool masked_out(const Phys &ph, const ShellArgs &sa, uint64_t ipix,
                                  bool writer) {
  if (!sa.mask || sa.mask[ipix] == 1.0)
    return false;
  if (writer)
    store_ray(ph, sa, ipix, 0., 0., 0., 0., 0.);
  return true;
}

// Left to itself the register allocator gives these generic marches 256
// VGPRs plus AGPRs (one wave per SIMD).  Two waves per SIMD (a few dozen
// spilled registers) measured 1.7x faster for a JF12 field on the
// random_forward workload; with the CRE grid's energy loop three waves
// measured 2.1x faster than one (2.3x slower for JF12, hence per case).
template <bool PACKED, bool CREG, bool HEAVY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CREG ? 3 : 2))) void
k_los_wave(Phys ph, Grids gr, ShellArgs sa) {
  // the synchrotron kernel tables of the CRE grid emissivities, in LDS
  __shared__ double stab[CREG ? kSyncTab : 1];
  if (CREG) {
    load_sync_tab(stab);
    __syncthreads();
  }
  const int lane = threadIdx.x & 63;
  const uint64_t ipix =
      static_cast<uint64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (ipix >= sa.hs.npix)
    return; // uniform per wavefront
  if (masked_out(ph, sa, ipix, lane == 0))
    return; // uniform per wavefront
  RayGeo r;
  ray_setup(ph, sa, ipix, r);
  double dm = 0., I = 0., Q = 0., U = 0., fd_part = 0., fd_carry = 0.;
  for (uint64_t base = r.k_lo; base < r.k_hi; base += 64) {
    const uint64_t k = base + lane;
    SampleOut s;
    bool valid = false;
    if (k < r.k_hi)
      valid = eval_sample<PACKED, CREG, HEAVY>(ph, gr, sa, r, k, s, stab);
    const double fdk = valid ? s.fdc : 0.;
    if (ph.do_sync) {
      const double incl = wave_incl_scan(fdk, lane);
      const double fd_here = fd_carry + incl;
      fd_carry += wave_last(incl);
      if (valid) {
        const double qui = (r.fd_inner + fd_here) * ph.lambda_sq + s.ipa;
        double s2, c2;
        sincos(2. * qui, &s2, &c2);
        I += s.jt;
        Q += c2 * s.jp;
        U += s2 * s.jp;
      }
    } else {
      fd_part += fdk;
    }
    if (valid && ph.do_dm)
      dm += s.te * sa.delta;
  }
  dm = wave_sum(dm);
  double fd = ph.do_sync ? fd_carry : wave_sum(fd_part);
  I = wave_sum(I);
  Q = wave_sum(Q);
  U = wave_sum(U);
  if (lane == 0)
    store_ray(ph, sa, ipix, dm, fd, I, Q, U);
}

// ---------------------------------------------------- specialised march --
// The common model set -- WMAP-LSA regular B, uniform-sphere thermal
// electrons, a CRE model with a constant spectral index, random grids (if
// any) packed in one box, synchrotron on -- with the per-sample arithmetic
// rewritten into fewer, cheaper operations that agree with the reference to
// rounding:
//   * the IAU angle enters only through cos/sin(2 ipa), formed from the two
//     projections x = -e_phi.B, y = -e_theta.B as ((x^2 - y^2), 2xy)/(x^2+y^2)
//     instead of atan2, so one sincos of the Faraday phase gives
//     cos/sin(2 qui) by the angle-sum identity;
//   * the emissivities fold every per-shell constant (Gamma factors, CGS
//     products, delta, i2bt, the CRE norm) into one factor each, and when
//     the exponent (index+1)/2 is -1 (alpha = 3) A^-1 is a product;
//   * divisions by constants that feed no decision become products by
//     reciprocals.  The galactocentric limits, the LSA radius window, the
//     TE sphere test, the te > 0 clamp and the grid-cell lookup (divisions
//     kept) decide exactly as in the reference.

template <bool GRIDS, bool PW1>
// Four waves per SIMD (at most 128 VGPRs; the allocator moves ~90 scalar
// values into a 96-byte private segment) measured 4% faster than three
// (7.30 -> 7.01 ms per pass on the random_forward workload), which in turn
// measured 8% faster than the unconstrained choice of two.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_los_fast(
    Phys ph, FastConst fc, Grids gr, ShellArgs sa) {
  // kFastRays consecutive rays per wavefront: their set-up (pointing,
  // versors, FD cache, culled step range) runs one ray per lane into LDS,
  // then the rays are marched one after another, 64 samples per step
  __shared__ RayGeo slots[4][kFastRays];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t p0 = (static_cast<uint64_t>(blockIdx.x) * 4 + wave) * kFastRays;
  if (p0 >= sa.hs.npix)
    return; // uniform per wavefront
  const uint64_t nr = (sa.hs.npix - p0 < static_cast<uint64_t>(kFastRays))
                          ? sa.hs.npix - p0
                          : static_cast<uint64_t>(kFastRays);
  if (static_cast<uint64_t>(lane) < nr)
    ray_setup(ph, sa, p0 + lane, slots[wave][lane]);
  wave_lds_sync();
  const double two_l2 = 2. * ph.lambda_sq;
  for (uint64_t j = 0; j < nr; ++j) {
    const uint64_t ipix = p0 + j;
    if (masked_out(ph, sa, ipix, lane == 0))
      continue; // uniform per wavefront
    const RayGeo r = slots[wave][j];
    double dm = 0., I = 0., Q = 0., U = 0., fd_carry = 0.;
    for (uint64_t base = r.k_lo; base < r.k_hi; base += 64) {
      const uint64_t k = base + lane;
      FastSample o;
      bool valid = false;
      if (k < r.k_hi)
        valid = eval_fast<GRIDS, PW1>(ph, fc, gr, sa, r, k, o);
      const double fdk = valid ? o.fdk : 0.;
      const double incl = wave_incl_scan(fdk, lane);
      const double fd_here = fd_carry + incl;
      fd_carry += wave_last(incl);
      if (valid) {
        double sph, cph;
        sincos_fast(two_l2 * (r.fd_inner + fd_here), sph, cph);
        I += o.jt;
        Q += (cph * o.c2i - sph * o.s2i) * o.jp;
        U += (sph * o.c2i + cph * o.s2i) * o.jp;
        dm += o.te;
      }
    }
    dm = wave_sum(dm);
    dm = sa.step ? dm * sa.delta : 0.; // a shell thinner than one step: no finite delta
    I = wave_sum(I);
    Q = wave_sum(Q);
    U = wave_sum(U);
    if (lane == 0)
      store_ray(ph, sa, ipix, dm, fd_carry, I, Q, U);
  }
}

template <bool PACKED, bool CREG, bool HEAVY>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(CREG ? 3 : 2))) void
k_los_thread(Phys ph, Grids gr, ShellArgs sa) {
  __shared__ double stab[CREG ? kSyncTab : 1];
  if (CREG) {
    load_sync_tab(stab);
    __syncthreads();
  }
  const uint64_t ipix = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (ipix >= sa.hs.npix)
    return;
  if (masked_out(ph, sa, ipix, true))
    return;
  RayGeo r;
  ray_setup(ph, sa, ipix, r);
  double dm = 0., fd = 0., I = 0., Q = 0., U = 0.;
  for (uint64_t k = r.k_lo; k < r.k_hi; ++k) {
    SampleOut s;
    if (!eval_sample<PACKED, CREG, HEAVY>(ph, gr, sa, r, k, s, stab))
      continue;
    if (ph.do_dm)
      dm += s.te * sa.delta;
    if (ph.do_fd || ph.do_sync)
      fd += s.fdc;
    if (ph.do_sync) {
      I += s.jt;
      const double qui = (r.fd_inner + fd) * ph.lambda_sq + s.ipa;
      Q += cos(2. * qui) * s.jp;
      U += sin(2. * qui) * s.jp;
    }
  }
  store_ray(ph, sa, ipix, dm, fd, I, Q, U);
}

// ------------------------------------------------------------- accumulate --

struct AccPairs {
  double *dst[5];
  const double *src[5];
  int n;
};

// dst += src (same nside) or dst += interpolate(src, pixel centre of dst)
// for up to 5 map pairs sharing the two resolutions (hamdis.h:698-721).
__global__ __launch_bounds__(256) void k_accumulate(AccPairs p, Healpix hd,
                                                    Healpix hsrc) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= hd.npix)
    return;
  if (hd.npix == hsrc.npix) {
    for (int m = 0; m < p.n; ++m)
      p.dst[m][i] = p.dst[m][i] + p.src[m][i];
    return;
  }
  double theta, phi;
  hp_pix2ang(hd, i, theta, phi);
  uint64_t pix[4];
  double wght[4];
  hp_interpol_weights(hsrc, theta, phi, pix, wght);
  for (int m = 0; m < p.n; ++m) {
    double wtot = 0.0, res = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double val = p.src[m][pix[q]];
      if (val > cgs::undef_threshold) {
        res += val * wght[q];
        wtot += wght[q];
      }
    }
    const double v = (wtot == 0.0) ? cgs::undef : res / wtot;
    p.dst[m][i] = p.dst[m][i] + v;
  }
}

// The mask at a shell's nside from the caller's map (Hampisk::duplicate).
__global__ __launch_bounds__(256) void k_mask_dup(double *out, Healpix ht, const double *piv,
                                                  Healpix hp) {
  const uint64_t m = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (m < ht.npix)
    out[m] = hp_mask_value(ht, m, piv, hp);
}

int mask_at_nside(hamx_ctx *ctx, double *out, const Healpix &ht, const double *piv,
                  const Healpix &hp) {
  hipLaunchKernelGGL(k_mask_dup, dim3(static_cast<unsigned>((ht.npix + 255) / 256)), dim3(256),
                     0, ctx->stream, out, ht, piv, hp);
  HAMX_HIP(ctx, hipGetLastError());
  return HAMX_OK;
}

// Pack SoA brnd (x,y,z) and ternd into one double4 record per vertex.
__global__ void k_pack4(double4 *__restrict__ out, const double *__restrict__ bx,
                        const double *__restrict__ by, const double *__restrict__ bz,
                        const double *__restrict__ te, uint64_t n) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < n; i += stride) {
    double4 v;
    v.x = bx ? bx[i] : 0.;
    v.y = by ? by[i] : 0.;
    v.z = bz ? bz[i] : 0.;
    v.w = te ? te[i] : 0.;
    out[i] = v;
  }
}

// ------------------------------------------------------------------ host --

bool same_box(const hamx_box &a, const hamx_box &b) {
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.x_min == b.x_min &&
         a.x_max == b.x_max && a.y_min == b.y_min && a.y_max == b.y_max &&
         a.z_min == b.z_min && a.z_max == b.z_max;
}

int check_box(hamx_ctx *ctx, const hamx_box &b, const char *what) {
  if (b.nx < 2 || b.ny < 2 || b.nz < 2)
    return set_err(ctx, HAMX_ERR_ARG, "%s grid needs at least 2 vertices per axis", what);
  if (!(b.x_max > b.x_min) || !(b.y_max > b.y_min) || !(b.z_max > b.z_min))
    return set_err(ctx, HAMX_ERR_ARG, "%s grid box limits are not increasing", what);
  // cell coordinates and the y-z plane index are 32-bit on the device
  if (b.nx > 0xffffffffull || b.ny > 0xffffffffull || b.nz > 0xffffffffull ||
      b.ny * b.nz > 0xffffffffull)
    return set_err(ctx, HAMX_ERR_ARG, "%s grid has 2^32 or more vertices per y-z plane", what);
  return HAMX_OK;
}

// Integrator::assemble_shell_ref (integrator.cc:231-244)
void shell_ref(const hamx_los_desc *d, uint32_t s, uint64_t &step, double &d_start,
               double &delta) {
  d_start = d->radii_shell[s];
  const double d_stop = d->radii_shell[s + 1];
  delta = d->oc_r_res;
  const double stepf = std::floor(d_stop / delta - d_start / delta);
  step = stepf > 0 ? static_cast<uint64_t>(stepf) : 0;
  delta = (d_stop - d_start) / static_cast<double>(step);
}

int validate_los(hamx_ctx *ctx, const hamx_los_desc *d) {
  if (d->n_shell < 1 || d->n_shell > HAMX_MAX_SHELL)
    return set_err(ctx, HAMX_ERR_ARG, "n_shell %u outside [1, %d]", d->n_shell,
                   HAMX_MAX_SHELL);
  for (uint32_t s = 0; s < d->n_shell; ++s)
    if (!is_pow2(d->nside_shell[s]))
      return set_err(ctx, HAMX_ERR_NSIDE, "shell %u nside %u is not a power of 2",
                     s, d->nside_shell[s]);
  if (!(d->oc_r_res > 0.))
    return set_err(ctx, HAMX_ERR_ARG, "oc_r_res must be positive");
  for (uint32_t s = 0; s < d->n_shell; ++s) {
    // the marches convert step indices with 32-bit conversions
    uint64_t step;
    double d_start, delta;
    shell_ref(d, s, step, d_start, delta);
    if (step >= (1ull << 31))
      return set_err(ctx, HAMX_ERR_ARG, "shell %u has 2^31 or more radial steps", s);
  }
  if (d->do_dm && !is_pow2(d->nside_dm))
    return set_err(ctx, HAMX_ERR_NSIDE, "nside_dm %u is not a power of 2", d->nside_dm);
  if (d->do_fd && !is_pow2(d->nside_fd))
    return set_err(ctx, HAMX_ERR_NSIDE, "nside_fd %u is not a power of 2", d->nside_fd);
  if (d->do_sync && !is_pow2(d->nside_sync))
    return set_err(ctx, HAMX_ERR_NSIDE, "nside_sync %u is not a power of 2",
                   d->nside_sync);
  if (d->do_sync && !(d->sync_freq > 0.))
    return set_err(ctx, HAMX_ERR_ARG, "sync frequency must be positive");
  if (d->breg_model < HAMX_BREG_NONE || d->breg_model > HAMX_BREG_JAFFE)
    return set_err(ctx, HAMX_ERR_UNSUPPORTED, "unsupported breg model %d", d->breg_model);
  if (d->tereg_model < HAMX_TEREG_NONE || d->tereg_model > HAMX_TEREG_YMW16)
    return set_err(ctx, HAMX_ERR_UNSUPPORTED, "unsupported tereg model %d", d->tereg_model);
  if (d->cre_model < HAMX_CRE_NONE || d->cre_model > HAMX_CRE_GRID)
    return set_err(ctx, HAMX_ERR_UNSUPPORTED, "unsupported cre model %d", d->cre_model);
  if (d->do_sync && d->cre_model == HAMX_CRE_NONE)
    return set_err(ctx, HAMX_ERR_ARG, "synchrotron output requires a CRE model");
  if (d->schedule < HAMX_SCHED_AUTO || d->schedule > HAMX_SCHED_WAVE)
    return set_err(ctx, HAMX_ERR_ARG, "unknown schedule %d", d->schedule);
  return HAMX_OK;
}

Phys make_phys(const hamx_los_desc *d) {
  Phys ph;
  std::memset(&ph, 0, sizeof(ph));
  ph.obs = Vec3{d->observer[0], d->observer[1], d->observer[2]};
  ph.gc_r_min = d->gc_r_min;
  ph.gc_r_max = d->gc_r_max;
  ph.gc_r_min2 = d->gc_r_min * d->gc_r_min;
  ph.gc_r_max2 = d->gc_r_max * d->gc_r_max;
  ph.gc_z_min = d->gc_z_min;
  ph.gc_z_max = d->gc_z_max;
  ph.do_dm = d->do_dm != 0;
  ph.do_fd = d->do_fd != 0;
  ph.do_sync = d->do_sync != 0;
  ph.breg_model = d->breg_model;
  ph.tereg_model = d->tereg_model;
  ph.cre_model = d->cre_model;
  std::memcpy(ph.breg, d->breg, sizeof(ph.breg));
  std::memcpy(ph.tereg, d->tereg, sizeof(ph.tereg));
  std::memcpy(ph.cre, d->cre, sizeof(ph.cre));
  ph.fd_fore = -(cgs::qe * cgs::qe * cgs::qe) / (2. * cgs::pi * cgs::mec2 * cgs::mec2);
  ph.qe3 = cgs::qe * cgs::qe * cgs::qe;
  ph.three_qe = 3. * cgs::qe;
  ph.fourpi_mec2 = 4. * cgs::pi * cgs::mec2;
  ph.sixteenpi_mec2 = 16. * cgs::pi * cgs::mec2;
  if (d->do_sync) {
    const double f = d->sync_freq;
    ph.lambda_sq = (cgs::c_light / f) * (cgs::c_light / f);
    ph.i2bt = cgs::c_light * cgs::c_light / (2. * cgs::kB * f * f);
    ph.A_num = 2. * cgs::pi * f * cgs::mec;
    ph.xft_num = 4. * cgs::pi * f * cgs::mec * cgs::mec2 * cgs::mec2;
    ph.xfp_num = 2. * cgs::mec * cgs::mec2 * cgs::mec2 * 2. * cgs::pi * f;
    ph.fore_num = 1.73205081 * cgs::qe * cgs::qe * cgs::qe;
    ph.fore_den = cgs::mec2 * cgs::c_light * cgs::GeV * cgs::m * cgs::m * cgs::sec;
    const double p = cgs::h_planck * f / (cgs::kB * 2.725);
    ph.tc_em1 = std::exp(p) - 1.;
    ph.tc_den = p * p * std::exp(p);
  }
  // CRE normalisation constants (cre_ana.cc:30-47, cre_unif.cc)
  if (d->cre_model == HAMX_CRE_ANA) {
    const double alpha = d->cre[0], beta = d->cre[1], theta = d->cre[2];
    const double E0 = d->cre[5], j0 = d->cre[6];
    ph.R0 = std::sqrt(d->observer[0] * d->observer[0] + d->observer[1] * d->observer[1]);
    const double gamma0 = E0 / cgs::mec2;
    const double beta0 = std::sqrt(1. - 1. / gamma0);
    const double unit = 4. * cgs::pi * cgs::mec /
                        (cgs::GeV * cgs::m * cgs::m * cgs::sec * beta0);
    const double r_obs = ph.R0 / cgs::kpc;
    const double z_obs = std::fabs(d->observer[2]) / cgs::kpc;
    const double idx_obs = -alpha + beta * r_obs + theta * z_obs;
    const double norm = j0 * std::pow(gamma0, -idx_obs);
    ph.cre_nu = norm * unit;
    ph.idx_const = (beta == 0. && theta == 0.) ? 1 : 0;
    ph.idx = -alpha;
  } else if (d->cre_model == HAMX_CRE_UNIF) {
    const double alpha = d->cre[0], E0 = d->cre[2], j0 = d->cre[3];
    const double gamma0 = E0 / cgs::mec2;
    const double beta0 = std::sqrt(1. - 1. / gamma0);
    const double unit = 4. * cgs::pi * cgs::mec /
                        (cgs::GeV * cgs::m * cgs::m * cgs::sec * beta0);
    const double norm = j0 * std::pow(gamma0, -(-alpha));
    ph.cre_nu = norm * unit;
    ph.idx_const = 1;
    ph.idx = -alpha;
  }
  if (ph.idx_const) {
    ph.g_t1 = std::tgamma(-0.25 * ph.idx + 19. / 12.);
    ph.g_t2 = std::tgamma(-0.25 * ph.idx - 1. / 12.);
    ph.g_p1 = std::tgamma(-0.25 * ph.idx + 7. / 12.);
  }
  return ph;
}

// Per-shell constants of the specialised march: the emissivity factors of
// integrator.cc:333-346,402-415 with every sample-independent factor folded.
FastConst make_fast(const hamx_los_desc *d, const Phys &ph, double delta, bool pw1) {
  FastConst fc;
  std::memset(&fc, 0, sizeof(fc));
  fc.b0 = d->breg[0];
  fc.psi0 = d->breg[1];
  fc.psi1 = d->breg[2];
  fc.chi0 = d->breg[3];
  // |ln(r / 8 kpc)| <= ln(8 / 3) over the LSA window 3 <= r <= 20 kpc
  fc.lsa_direct = (std::fabs(fc.psi0) + std::fabs(fc.psi1) * 0.99 <= 0.25 * cgs::pi &&
                   std::fabs(fc.chi0) <= 0.25 * cgs::pi)
                      ? 1
                      : 0;
  fc.inv_8kpc = 1.0 / (8. * cgs::kpc);
  fc.inv_kpc = 1.0 / (1. * cgs::kpc);
  fc.n0 = d->tereg[0];
  fc.te_r0 = d->tereg[1];
  fc.te_r02 = d->tereg[1] * d->tereg[1];
  fc.cre_unif = d->cre_model == HAMX_CRE_UNIF;
  if (fc.cre_unif) {
    fc.cre_r0 = d->cre[1];
    fc.cre_r02 = d->cre[1] * d->cre[1];
  } else {
    fc.R0 = ph.R0;
    fc.inv_cre_r0 = 1.0 / d->cre[3];
    fc.inv_cre_z0 = 1.0 / d->cre[4];
    fc.cre_z_kpc = fc.inv_cre_z0 == fc.inv_kpc ? 1 : 0;
  }
  fc.pw_e = 0.5 * (ph.idx + 1);
  fc.a_coef = ph.A_num / ph.three_qe;
  const double common = ph.cre_nu * 1.73205081 * ph.qe3 * delta * ph.i2bt;
  fc.kt = common * ph.g_t1 * ph.g_t2 / (ph.fourpi_mec2 * (1. - ph.idx));
  fc.kp = common * ph.g_p1 * ph.g_t2 / ph.sixteenpi_mec2;
  if (pw1) { // A^-1 = |B_perp| / a_coef
    fc.kt /= fc.a_coef;
    fc.kp /= fc.a_coef;
  }
  fc.fd_scale = ph.fd_fore * delta;
  return fc;
}

int scratch_get(hamx_ctx *ctx, Arena *ar, int slot, size_t bytes, void **out) {
  if (!ar)
    return ws_get(ctx, slot, bytes, out);
  const size_t off = (ar->used + 255) & ~static_cast<size_t>(255);
  if (off > ar->cap || bytes > ar->cap - off)
    return set_err(ctx, HAMX_ERR_ARG,
                   "the pass needs more than the %zu bytes of static scratch "
                   "(%zu in use, %zu more requested): shells above nside %d need "
                   "hamx_los_forward", ar->cap, off, bytes, HAMX_STATIC_MAX_NSIDE);
  *out = ar->base + off;
  ar->used = off + bytes;
  return HAMX_OK;
}

int stage_fields(hamx_ctx *ctx, const hamx_los_desc *d, const hamx_fields &f,
                 bool in_dev, Grids &gr, bool &packed, bool &grids_any, Arena *ar) {
  std::memset(&gr, 0, sizeof(gr));
  const bool has_brnd = f.brnd_x || f.brnd_y || f.brnd_z;
  if (has_brnd && !(f.brnd_x && f.brnd_y && f.brnd_z))
    return set_err(ctx, HAMX_ERR_ARG, "brnd grid needs all three components");
  if (has_brnd)
    HAMX_TRY(check_box(ctx, f.brnd_box, "brnd"));
  if (f.ternd)
    HAMX_TRY(check_box(ctx, f.ternd_box, "ternd"));
  if (d->breg_model == HAMX_BREG_GRID) {
    if (!(f.breg_x && f.breg_y && f.breg_z))
      return set_err(ctx, HAMX_ERR_ARG, "breg grid model without breg grid arrays");
    HAMX_TRY(check_box(ctx, f.breg_box, "breg"));
  }
  if (d->tereg_model == HAMX_TEREG_GRID) {
    if (!f.tereg)
      return set_err(ctx, HAMX_ERR_ARG, "tereg grid model without tereg grid array");
    HAMX_TRY(check_box(ctx, f.tereg_box, "tereg"));
  }
  const void *p;
  if (has_brnd) {
    const size_t nb = box_size(f.brnd_box) * sizeof(double);
    HAMX_TRY(stage_in(ctx, kSlotBrndX, f.brnd_x, nb, in_dev, &p));
    gr.brnd_x = static_cast<const double *>(p);
    HAMX_TRY(stage_in(ctx, kSlotBrndY, f.brnd_y, nb, in_dev, &p));
    gr.brnd_y = static_cast<const double *>(p);
    HAMX_TRY(stage_in(ctx, kSlotBrndZ, f.brnd_z, nb, in_dev, &p));
    gr.brnd_z = static_cast<const double *>(p);
    gr.g_brnd = grid_geo(f.brnd_box);
  }
  if (f.ternd) {
    HAMX_TRY(stage_in(ctx, kSlotTernd, f.ternd, box_size(f.ternd_box) * sizeof(double),
                      in_dev, &p));
    gr.ternd = static_cast<const double *>(p);
    gr.g_ternd = grid_geo(f.ternd_box);
  }
  if (d->breg_model == HAMX_BREG_GRID) {
    const size_t nb = box_size(f.breg_box) * sizeof(double);
    HAMX_TRY(stage_in(ctx, kSlotBregX, f.breg_x, nb, in_dev, &p));
    gr.breg_x = static_cast<const double *>(p);
    HAMX_TRY(stage_in(ctx, kSlotBregY, f.breg_y, nb, in_dev, &p));
    gr.breg_y = static_cast<const double *>(p);
    HAMX_TRY(stage_in(ctx, kSlotBregZ, f.breg_z, nb, in_dev, &p));
    gr.breg_z = static_cast<const double *>(p);
    gr.g_breg = grid_geo(f.breg_box);
  }
  if (d->tereg_model == HAMX_TEREG_GRID) {
    HAMX_TRY(stage_in(ctx, kSlotTereg, f.tereg, box_size(f.tereg_box) * sizeof(double),
                      in_dev, &p));
    gr.tereg = static_cast<const double *>(p);
    gr.g_tereg = grid_geo(f.tereg_box);
  }
  if (d->cre_model == HAMX_CRE_GRID) {
    if (!f.cre)
      return set_err(ctx, HAMX_ERR_ARG, "cre grid model without cre grid array");
    HAMX_TRY(check_box(ctx, f.cre_box, "cre"));
    if (f.cre_nE < 2)
      return set_err(ctx, HAMX_ERR_ARG, "cre grid needs at least 2 energies");
    if (!(f.cre_E_min > cgs::mec2) || !(f.cre_E_fact > 0.))
      return set_err(ctx, HAMX_ERR_ARG,
                     "cre energy axis must start above mec2 and increase");
    const uint64_t nE = f.cre_nE;
    HAMX_TRY(stage_in(ctx, kSlotCre, f.cre, box_size(f.cre_box) * nE * sizeof(double), in_dev,
                      &p));
    gr.cre = static_cast<const double *>(p);
    gr.g_cre = grid_geo(f.cre_box);
    gr.cre_nE = nE;
    // E_i, x denominators and beta as integrator.cc:300-304 forms them
    std::vector<double> &tab = ctx->host_tab;
    tab.assign(4 * nE, 0.);
    std::vector<double> KE(nE);
    for (uint64_t i = 0; i != nE; ++i) {
      KE[i] = f.cre_E_min * std::exp(i * f.cre_E_fact);
      tab[i] = KE[i] * KE[i];
      tab[nE + i] = std::sqrt(1 - cgs::mec2 / KE[i]);
    }
    for (uint64_t i = 0; i + 1 < nE; ++i) {
      tab[2 * nE + i] = std::fabs(KE[i + 1] - KE[i]);
      tab[3 * nE + i] = std::log(0.5 * (1. / tab[i + 1] + 1. / tab[i]));
    }
    if (ar)
      return set_err(ctx, HAMX_ERR_UNSUPPORTED,
                     "a CRE grid needs host-built energy tables: use hamx_los_forward");
    void *tp;
    HAMX_TRY(ws_get(ctx, kSlotCreTab, tab.size() * sizeof(double), &tp));
    HAMX_HIP(ctx, hipMemcpyAsync(tp, tab.data(), tab.size() * sizeof(double),
                                 hipMemcpyHostToDevice, ctx->stream));
    gr.cre_tab = static_cast<const double *>(tp);
  }
  // Pack the random B components and TE density into one record per vertex
  // when they share a box: 64 B contiguous per z-pair of corners.
  packed = false;
  if (!ar && d->cre_model != HAMX_CRE_GRID && has_brnd &&
      (!f.ternd || same_box(f.brnd_box, f.ternd_box))) {
    const uint64_t n = box_size(f.brnd_box);
    void *pk;
    HAMX_TRY(ws_get(ctx, kSlotPacked, n * sizeof(double4), &pk));
    uint64_t blocks = (n + 255) / 256;
    if (blocks > 16384)
      blocks = 16384;
    hipLaunchKernelGGL(k_pack4, dim3((unsigned)blocks), dim3(256), 0, ctx->stream,
                       static_cast<double4 *>(pk), gr.brnd_x, gr.brnd_y, gr.brnd_z,
                       gr.ternd, n);
    HAMX_HIP(ctx, hipGetLastError());
    gr.packed = static_cast<const double4 *>(pk);
    gr.g_packed = gr.g_brnd;
    packed = true;
  }
  grids_any = has_brnd || f.ternd;
  return HAMX_OK;
}

} // namespace hamx

using namespace hamx;

extern "C" int hamx_los_samples(const hamx_los_desc *d, uint64_t *samples) {
  if (!d || !samples)
    return HAMX_ERR_ARG;
  if (d->n_shell < 1 || d->n_shell > HAMX_MAX_SHELL || !(d->oc_r_res > 0.))
    return HAMX_ERR_ARG;
  uint64_t total = 0;
  for (uint32_t s = 0; s < d->n_shell; ++s) {
    uint64_t step;
    double d0, delta;
    shell_ref(d, s, step, d0, delta);
    const uint64_t ns = d->nside_shell[s];
    total += 12 * ns * ns * step;
  }
  *samples = total;
  return HAMX_OK;
}

namespace hamx {
namespace {
// A NaN or an infinity in an input grid or parameter reaches the shell maps
// (the reference's debug build asserts isfinite on B_par, B_per, te and the
// phase, integrator.cc:192,195,202,223; its release build carries the NaN):
// flag[0] becomes 1 when any of the given values is not finite.
struct FiniteArgs {
  const double *m[5];
  uint64_t n[5];
  int k;
};

__global__ __launch_bounds__(256) void k_nonfinite(FiniteArgs a, double *flag) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  bool bad = false;
  for (int j = 0; j < a.k; ++j)
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < a.n[j]; i += stride)
      bad |= !isfinite(a.m[j][i]);
  if (bad)
    flag[0] = 1.;
}
} // namespace

int los_forward_impl(hamx_ctx *ctx, const hamx_los_desc *d, const hamx_fields *fields,
                     hamx_maps *out, uint32_t flags, double *ckpt, double *shell_fd,
                     Arena *ar) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!d || !out)
    return set_err(ctx, HAMX_ERR_ARG, "null descriptor or output maps");
  HAMX_TRY(validate_los(ctx, d));
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  const bool in_dev = (flags & HAMX_IN_DEVICE) != 0;
  const bool out_dev = (flags & HAMX_OUT_DEVICE) != 0;
  hamx_fields none;
  std::memset(&none, 0, sizeof(none));
  const hamx_fields &f = fields ? *fields : none;

  // ---- field grids
  Grids gr;
  bool packed = false, grids_any = false;
  HAMX_TRY(stage_fields(ctx, d, f, in_dev, gr, packed, grids_any, ar));

  // ---- mask, duplicated to every shell's nside (integrator.cc:36-39)
  const double *mask_piv = nullptr;
  Healpix h_mask;
  std::memset(&h_mask, 0, sizeof(h_mask));
  if (f.mask) {
    if (!is_pow2(f.nside_mask))
      return set_err(ctx, HAMX_ERR_NSIDE, "nside_mask %u is not a power of 2", f.nside_mask);
    h_mask = healpix_make(f.nside_mask);
    const void *mp;
    HAMX_TRY(stage_in(ctx, kSlotMaskPivot, f.mask, h_mask.npix * sizeof(double), in_dev, &mp));
    mask_piv = static_cast<const double *>(mp);
  }

  // ---- output maps (device)
  const Phys ph = make_phys(d);
  const uint64_t nside_fd_map = d->do_fd ? d->nside_fd : d->nside_sync;
  const Healpix h_dm = healpix_make(d->do_dm ? d->nside_dm : 1);
  const Healpix h_sync = healpix_make(d->do_sync ? d->nside_sync : 1);
  const Healpix h_fd = healpix_make((d->do_fd || d->do_sync) ? nside_fd_map : 1);
  double *o_dm = nullptr, *o_fd = nullptr, *o_is = nullptr, *o_qs = nullptr,
         *o_us = nullptr;
  void *q;
  auto out_map = [&](int slot, double *user, const Healpix &h, double **dst) -> int {
    if (out_dev && user) {
      *dst = user;
    } else {
      HAMX_TRY(scratch_get(ctx, ar, slot, h.npix * sizeof(double), &q));
      *dst = static_cast<double *>(q);
    }
    return HAMX_OK;
  };
  if (d->do_dm)
    HAMX_TRY(out_map(kSlotOutDm, out->dm, h_dm, &o_dm));
  if (d->do_sync) {
    HAMX_TRY(out_map(kSlotOutIs, out->is, h_sync, &o_is));
    HAMX_TRY(out_map(kSlotOutQs, out->qs, h_sync, &o_qs));
    HAMX_TRY(out_map(kSlotOutUs, out->us, h_sync, &o_us));
  }
  if (d->do_fd || d->do_sync)
    HAMX_TRY(out_map(kSlotOutFd, out->fd, h_fd, &o_fd));

  // ---- shell maps
  uint64_t max_npix = 0;
  for (uint32_t s = 0; s < d->n_shell; ++s) {
    const uint64_t ns = d->nside_shell[s];
    if (12 * ns * ns > max_npix)
      max_npix = 12 * ns * ns;
  }
  double *t_dm = nullptr, *t_fd = nullptr, *t_is = nullptr, *t_qs = nullptr,
         *t_us = nullptr;
  const size_t tb = max_npix * sizeof(double);
  if (d->do_dm) {
    HAMX_TRY(scratch_get(ctx, ar, kSlotTmpDm, tb, &q));
    t_dm = static_cast<double *>(q);
  }
  if ((d->do_fd || d->do_sync) && !shell_fd) {
    HAMX_TRY(scratch_get(ctx, ar, kSlotTmpFd, tb, &q));
    t_fd = static_cast<double *>(q);
  }
  if (d->do_sync) {
    HAMX_TRY(scratch_get(ctx, ar, kSlotTmpIs, tb, &q));
    t_is = static_cast<double *>(q);
    HAMX_TRY(scratch_get(ctx, ar, kSlotTmpQs, tb, &q));
    t_qs = static_cast<double *>(q);
    HAMX_TRY(scratch_get(ctx, ar, kSlotTmpUs, tb, &q));
    t_us = static_cast<double *>(q);
  }
  // the mask at each shell's nside (one buffer, sized for the largest shell)
  double *mask_shell = nullptr;
  if (mask_piv) {
    HAMX_TRY(scratch_get(ctx, ar, kSlotMaskShell, tb, &q));
    mask_shell = static_cast<double *>(q);
  }
  HAMX_TRY(scratch_get(ctx, ar, kSlotFlag, sizeof(double), &q));
  double *nonfinite = static_cast<double *>(q);
  HAMX_TRY(fill_f64(ctx, nonfinite, 1, 0.));
  // every buffer is in place: zero the output maps
  if (o_dm)
    HAMX_TRY(fill_f64(ctx, o_dm, h_dm.npix, 0.));
  if (o_is) {
    HAMX_TRY(fill_f64(ctx, o_is, h_sync.npix, 0.));
    HAMX_TRY(fill_f64(ctx, o_qs, h_sync.npix, 0.));
    HAMX_TRY(fill_f64(ctx, o_us, h_sync.npix, 0.));
  }
  if (o_fd)
    HAMX_TRY(fill_f64(ctx, o_fd, h_fd.npix, 0.));

  const int sched = d->schedule == HAMX_SCHED_THREAD ? HAMX_SCHED_THREAD : HAMX_SCHED_WAVE;
  // the specialised kernel covers LSA + uniform TE + constant-index CRE with
  // synchrotron on and the random grids (if any) packed in one box
  const bool fast = d->schedule == HAMX_SCHED_AUTO && d->breg_model == HAMX_BREG_LSA &&
                    d->tereg_model == HAMX_TEREG_UNIF && ph.idx_const && d->do_sync &&
                    (d->cre_model == HAMX_CRE_ANA || d->cre_model == HAMX_CRE_UNIF) &&
                    (packed || !grids_any);
  const bool pw1 = fast && 0.5 * (ph.idx + 1) == -1.0;
  const bool creg = d->cre_model == HAMX_CRE_GRID && d->do_sync;
  // models with large evaluators get kernels of their own (register use)
  const bool heavy = d->breg_model == HAMX_BREG_JF12 || d->breg_model == HAMX_BREG_JAFFE ||
                     d->tereg_model == HAMX_TEREG_YMW16;
  for (uint32_t s = 0; s < d->n_shell; ++s) {
    ShellArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.hs = healpix_make(d->nside_shell[s]);
    shell_ref(d, s, sa.step, sa.d_start, sa.delta);
    // FD cache: only a carried FD map can be nonzero (Appendix of the
    // reference pipeline: passes without do_fd read a zero map).
    sa.fd_map = (d->do_fd && s > 0) ? o_fd : nullptr;
    sa.hfd = h_fd;
    sa.t_dm = t_dm;
    // with a tape, each shell's FD map is written into the tape itself
    sa.t_fd = shell_fd ? shell_fd + shell_offset(d, s) : t_fd;
    sa.t_is = t_is;
    sa.t_qs = t_qs;
    sa.t_us = t_us;
    const uint64_t npix = sa.hs.npix;
    if (mask_piv) {
      hipLaunchKernelGGL(k_mask_dup, dim3(static_cast<unsigned>((npix + 255) / 256)), dim3(256),
                         0, ctx->stream, mask_shell, sa.hs, mask_piv, h_mask);
      HAMX_HIP(ctx, hipGetLastError());
      sa.mask = mask_shell;
    }
    if (ckpt && d->do_fd) // FD map before shell s, for the adjoint
      HAMX_HIP(ctx, hipMemcpyAsync(ckpt + s * h_fd.npix, o_fd, h_fd.npix * sizeof(double),
                                   hipMemcpyDeviceToDevice, ctx->stream));
    ProfMark pm = prof_begin(ctx);
    if (fast) {
      const FastConst fc = make_fast(d, ph, sa.delta, pw1);
      const unsigned blocks =
          static_cast<unsigned>((npix + 4 * kFastRays - 1) / (4 * kFastRays));
      if (packed && pw1)
        hipLaunchKernelGGL((k_los_fast<true, true>), dim3(blocks), dim3(256), 0, ctx->stream,
                           ph, fc, gr, sa);
      else if (packed)
        hipLaunchKernelGGL((k_los_fast<true, false>), dim3(blocks), dim3(256), 0,
                           ctx->stream, ph, fc, gr, sa);
      else if (pw1)
        hipLaunchKernelGGL((k_los_fast<false, true>), dim3(blocks), dim3(256), 0,
                           ctx->stream, ph, fc, gr, sa);
      else
        hipLaunchKernelGGL((k_los_fast<false, false>), dim3(blocks), dim3(256), 0,
                           ctx->stream, ph, fc, gr, sa);
    } else if (sched == HAMX_SCHED_WAVE) {
      const unsigned blocks = static_cast<unsigned>((npix + 3) / 4);
#define HAMX_LAUNCH_WAVE(P, C, H)                                                          \
  hipLaunchKernelGGL((k_los_wave<P, C, H>), dim3(blocks), dim3(256), 0, ctx->stream, ph, gr, sa)
      if (creg)
        {
        if (heavy)
          HAMX_LAUNCH_WAVE(false, true, true);
        else
          HAMX_LAUNCH_WAVE(false, true, false);
      }
      else if (packed)
        {
        if (heavy)
          HAMX_LAUNCH_WAVE(true, false, true);
        else
          HAMX_LAUNCH_WAVE(true, false, false);
      }
      else
        {
        if (heavy)
          HAMX_LAUNCH_WAVE(false, false, true);
        else
          HAMX_LAUNCH_WAVE(false, false, false);
      }
#undef HAMX_LAUNCH_WAVE
    } else {
      const unsigned blocks = static_cast<unsigned>((npix + 255) / 256);
#define HAMX_LAUNCH_THREAD(P, C, H)                                                        \
  hipLaunchKernelGGL((k_los_thread<P, C, H>), dim3(blocks), dim3(256), 0, ctx->stream, ph, gr, \
                     sa)
      if (creg)
        {
        if (heavy)
          HAMX_LAUNCH_THREAD(false, true, true);
        else
          HAMX_LAUNCH_THREAD(false, true, false);
      }
      else if (packed)
        {
        if (heavy)
          HAMX_LAUNCH_THREAD(true, false, true);
        else
          HAMX_LAUNCH_THREAD(true, false, false);
      }
      else
        {
        if (heavy)
          HAMX_LAUNCH_THREAD(false, false, true);
        else
          HAMX_LAUNCH_THREAD(false, false, false);
      }
#undef HAMX_LAUNCH_THREAD
    }
    HAMX_HIP(ctx, hipGetLastError());
    prof_end(ctx, pm,
             fast ? "los_fast" : (sched == HAMX_SCHED_WAVE ? "los_wave" : "los_thread"), s,
             npix * sa.step);
    {
      // every ray's observables finite (checked here: the HEALPix
      // interpolation of the accumulation skips a NaN as an undefined pixel)
      FiniteArgs fa;
      std::memset(&fa, 0, sizeof(fa));
      for (const double *m : {static_cast<const double *>(sa.t_dm),
                              static_cast<const double *>(sa.t_fd),
                              static_cast<const double *>(sa.t_is),
                              static_cast<const double *>(sa.t_qs),
                              static_cast<const double *>(sa.t_us)})
        if (m) {
          fa.m[fa.k] = m;
          fa.n[fa.k++] = npix;
        }
      const unsigned fb = static_cast<unsigned>(std::min<uint64_t>((npix + 255) / 256, 1024));
      hipLaunchKernelGGL(k_nonfinite, dim3(fb), dim3(256), 0, ctx->stream, fa, nonfinite);
      HAMX_HIP(ctx, hipGetLastError());
    }
    // accumulate shell maps into the outputs (integrator.cc:122-132)
    auto acc = [&](AccPairs &pairs, const Healpix &hd) -> int {
      if (pairs.n == 0)
        return HAMX_OK;
      const unsigned blocks = static_cast<unsigned>((hd.npix + 255) / 256);
      ProfMark am = prof_begin(ctx);
      hipLaunchKernelGGL(k_accumulate, dim3(blocks), dim3(256), 0, ctx->stream,
                         pairs, hd, sa.hs);
      HAMX_HIP(ctx, hipGetLastError());
      prof_end(ctx, am, "accumulate", s, hd.npix);
      return HAMX_OK;
    };
    AccPairs pr;
    std::memset(&pr, 0, sizeof(pr));
    if (d->do_sync) {
      pr.dst[0] = o_is; pr.src[0] = t_is;
      pr.dst[1] = o_qs; pr.src[1] = t_qs;
      pr.dst[2] = o_us; pr.src[2] = t_us;
      pr.n = 3;
      if (d->do_dm && d->nside_dm == d->nside_sync) {
        pr.dst[pr.n] = o_dm; pr.src[pr.n] = t_dm; ++pr.n;
      }
      if (d->do_fd && d->nside_fd == d->nside_sync) {
        pr.dst[pr.n] = o_fd; pr.src[pr.n] = sa.t_fd; ++pr.n;
      }
      HAMX_TRY(acc(pr, h_sync));
    }
    if (d->do_dm && !(d->do_sync && d->nside_dm == d->nside_sync)) {
      AccPairs p2;
      std::memset(&p2, 0, sizeof(p2));
      p2.dst[0] = o_dm; p2.src[0] = t_dm; p2.n = 1;
      if (d->do_fd && d->nside_fd == d->nside_dm &&
          !(d->do_sync && d->nside_fd == d->nside_sync)) {
        p2.dst[1] = o_fd; p2.src[1] = sa.t_fd; p2.n = 2;
      }
      HAMX_TRY(acc(p2, h_dm));
    }
    const bool fd_done = (d->do_sync && d->nside_fd == d->nside_sync) ||
                         (d->do_dm && d->nside_fd == d->nside_dm);
    if (d->do_fd && !fd_done) {
      AccPairs p3;
      std::memset(&p3, 0, sizeof(p3));
      p3.dst[0] = o_fd; p3.src[0] = sa.t_fd; p3.n = 1;
      HAMX_TRY(acc(p3, h_fd));
    }
  }

  // ---- the shells' non-finite flag
  double host_flag = 0.;
  HAMX_HIP(ctx, hipMemcpyAsync(&host_flag, nonfinite, sizeof(double), hipMemcpyDeviceToHost,
                               ctx->stream));

  // ---- results to the caller
  auto copy_out = [&](double *user, const double *dev, const Healpix &h) -> int {
    if (!user || out_dev)
      return HAMX_OK;
    HAMX_HIP(ctx, hipMemcpyAsync(user, dev, h.npix * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
    return HAMX_OK;
  };
  if (d->do_dm)
    HAMX_TRY(copy_out(out->dm, o_dm, h_dm));
  if (d->do_sync) {
    HAMX_TRY(copy_out(out->is, o_is, h_sync));
    HAMX_TRY(copy_out(out->qs, o_qs, h_sync));
    HAMX_TRY(copy_out(out->us, o_us, h_sync));
  }
  if (d->do_fd || d->do_sync)
    HAMX_TRY(copy_out(out->fd, o_fd, h_fd));
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (host_flag != 0.)
    return set_err(ctx, HAMX_ERR_ARG,
                   "non-finite values on the lines of sight: an input grid or parameter "
                   "holds a NaN or an infinity");
  return HAMX_OK;
}
} // namespace hamx

extern "C" int hamx_los_forward(hamx_ctx *ctx, const hamx_los_desc *d,
                                const hamx_fields *fields, hamx_maps *out,
                                uint32_t flags) {
  if (ctx && (flags & HAMX_KEEP_TAPE))
    return los_forward_taped(ctx, d, fields, out, flags);
  return los_forward_impl(ctx, d, fields, out, flags, nullptr, nullptr);
}

// ------------------------------------------- the allocation-free forward --
namespace hamx {
namespace {

// The scratch of hamx_los_forward_static: the five shell maps, the mask at
// the shell's nside (for shells up to nside HAMX_STATIC_MAX_NSIDE) and the
// non-finite output flag.  Every GPU the library runs on holds its own copy
// until the process ends.  In host code `scratch` names a host-side stand-in,
// not device memory: kernels and copies take the device address looked up by
// static_scratch_init.
static_assert(HAMX_STATIC_SCRATCH_BYTES >=
                  6 * (12ull * HAMX_STATIC_MAX_NSIDE * HAMX_STATIC_MAX_NSIDE * sizeof(double) +
                       256) + 512,
              "static scratch below six maps at HAMX_STATIC_MAX_NSIDE and the flag");
static __device__ __align__(256) double scratch[HAMX_STATIC_SCRATCH_BYTES / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_dev[kMaxDevices] = {};  // device address of `scratch` per device
int g_scratch_err[kMaxDevices] = {};    // hipError_t of a failed lookup

} // namespace

int static_scratch_init(hamx_ctx *ctx) {
  // with ctx->device current (hamx_open)
  if (ctx->device < 0 || ctx->device >= kMaxDevices)
    return HAMX_OK; // hamx_los_forward_static reports it
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[ctx->device])
    return HAMX_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    g_scratch_err[ctx->device] = e != hipSuccess ? static_cast<int>(e) : -1;
    return HAMX_OK; // the other entry points do not need it
  }
  g_scratch_dev[ctx->device] = p;
  return HAMX_OK;
}

} // namespace hamx

extern "C" int hamx_los_forward_static(hamx_ctx *ctx, const hamx_los_desc *d,
                                       const hamx_fields *fields, hamx_maps *out,
                                       uint32_t flags) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!d || !out)
    return set_err(ctx, HAMX_ERR_ARG, "null descriptor or output maps");
  const uint32_t dev_io = HAMX_IN_DEVICE | HAMX_OUT_DEVICE;
  if ((flags & dev_io) != dev_io || (flags & ~dev_io))
    return set_err(ctx, HAMX_ERR_ARG,
                   "hamx_los_forward_static takes device inputs and outputs "
                   "(flags HAMX_IN_DEVICE | HAMX_OUT_DEVICE) and no other flag");
  if ((d->do_dm && !out->dm) || (d->do_sync && !(out->is && out->qs && out->us)) ||
      ((d->do_fd || d->do_sync) && !out->fd))
    return set_err(ctx, HAMX_ERR_ARG,
                   "hamx_los_forward_static needs every output map of the pass "
                   "(dm with do_dm, is/qs/us with do_sync, fd with do_fd or do_sync)");
  for (uint32_t s = 0; s < d->n_shell && s < HAMX_MAX_SHELL; ++s)
    if (d->nside_shell[s] > HAMX_STATIC_MAX_NSIDE)
      return set_err(ctx, HAMX_ERR_ARG,
                     "shell %u at nside %u: hamx_los_forward_static holds shells up to "
                     "nside %d in its %d bytes of scratch (use hamx_los_forward)",
                     s, d->nside_shell[s], HAMX_STATIC_MAX_NSIDE, HAMX_STATIC_SCRATCH_BYTES);
  void *base = nullptr;
  int err = 0;
  if (ctx->device >= 0 && ctx->device < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    base = g_scratch_dev[ctx->device];
    err = g_scratch_err[ctx->device];
  }
  if (!base)
    return set_err(ctx, HAMX_ERR_DEVICE,
                   "no static scratch on device %d (address lookup returned %d)",
                   ctx->device, err);
  Arena ar{static_cast<char *>(base), HAMX_STATIC_SCRATCH_BYTES, 0};
  return los_forward_impl(ctx, d, fields, out, flags, nullptr, nullptr, &ar);
}

// ------------------------------------------------------ CRE grid helpers --
namespace hamx {
namespace {

struct CreWriteArgs {
  hamx_box box;
  double lx, ly, lz;
  const double *E; // E_m [erg], host-formed
  uint64_t nE, n;
  double norm;     // j0 * gamma0^(-flux_idx(observer))
  double den0;     // 1 - 1/gamma0
};

// CREfield::write_grid (cre.cc:204-225) with CRE_ana / CRE_unif::write_field
// (cre_ana.cc:52-63, cre_unif.cc:48-59): one thread per output value, E
// fastest, so the stores are contiguous.
__global__ __launch_bounds__(256) void k_cre_write(Phys ph, CreWriteArgs a, double *out) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < a.n;
       t += stride) {
    const uint64_t m = t % a.nE, v = t / a.nE;
    const uint64_t k = v % a.box.nz, j = (v / a.box.nz) % a.box.ny,
                   i = v / (a.box.nz * a.box.ny);
    const Vec3 pos{a.lx * static_cast<double>(i) / static_cast<double>(a.box.nx - 1) + a.box.x_min,
                   a.ly * static_cast<double>(j) / static_cast<double>(a.box.ny - 1) + a.box.y_min,
                   a.lz * static_cast<double>(k) / static_cast<double>(a.box.nz - 1) + a.box.z_min};
    const double gamma = a.E[m] / cgs::mec2;
    const double beta_ratio = sqrt((1. - 1. / gamma) / a.den0);
    const double idx = ph.cre_model == HAMX_CRE_ANA ? cre_index(ph, pos) : -(ph.cre[0]);
    out[t] = a.norm * beta_ratio * pow(gamma, idx) * cre_profile(ph, pos);
  }
}

__global__ __launch_bounds__(256) void k_sync_fg(const double *x, uint64_t n, double *F,
                                                 double *G) {
  __shared__ double tab[kSyncTab];
  load_sync_tab(tab);
  __syncthreads();
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += stride) {
    sync_fg(tab, log(x[t]), x[t], F[t], G[t]);
  }
}

unsigned grid_blocks(uint64_t n) {
  uint64_t b = (n + 255) / 256;
  if (b > 65536)
    b = 65536;
  return static_cast<unsigned>(b < 1 ? 1 : b);
}

} // namespace
} // namespace hamx

extern "C" int hamx_cre_write_grid(hamx_ctx *ctx, const hamx_los_desc *d, const hamx_box *box,
                                   double E_min, double E_fact, uint32_t nE, double *out,
                                   uint32_t flags) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!d || !box || !out)
    return set_err(ctx, HAMX_ERR_ARG, "hamx_cre_write_grid: null argument");
  if (d->cre_model != HAMX_CRE_ANA && d->cre_model != HAMX_CRE_UNIF)
    return set_err(ctx, HAMX_ERR_UNSUPPORTED,
                   "hamx_cre_write_grid: cre model %d is not analytic or uniform", d->cre_model);
  HAMX_TRY(check_box(ctx, *box, "cre"));
  if (nE < 1 || !(E_min > 0.))
    return set_err(ctx, HAMX_ERR_ARG, "hamx_cre_write_grid: empty energy axis");
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  const bool out_dev = (flags & HAMX_OUT_DEVICE) != 0;
  hamx_los_desc dd = *d;
  dd.do_sync = 0;
  const Phys ph = make_phys(&dd);
  CreWriteArgs a;
  std::memset(&a, 0, sizeof(a));
  a.box = *box;
  a.lx = box->x_max - box->x_min;
  a.ly = box->y_max - box->y_min;
  a.lz = box->z_max - box->z_min;
  a.nE = nE;
  a.n = box_size(*box) * nE;
  const bool ana = d->cre_model == HAMX_CRE_ANA;
  const double E0 = ana ? d->cre[5] : d->cre[2];
  const double j0 = ana ? d->cre[6] : d->cre[3];
  const double gamma0 = E0 / cgs::mec2;
  double idx_obs = -(d->cre[0]);
  if (ana) {
    const double r_obs =
        std::sqrt(d->observer[0] * d->observer[0] + d->observer[1] * d->observer[1]) / cgs::kpc;
    const double z_obs = std::fabs(d->observer[2]) / cgs::kpc;
    idx_obs = -d->cre[0] + d->cre[1] * r_obs + d->cre[2] * z_obs;
  }
  a.norm = j0 * std::pow(gamma0, -idx_obs);
  a.den0 = 1. - 1. / gamma0;
  std::vector<double> &tab = ctx->host_tab;
  tab.resize(nE);
  for (uint32_t m = 0; m != nE; ++m)
    tab[m] = E_min * std::exp(m * E_fact);
  void *w;
  HAMX_TRY(ws_get(ctx, kSlotCreTab, nE * sizeof(double), &w));
  HAMX_HIP(ctx, hipMemcpyAsync(w, tab.data(), nE * sizeof(double), hipMemcpyHostToDevice,
                               ctx->stream));
  a.E = static_cast<const double *>(w);
  double *o = out;
  if (!out_dev) {
    HAMX_TRY(ws_get(ctx, kSlotCreOut, a.n * sizeof(double), &w));
    o = static_cast<double *>(w);
  }
  hipLaunchKernelGGL(k_cre_write, dim3(grid_blocks(a.n)), dim3(256), 0, ctx->stream, ph, a, o);
  HAMX_HIP(ctx, hipGetLastError());
  if (!out_dev)
    HAMX_HIP(ctx, hipMemcpyAsync(out, o, a.n * sizeof(double), hipMemcpyDeviceToHost,
                                 ctx->stream));
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}

extern "C" int hamx_sync_kernels(hamx_ctx *ctx, const double *x, uint64_t n, double *F,
                                 double *G, uint32_t flags) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!x || !F || !G)
    return set_err(ctx, HAMX_ERR_ARG, "hamx_sync_kernels: null argument");
  if (n == 0)
    return HAMX_OK;
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  const bool in_dev = (flags & HAMX_IN_DEVICE) != 0;
  const bool out_dev = (flags & HAMX_OUT_DEVICE) != 0;
  void *w;
  HAMX_TRY(ws_get(ctx, kSlotSyncIO, 3 * n * sizeof(double), &w));
  double *buf = static_cast<double *>(w);
  const double *xd = x;
  if (!in_dev) {
    HAMX_HIP(ctx, hipMemcpyAsync(buf, x, n * sizeof(double), hipMemcpyHostToDevice,
                                 ctx->stream));
    xd = buf;
  }
  double *Fd = out_dev ? F : buf + n, *Gd = out_dev ? G : buf + 2 * n;
  hipLaunchKernelGGL(k_sync_fg, dim3(grid_blocks(n)), dim3(256), 0, ctx->stream, xd, n, Fd, Gd);
  HAMX_HIP(ctx, hipGetLastError());
  if (!out_dev) {
    HAMX_HIP(ctx, hipMemcpyAsync(F, Fd, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HAMX_HIP(ctx, hipMemcpyAsync(G, Gd, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  }
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}
