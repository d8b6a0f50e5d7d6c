// Reverse mode of the LOS integrator: the gradient of a scalar L of the
// output maps with respect to the LSA field parameters, the CRE spectral
// index, the uniform TE density and the random-field grids.  The reference
// has no gradients; the forward being differentiated is integrator.cc:23-228.
//
// Forward structure (per shell s, FD carried through the accumulated map):
//   f_k   = te_k B_par,k C delta                  (Faraday depth of sample k)
//   FD_k  = sum_{j<=k} f_j,  qui_k = (fd_in + FD_k) lambda^2 + ipa_k
//   Q = sum_k cos(2 qui_k) jp_k, U = sum_k sin(2 qui_k) jp_k, I = sum_k jt_k
//   maps  = sum_s A_s(shell maps),  fd_in(s) = Interp_s(F_s),
//   F_s   = sum_{t<s} A_t(shell FD_t)
// Reverse:
//   dL/dqui_k = 2 jp_k (-gQ sin 2qui_k + gU cos 2qui_k)
//   dL/df_j   = lambda^2 (T - sum_{k<j} dL/dqui_k) + gFD,  T = sum_k dL/dqui_k
//   dL/dfd_in = lambda^2 T
// so each ray is marched twice: once for T, once for the per-sample
// gradients.  Shells are processed outermost first; the FD adjoint G
// (nside of the FD map) collects A^T of the output FD adjoint plus
// Interp_s^T(dL/dfd_in) of every shell already processed.  The forward is
// recomputed with the FD map checkpointed before each shell.
//
// Scatter: the shell-map adjoints come from A^T (transposed bilinear
// HEALPix interpolation, FP64 atomics), the grid gradient from the
// transposed trilinear gather (FP64 atomics into a packed {x,y,z,te} record
// per vertex), parameter gradients from per-ray partial sums reduced by a
// second kernel.
//
// Polynomial steps with the compiler's own fused multiply-add selection
// (measured faster here than the scalar-operand form the marches use).
#define HAMX_FMA_K_ASM 0

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>

#include "hamx_march.h"

namespace hamx {

struct AdjConst {
  double tc;        // temp_convert factor of I, Q, U
  double l2, two_l2;
  double delta;
  double alpha1;    // 1 - index = 1 + alpha
  double lt, lp;    // d ln jt / d alpha, d ln jp / d alpha without the ln|B_perp| term
  double b0;
};

struct AdjShell {
  const double *gi, *gq, *gu, *gdm, *gfd; // shell-map adjoints (nside_s)
  const double *fd_tot;                   // FD of each ray over this shell (forward tape)
  double *gfdin;                          // out: dL/d(fd_inner) per pixel
  double *ray_par;                        // out: 8 parameter partials per ray
  double *ggrid;                          // packed {x,y,z,te} grid gradient
  uint64_t sy, sx;                        // corner strides of the packed grid
};

template <bool GRIDS, bool PW1, bool MODELS>
__global__ __launch_bounds__(256) void k_los_adj(Phys ph, FastConst fc, AdjConst ac,
                                                 Grids gr, ShellArgs sa, AdjShell as) {
  // kAdjRays consecutive rays per wavefront (set up one per lane), and a
  // per-wavefront cell cache that merges their grid-gradient contributions
  __shared__ RayGeo slots[4][kAdjRays];
  __shared__ __align__(16) unsigned char cache_mem[4][staged_cache_bytes()];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t p0 = (static_cast<uint64_t>(blockIdx.x) * 4 + wave) * kAdjRays;
  if (p0 >= sa.hs.npix)
    return; // uniform per wavefront
  const uint64_t nr = (sa.hs.npix - p0 < static_cast<uint64_t>(kAdjRays))
                          ? sa.hs.npix - p0
                          : static_cast<uint64_t>(kAdjRays);
  if (static_cast<uint64_t>(lane) < nr)
    ray_setup(ph, sa, p0 + lane, slots[wave][lane]);
  const StagedCache cache = staged_at(cache_mem[wave]);
  staged_init(cache, lane);
  for (uint64_t jr = 0; jr < nr; ++jr) {
  const uint64_t ipix = p0 + jr;
  const RayGeo r = slots[wave][jr];
  if (sa.mask && sa.mask[ipix] != 1.0) {
    // a masked-out ray is zero whatever the fields (integrator.cc:63-70):
    // no partials, no grid gradient, no dependence on the inner FD
    if (lane == 0) {
      double *rp = as.ray_par + 8 * ipix;
      for (int m = 0; m < 6; ++m)
        rp[m] = 0.;
      if (as.gfdin)
        as.gfdin[ipix] = 0.;
    }
    continue; // uniform: one ray per wavefront at a time
  }
  const double gq = as.gq ? as.gq[ipix] * ac.tc : 0.;
  const double gu = as.gu ? as.gu[ipix] * ac.tc : 0.;
  const double gi = as.gi ? as.gi[ipix] * ac.tc : 0.;
  const double gdm = as.gdm ? as.gdm[ipix] : 0.;
  const double gfd = as.gfd ? as.gfd[ipix] : 0.;
  // One sweep from the far end of the ray to the observer.  The running FD
  // at sample k is the ray's total (from the forward's tape) minus the
  // samples beyond k, and the suffix sum of dL/dqui is carried the same way,
  // so every sample gets its full gradient in one evaluation.
  const uint64_t n = r.k_hi - r.k_lo;
  const double fd_total = as.fd_tot[ipix];
  double p0 = 0., p1 = 0., p2 = 0., p3 = 0., p4 = 0., p5 = 0.;
  double after = 0.;    // FD of the samples beyond the current chunk
  double dq_after = 0.; // sum of dL/dqui beyond the current chunk
  for (uint64_t c = (n + 63) / 64; c-- > 0;) {
    const uint64_t k = r.k_lo + 64 * c + lane;
    FastSample o;
    bool valid = false;
    if (k < r.k_hi)
      valid = eval_fast<GRIDS, PW1, MODELS>(ph, fc, gr, sa, r, k, o);
    const double incl = wave_incl_scan(valid ? o.fdk : 0., lane);
    const double chunk = wave_last(incl);
    const double fd_here = (fd_total - after - chunk) + incl;
    after += chunk;
    double c2q = 1., s2q = 0., dq = 0.;
    if (valid) {
      double sph, cph;
      sincos_fast(ac.two_l2 * (r.fd_inner + fd_here), sph, cph);
      c2q = cph * o.c2i - sph * o.s2i;
      s2q = sph * o.c2i + cph * o.s2i;
      dq = 2. * o.jp * (-gq * s2q + gu * c2q);
    }
    const double dq_scan = wave_suffix_scan(dq, lane);
    const double dq_suffix = dq_scan + dq_after; // sum over samples k' >= k
    dq_after += __shfl(dq_scan, 0, 64);
    double dte = 0.;
    Vec3 dB{0., 0., 0.};
    if (valid) {
      const double R = ac.l2 * dq_suffix + gfd; // dL/df_k
      const double djp = gq * c2q + gu * s2q;
      const double djt = gi;
      dte = (o.te_raw > 0.) ? (R * o.b_par * fc.fd_scale + gdm * ac.delta) : 0.;
      // dL/dB
      const double ft = R * o.te * fc.fd_scale;
      dB = Vec3{ft * r.l.x, ft * r.l.y, ft * r.l.z};
      if (o.bper2 > 0.) {
        const double coef = (djt * o.jt + djp * o.jp) * ac.alpha1 / (2. * o.bper2);
        dB.x += coef * (o.B.x - o.b_par * r.l.x);
        dB.y += coef * (o.B.y - o.b_par * r.l.y);
        dB.z += coef * (o.B.z - o.b_par * r.l.z);
        const double lnb = 0.25 * log_pos(o.bper2); // ln |B_perp|
        p4 += djt * o.jt * (ac.lt + lnb) + djp * o.jp * (ac.lp + lnb);
      }
      if (o.n2 > 0.) {
        const double c = dq / o.n2;
        dB.x += c * (o.y * r.eph.x - o.x * r.eth.x);
        dB.y += c * (o.y * r.eph.y - o.x * r.eth.y);
        dB.z += c * (o.y * r.eph.z - o.x * r.eth.z);
      }
      if (o.te_in)
        p5 += dte;
      if (o.lsa_in) {
        const double u = o.sps * o.cos_p - o.cps * o.sin_p; // B_x / (b0 cos chi)
        const double v = o.sps * o.sin_p + o.cps * o.cos_p; // B_y / (b0 cos chi)
        p0 += dB.x * (o.cch * u) + dB.y * (o.cch * v) + dB.z * o.sch;
        const double gpsi = ac.b0 * o.cch * (dB.x * v - dB.y * u);
        p1 += gpsi;
        p2 += gpsi * o.lnr;
        const double gchi = ac.b0 * (-o.sch * (dB.x * u + dB.y * v) + o.cch * dB.z);
        p3 += gchi * o.tnh;
      }
    }
    if (GRIDS && as.ggrid) { // wave-cooperative: every lane calls it
      const bool has = valid && o.cell_ok;
      scatter_staged(cache, as.ggrid, has, has ? o.cell.idx : 0, o.cell.xd, o.cell.yd,
                     o.cell.zd, dB.x, dB.y, dB.z, dte, as.sx, as.sy, lane);
    }
  }
  p0 = wave_sum(p0);
  p1 = wave_sum(p1);
  p2 = wave_sum(p2);
  p3 = wave_sum(p3);
  p4 = wave_sum(p4);
  p5 = wave_sum(p5);
  if (lane == 0) {
    double *rp = as.ray_par + 8 * ipix;
    rp[0] = p0;
    rp[1] = p1;
    rp[2] = p2;
    rp[3] = p3;
    rp[4] = p4;
    rp[5] = p5;
    if (as.gfdin)
      as.gfdin[ipix] = ac.l2 * dq_after;
  }
  } // rays of the wavefront
  if (GRIDS && as.ggrid)
    staged_flush_all(cache, as.ggrid, as.sx, as.sy, lane);
}

// dst (shell maps) += A^T(src) for up to 5 map pairs (the transpose of
// k_accumulate: same-nside add, else transposed bilinear interpolation).
struct AccTPairs {
  double *dst[5];
  const double *src[5];
  int n;
};

__global__ __launch_bounds__(256) void k_accumulate_t(AccTPairs p, Healpix hd, Healpix hs) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= hd.npix)
    return;
  if (hd.npix == hs.npix) {
    for (int m = 0; m < p.n; ++m)
      p.dst[m][i] += p.src[m][i];
    return;
  }
  double theta, phi;
  hp_pix2ang(hd, i, theta, phi);
  uint64_t pix[4];
  double w[4];
  hp_interpol_weights(hs, theta, phi, pix, w);
  const double wtot = ((w[0] + w[1]) + w[2]) + w[3];
  for (int m = 0; m < p.n; ++m) {
    const double g = p.src[m][i] / wtot;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      unsafeAtomicAdd(&p.dst[m][pix[q]], w[q] * g);
  }
}

// G (FD-map adjoint) += Interp_s^T(gfdin): transposed interpolation of the
// FD cache at the shell's pixel centres.
__global__ __launch_bounds__(256) void k_fdcache_t(double *G, Healpix hfd,
                                                   const double *gfdin, Healpix hs) {
  const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= hs.npix)
    return;
  const double g = gfdin[p];
  if (g == 0.)
    return;
  double theta, phi;
  hp_pix2ang(hs, p, theta, phi);
  uint64_t pix[4];
  double w[4];
  hp_interpol_weights(hfd, theta, phi, pix, w);
  const double wtot = ((w[0] + w[1]) + w[2]) + w[3];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    unsafeAtomicAdd(&G[pix[q]], w[q] * (g / wtot));
}

// Column sums of the per-ray parameter partials into acc[0..5].
__global__ __launch_bounds__(256) void k_reduce_params(const double *ray_par, uint64_t n,
                                                       double *acc) {
  __shared__ double sm[4];
  const int j = blockIdx.y;
  double v = 0.;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride)
    v += ray_par[8 * i + j];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0)
    sm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    unsafeAtomicAdd(&acc[j], sm[0] + sm[1] + sm[2] + sm[3]);
}

__global__ void k_unpack4(const double *__restrict__ g, double *__restrict__ gx,
                          double *__restrict__ gy, double *__restrict__ gz,
                          double *__restrict__ gt, uint64_t n) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += stride) {
    if (gx) gx[i] = g[4 * i + 0];
    if (gy) gy[i] = g[4 * i + 1];
    if (gz) gz[i] = g[4 * i + 2];
    if (gt) gt[i] = g[4 * i + 3];
  }
}

bool same_box_bits(const hamx_box &a, const hamx_box &b) {
  return a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && a.x_min == b.x_min &&
         a.x_max == b.x_max && a.y_min == b.y_min && a.y_max == b.y_max &&
         a.z_min == b.z_min && a.z_max == b.z_max;
}

// Field-by-field equality of two pass descriptors (the march schedule aside).
bool same_desc(const hamx_los_desc &a, const hamx_los_desc &b) {
  bool eq = a.n_shell == b.n_shell && a.oc_r_res == b.oc_r_res && a.gc_r_min == b.gc_r_min &&
            a.gc_r_max == b.gc_r_max && a.gc_z_min == b.gc_z_min &&
            a.gc_z_max == b.gc_z_max && a.do_dm == b.do_dm && a.do_fd == b.do_fd &&
            a.do_sync == b.do_sync && a.nside_dm == b.nside_dm && a.nside_fd == b.nside_fd &&
            a.nside_sync == b.nside_sync && a.sync_freq == b.sync_freq &&
            a.breg_model == b.breg_model && a.tereg_model == b.tereg_model &&
            a.cre_model == b.cre_model;
  for (int i = 0; i < 3; ++i)
    eq = eq && a.observer[i] == b.observer[i];
  for (int i = 0; i < HAMX_MAX_SHELL; ++i)
    eq = eq && a.nside_shell[i] == b.nside_shell[i];
  for (int i = 0; i <= HAMX_MAX_SHELL; ++i)
    eq = eq && a.radii_shell[i] == b.radii_shell[i];
  for (int i = 0; i < HAMX_BREG_NPARAM; ++i)
    eq = eq && a.breg[i] == b.breg[i];
  for (int i = 0; i < 8; ++i)
    eq = eq && a.cre[i] == b.cre[i];
  for (int i = 0; i < HAMX_TEREG_NPARAM; ++i)
    eq = eq && a.tereg[i] == b.tereg[i];
  return eq;
}

bool same_fields(const hamx_fields &a, const hamx_fields &b) {
  return a.brnd_x == b.brnd_x && a.brnd_y == b.brnd_y && a.brnd_z == b.brnd_z &&
         a.ternd == b.ternd && a.breg_x == b.breg_x && a.breg_y == b.breg_y &&
         a.breg_z == b.breg_z && a.tereg == b.tereg && same_box_bits(a.brnd_box, b.brnd_box) &&
         same_box_bits(a.ternd_box, b.ternd_box) && same_box_bits(a.breg_box, b.breg_box) &&
         same_box_bits(a.tereg_box, b.tereg_box) && a.mask == b.mask &&
         a.nside_mask == b.nside_mask;
}

int los_forward_taped(hamx_ctx *ctx, const hamx_los_desc *d, const hamx_fields *fields,
                      hamx_maps *out, uint32_t flags) {
  HAMX_TRY(validate_los(ctx, d));
  ctx->tape_valid = false;
  const bool fd_map = d->do_fd || d->do_sync;
  const Healpix h_fd = healpix_make(d->do_fd ? d->nside_fd : d->nside_sync);
  void *q;
  double *ckpt = nullptr, *shell_fd = nullptr;
  if (d->do_fd) {
    HAMX_TRY(ws_get(ctx, kSlotAdjCkpt, d->n_shell * h_fd.npix * sizeof(double), &q));
    ckpt = static_cast<double *>(q);
  }
  if (fd_map) {
    HAMX_TRY(ws_get(ctx, kSlotTapeFd, shell_offset(d, d->n_shell) * sizeof(double), &q));
    shell_fd = static_cast<double *>(q);
  }
  HAMX_TRY(los_forward_impl(ctx, d, fields, out, flags & ~HAMX_KEEP_TAPE, ckpt, shell_fd));
  if (fd_map) {
    ctx->tape_valid = true;
    ctx->tape_desc = *d;
    std::memset(&ctx->tape_fields, 0, sizeof(ctx->tape_fields));
    if (fields)
      ctx->tape_fields = *fields;
  }
  return HAMX_OK;
}

static unsigned blocks_for(uint64_t n, unsigned per = 256) {
  return static_cast<unsigned>((n + per - 1) / per);
}

} // namespace hamx

using namespace hamx;

extern "C" int hamx_los_adjoint(hamx_ctx *ctx, const hamx_los_desc *d,
                                const hamx_fields *fields, const hamx_maps *adj,
                                double *dparams, double *dbrnd_x, double *dbrnd_y,
                                double *dbrnd_z, double *dternd, uint32_t flags) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!d || !adj || !dparams)
    return set_err(ctx, HAMX_ERR_ARG, "hamx_los_adjoint: null descriptor, adjoint maps or "
                                      "parameter gradient");
  HAMX_TRY(validate_los(ctx, d));
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  const bool in_dev = (flags & HAMX_IN_DEVICE) != 0;
  const bool out_dev = (flags & HAMX_OUT_DEVICE) != 0;
  hamx_fields none;
  std::memset(&none, 0, sizeof(none));
  const hamx_fields &f = fields ? *fields : none;
  const Phys ph = make_phys(d);
  if (!(ph.idx_const && d->do_sync &&
        (d->cre_model == HAMX_CRE_ANA || d->cre_model == HAMX_CRE_UNIF)))
    return set_err(ctx, HAMX_ERR_UNSUPPORTED,
                   "hamx_los_adjoint supports the analytic or uniform CRE with a constant "
                   "index, with synchrotron output");
  // any regular B and TE model: the LSA parameters' partials are zero for
  // the other B models, n0's for TE models other than the uniform sphere
  const bool models = !(d->breg_model == HAMX_BREG_LSA && d->tereg_model == HAMX_TEREG_UNIF);

  // ---- the forward's tape
  const uint64_t nside_fd_map = d->do_fd ? d->nside_fd : d->nside_sync;
  const Healpix h_fd = healpix_make(nside_fd_map);
  const Healpix h_sync = healpix_make(d->nside_sync);
  const Healpix h_dm = healpix_make(d->do_dm ? d->nside_dm : 1);
  void *q;
  if (flags & HAMX_USE_TAPE) {
    if (!ctx->tape_valid || !same_desc(ctx->tape_desc, *d) ||
        !same_fields(ctx->tape_fields, f))
      return set_err(ctx, HAMX_ERR_ARG,
                     "hamx_los_adjoint: no tape of a forward pass with this descriptor and "
                     "these field arrays (run hamx_los_forward with HAMX_KEEP_TAPE)");
  } else {
    hamx_maps none_maps;
    std::memset(&none_maps, 0, sizeof(none_maps));
    HAMX_TRY(los_forward_taped(ctx, d, fields, &none_maps, flags & HAMX_IN_DEVICE));
  }
  double *ckpt = static_cast<double *>(ctx->ws[kSlotAdjCkpt]);
  const double *tape_fd = static_cast<const double *>(ctx->ws[kSlotTapeFd]);

  // ---- grids and adjoint inputs on the device
  Grids gr;
  bool packed = false, grids_any = false;
  HAMX_TRY(stage_fields(ctx, d, f, in_dev, gr, packed, grids_any));
  if (grids_any && !packed)
    return set_err(ctx, HAMX_ERR_UNSUPPORTED,
                   "hamx_los_adjoint needs the random B and TE grids on one box");
  // the mask, duplicated to each shell's nside as the forward does
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
  const void *cp;
  const double *a_is = nullptr, *a_qs = nullptr, *a_us = nullptr, *a_dm = nullptr,
               *a_fd = nullptr;
  HAMX_TRY(stage_in(ctx, kSlotAdjIs, adj->is, h_sync.npix * sizeof(double), in_dev, &cp));
  a_is = static_cast<const double *>(cp);
  HAMX_TRY(stage_in(ctx, kSlotAdjQs, adj->qs, h_sync.npix * sizeof(double), in_dev, &cp));
  a_qs = static_cast<const double *>(cp);
  HAMX_TRY(stage_in(ctx, kSlotAdjUs, adj->us, h_sync.npix * sizeof(double), in_dev, &cp));
  a_us = static_cast<const double *>(cp);
  if (d->do_dm) {
    HAMX_TRY(stage_in(ctx, kSlotAdjDm, adj->dm, h_dm.npix * sizeof(double), in_dev, &cp));
    a_dm = static_cast<const double *>(cp);
  }
  if (d->do_fd) {
    HAMX_TRY(stage_in(ctx, kSlotAdjFd, adj->fd, h_fd.npix * sizeof(double), in_dev, &cp));
    a_fd = static_cast<const double *>(cp);
  }

  // ---- work buffers
  uint64_t max_npix = 0;
  for (uint32_t s = 0; s < d->n_shell; ++s) {
    const uint64_t ns = d->nside_shell[s];
    if (12 * ns * ns > max_npix)
      max_npix = 12 * ns * ns;
  }
  double *sg[5];
  for (int m = 0; m < 5; ++m) {
    HAMX_TRY(ws_get(ctx, kSlotAdjShell0 + m, max_npix * sizeof(double), &q));
    sg[m] = static_cast<double *>(q);
  }
  double *mask_shell = nullptr;
  if (mask_piv) {
    HAMX_TRY(ws_get(ctx, kSlotMaskShell, max_npix * sizeof(double), &q));
    mask_shell = static_cast<double *>(q);
  }
  HAMX_TRY(ws_get(ctx, kSlotAdjG, h_fd.npix * sizeof(double), &q));
  double *G = static_cast<double *>(q);
  HAMX_TRY(ws_get(ctx, kSlotAdjFdin, max_npix * sizeof(double), &q));
  double *gfdin = static_cast<double *>(q);
  HAMX_TRY(ws_get(ctx, kSlotAdjRayPar, 8 * max_npix * sizeof(double), &q));
  double *ray_par = static_cast<double *>(q);
  HAMX_TRY(ws_get(ctx, kSlotAdjPar, 8 * sizeof(double), &q));
  double *acc = static_cast<double *>(q);
  HAMX_TRY(fill_f64(ctx, acc, 8, 0.));
  double *ggrid = nullptr;
  uint64_t ngrid = 0;
  if (packed) {
    ngrid = box_size(f.brnd_box);
    HAMX_TRY(ws_get(ctx, kSlotAdjGrid, 4 * ngrid * sizeof(double), &q));
    ggrid = static_cast<double *>(q);
    HAMX_TRY(fill_f64(ctx, ggrid, 4 * ngrid, 0.));
  }
  if (d->do_fd) {
    if (a_fd)
      HAMX_HIP(ctx, hipMemcpyAsync(G, a_fd, h_fd.npix * sizeof(double),
                                   hipMemcpyDeviceToDevice, ctx->stream));
    else
      HAMX_TRY(fill_f64(ctx, G, h_fd.npix, 0.));
  }

  // ---- constants
  AdjConst ac;
  std::memset(&ac, 0, sizeof(ac));
  ac.tc = ph.tc_em1 * ph.tc_em1 / ph.tc_den;
  ac.l2 = ph.lambda_sq;
  ac.two_l2 = 2. * ph.lambda_sq;
  ac.alpha1 = 1. - ph.idx;
  ac.b0 = d->breg[0];
  {
    const double alpha = -ph.idx;
    const double E0 = d->cre_model == HAMX_CRE_UNIF ? d->cre[2] : d->cre[5];
    const double ln_g0 = std::log(E0 / cgs::mec2);
    const double half_ln_a = 0.5 * std::log(ph.A_num / ph.three_qe);
    const double psi_t = 0.25 * digamma(alpha / 4. + 19. / 12.);
    const double psi_p = 0.25 * digamma(alpha / 4. + 7. / 12.);
    const double psi_c = 0.25 * digamma(alpha / 4. - 1. / 12.);
    ac.lt = ln_g0 - half_ln_a - 1. / (1. + alpha) + psi_t + psi_c;
    ac.lp = ln_g0 - half_ln_a + psi_p + psi_c;
  }
  const bool pw1 = 0.5 * (ph.idx + 1) == -1.0;

  // ---- shells, outermost first
  for (int s = static_cast<int>(d->n_shell) - 1; s >= 0; --s) {
    ShellArgs sa;
    std::memset(&sa, 0, sizeof(sa));
    sa.hs = healpix_make(d->nside_shell[s]);
    shell_ref(d, static_cast<uint32_t>(s), sa.step, sa.d_start, sa.delta);
    const bool has_fdin = d->do_fd && s > 0;
    sa.fd_map = has_fdin ? ckpt + static_cast<uint64_t>(s) * h_fd.npix : nullptr;
    sa.hfd = h_fd;
    const uint64_t npix = sa.hs.npix;
    if (mask_piv) {
      HAMX_TRY(mask_at_nside(ctx, mask_shell, sa.hs, mask_piv, h_mask));
      sa.mask = mask_shell;
    }
    for (int m = 0; m < 5; ++m)
      HAMX_TRY(fill_f64(ctx, sg[m], npix, 0.));
    // shell-map adjoints: A_s^T of the output adjoints
    auto acc_t = [&](AccTPairs &pairs, const Healpix &hd) -> int {
      if (pairs.n == 0)
        return HAMX_OK;
      hipLaunchKernelGGL(k_accumulate_t, dim3(blocks_for(hd.npix)), dim3(256), 0,
                         ctx->stream, pairs, hd, sa.hs);
      HAMX_HIP(ctx, hipGetLastError());
      return HAMX_OK;
    };
    AccTPairs p1;
    std::memset(&p1, 0, sizeof(p1));
    const double *sync_src[3] = {a_is, a_qs, a_us};
    for (int m = 0; m < 3; ++m)
      if (sync_src[m]) { // a NULL adjoint map is zero
        p1.dst[p1.n] = sg[m];
        p1.src[p1.n] = sync_src[m];
        ++p1.n;
      }
    HAMX_TRY(acc_t(p1, h_sync));
    if (d->do_dm && a_dm) {
      AccTPairs p2;
      std::memset(&p2, 0, sizeof(p2));
      p2.dst[0] = sg[3]; p2.src[0] = a_dm; p2.n = 1;
      HAMX_TRY(acc_t(p2, h_dm));
    }
    if (d->do_fd) {
      AccTPairs p3;
      std::memset(&p3, 0, sizeof(p3));
      p3.dst[0] = sg[4]; p3.src[0] = G; p3.n = 1;
      HAMX_TRY(acc_t(p3, h_fd));
    }
    AdjShell as;
    std::memset(&as, 0, sizeof(as));
    as.gi = sg[0];
    as.gq = sg[1];
    as.gu = sg[2];
    as.gdm = d->do_dm ? sg[3] : nullptr;
    as.gfd = d->do_fd ? sg[4] : nullptr;
    as.gfdin = has_fdin ? gfdin : nullptr;
    as.fd_tot = tape_fd + shell_offset(d, static_cast<uint32_t>(s));
    as.ray_par = ray_par;
    as.ggrid = ggrid;
    as.sy = gr.g_packed.nz;
    as.sx = gr.g_packed.ny * gr.g_packed.nz;
    const FastConst fc = make_fast(d, ph, sa.delta, pw1);
    ac.delta = sa.delta;
    const unsigned blocks =
        static_cast<unsigned>((npix + 4 * kAdjRays - 1) / (4 * kAdjRays));
    ProfMark pm = prof_begin(ctx);
#define HAMX_ADJ_LAUNCH(G, P, M)                                                           \
  hipLaunchKernelGGL((k_los_adj<G, P, M>), dim3(blocks), dim3(256), 0, ctx->stream, ph, fc, ac, \
                     gr, sa, as)
#define HAMX_ADJ_LAUNCH_M(G, P)                                                            \
  do {                                                                                     \
    if (models)                                                                            \
      HAMX_ADJ_LAUNCH(G, P, true);                                                         \
    else                                                                                   \
      HAMX_ADJ_LAUNCH(G, P, false);                                                        \
  } while (0)
    if (packed && pw1)
      HAMX_ADJ_LAUNCH_M(true, true);
    else if (packed)
      HAMX_ADJ_LAUNCH_M(true, false);
    else if (pw1)
      HAMX_ADJ_LAUNCH_M(false, true);
    else
      HAMX_ADJ_LAUNCH_M(false, false);
#undef HAMX_ADJ_LAUNCH_M
#undef HAMX_ADJ_LAUNCH
    HAMX_HIP(ctx, hipGetLastError());
    prof_end(ctx, pm, "los_adjoint", static_cast<uint32_t>(s), npix * sa.step);
    unsigned rb = blocks_for(npix);
    if (rb > 256)
      rb = 256;
    hipLaunchKernelGGL(k_reduce_params, dim3(rb, 6), dim3(256), 0, ctx->stream,
                       static_cast<const double *>(ray_par), npix, acc);
    HAMX_HIP(ctx, hipGetLastError());
    if (has_fdin) {
      hipLaunchKernelGGL(k_fdcache_t, dim3(blocks_for(npix)), dim3(256), 0, ctx->stream, G,
                         h_fd, static_cast<const double *>(gfdin), sa.hs);
      HAMX_HIP(ctx, hipGetLastError());
    }
  }

  // ---- outputs
  if (packed && (dbrnd_x || dbrnd_y || dbrnd_z || dternd)) {
    double *ox = dbrnd_x, *oy = dbrnd_y, *oz = dbrnd_z, *ot = dternd;
    if (!out_dev) {
      HAMX_TRY(ws_get(ctx, kSlotRndOut0, 4 * ngrid * sizeof(double), &q));
      double *stage = static_cast<double *>(q);
      ox = dbrnd_x ? stage : nullptr;
      oy = dbrnd_y ? stage + ngrid : nullptr;
      oz = dbrnd_z ? stage + 2 * ngrid : nullptr;
      ot = dternd ? stage + 3 * ngrid : nullptr;
    }
    unsigned ub = blocks_for(ngrid);
    if (ub > 16384)
      ub = 16384;
    hipLaunchKernelGGL(k_unpack4, dim3(ub), dim3(256), 0, ctx->stream,
                       static_cast<const double *>(ggrid), ox, oy, oz, ot, ngrid);
    HAMX_HIP(ctx, hipGetLastError());
    if (!out_dev) {
      const size_t nb = ngrid * sizeof(double);
      if (dbrnd_x)
        HAMX_HIP(ctx, hipMemcpyAsync(dbrnd_x, ox, nb, hipMemcpyDeviceToHost, ctx->stream));
      if (dbrnd_y)
        HAMX_HIP(ctx, hipMemcpyAsync(dbrnd_y, oy, nb, hipMemcpyDeviceToHost, ctx->stream));
      if (dbrnd_z)
        HAMX_HIP(ctx, hipMemcpyAsync(dbrnd_z, oz, nb, hipMemcpyDeviceToHost, ctx->stream));
      if (dternd)
        HAMX_HIP(ctx, hipMemcpyAsync(dternd, ot, nb, hipMemcpyDeviceToHost, ctx->stream));
    }
  }
  if (out_dev)
    HAMX_HIP(ctx, hipMemcpyAsync(dparams, acc, HAMX_NPARAM * sizeof(double),
                                 hipMemcpyDeviceToDevice, ctx->stream));
  else
    HAMX_HIP(ctx, hipMemcpyAsync(dparams, acc, HAMX_NPARAM * sizeof(double),
                                 hipMemcpyDeviceToHost, ctx->stream));
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}
