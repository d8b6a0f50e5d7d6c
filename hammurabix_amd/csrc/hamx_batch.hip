// Batched LOS integration: many instances (realisations x parameter sets x
// frequency lists) per replay of one captured HIP graph.
//
// Per instance the result is Pipeline::assemble_obs (pipeline.cc:165-190)
// over its frequency list: the last-listed frequency's pass first, with DM,
// FD and the FD cache of the inner shells (integrator.cc:75-86), then one
// pass per remaining frequency without DM / FD and with a zero FD cache.
// All passes of an instance share one march: position, limit tests, fields,
// FD and the emissivity shape are evaluated once per sample; per frequency
// only the polarisation phase (a function of the running FD) and constant
// factors differ.
//
// Work is organised by shell level: level s holds shell s of every instance
// that has one.  A level is one march launch over "items" (a few consecutive
// rays of one instance, one 64-lane wavefront per ray, lanes = consecutive
// samples as in k_los_fast), followed by one accumulate launch that adds the
// shell maps of all instances into their outputs (and so into the FD cache
// read by the next level).  The graph is: zero outputs and tickets, pack the
// grids, then march / accumulate per level.
//
// Launch shapes (chosen by timing every candidate on the batch itself):
//   one-shot    one wave per item, the hardware dispatcher balances;
//               optionally XCD-aware (workgroups numbered so that each of the
//               8 XCDs, which take workgroups round-robin, gets a contiguous
//               run of items and so a compact region of sky and grid in its
//               own L2);
//   persistent  as many workgroups as fit on the device; each wave takes
//               tickets from an atomic counter; optionally one queue per XCD
//               (a contiguous, cost-balanced slice of the items) with
//               stealing from the other queues when its own is empty.
//
// Precision HAMX_PREC_F32 evaluates the fields and the emissivity shape in
// FP32 in kpc / muG units on FP32 grids, with the hardware sin/cos/exp/log;
// positions, the galactocentric limit tests, the grid-cell lookup and every
// in/out decision stay FP64 exactly as in the plain path, and so do the FD
// scan, the polarisation phase (reduced to one revolution in FP64 before the
// FP32 sine / cosine) and the cross-lane map sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <new>
#include <numeric>
#include <string>
#include <type_traits>
#include <vector>

#include "hamx_march.h"

namespace hamx {
namespace {

constexpr int kMaxF = HAMX_MAX_FREQ;
constexpr int kQueues = 8;   // one ticket queue per XCD
constexpr int kAccMaps = 12; // 3 Stokes maps x 4 frequencies
constexpr int kItemRaysMax = 16; // rays per work item (set up one per lane)

struct BShell {
  Healpix hs;
  uint64_t step;
  double d_start, delta;
  double fdu; // FP32 mode: te * B_par[muG] -> FD contribution, fd_fore * 1e-6 * delta
};

// FP32 field constants in kpc / muG units.
struct F32Const {
  double b0d, n0d; // b0 [muG] and n0 for the FD product, kept in FP64
  float psi0, psi1, chi0, b0u;
  // the LSA angles psi(r) and chi(z) stay within [-pi/4, pi/4] over the
  // field's window (checked on the host): their sines and cosines are then
  // the series alone, with no quadrant reduction
  int lsa_direct;
  float n0;
  float R0k, inv_r0k, inv_z0k;
  float hexp; // (1 - pw_e) / 2: the emissivity shape is prof * (B_perp^2)^hexp
};

// Reverse-mode state of one instance (device pointers into the plan).
struct BAdj {
  double *g_i[kMaxF], *g_q[kMaxF], *g_u[kMaxF]; // shell-map adjoints of the level
  double *g_dm, *g_fd;
  double *G;        // FD-map adjoint (nside_fd)
  double *gfdin;    // dL/d(FD cache) per ray of the level
  double *ckpt;     // FD map before each level (pass 0), n_shell * npix(fd)
  double *ray_par;  // 6 parameter partials per ray, laid out as t_fd
  void *ggrid;      // packed grid gradient {bx, by, bz, te} (float4 / double4)
  double lt[kMaxF]; // d ln(jt_f)/d alpha without the ln|B_perp| term
  double lp[kMaxF]; // d ln(jp_f)/d alpha without the ln|B_perp| term
  double l2[kMaxF]; // lambda_f^2
  double b0;        // LSA b0 [G]
};

struct BInst {
  Phys ph;                      // pass-0 physics (the last-listed frequency)
  FastConst fc[HAMX_MAX_SHELL]; // pass-0 specialised constants per shell
  BShell sh[HAMX_MAX_SHELL];
  F32Const f;
  // FP32 mode: squared limits and grid scales (the decisions are made on
  // FP64 positions; comparing squares instead of square roots, and scaling
  // by a reciprocal, move a decision only for a sample within rounding of a
  // boundary)
  double r_min2, r_max2, te_r02, cre_r02;
  double gx0, gy0, gz0, gsx, gsy, gsz; // grid origin and (n - 1) / extent
  int n_shell, n_freq, grids, pw1;
  int do_dm, do_fd, pad0, pad1;
  // per pass (index 0 = first pass)
  double two_l2[kMaxF]; // FP64 phase: 2 lambda^2
  double rev[kMaxF];    // FP32 mode phase in revolutions per unit FD: lambda^2 / pi
  int fdin[kMaxF];      // the pass reads the FD cache
  double st[kMaxF];     // map scale of the I sum (temp_convert included)
  double sp[kMaxF];     // map scale of the Q / U sums
  const void *packed;   // double4 (F64) or float4 (F32) {bx, by, bz, te} records
  GridGeo geo;
  Healpix hfd;
  const double *fd_cache; // accumulated FD map of the inner shells
  double *t_dm;
  double *t_fd;                    // each ray's FD per level (the adjoint's tape)
  uint64_t fd_off[HAMX_MAX_SHELL]; // first ray of level s in t_fd / ray_par
  double *t_i[kMaxF], *t_q[kMaxF], *t_u[kMaxF];
  BAdj adj;
};

struct LevelArgs {
  const BInst *inst;
  const uint32_t *ids;    // instances of this level, in item order
  const uint64_t *prefix; // [n_ids + 1] first item of each instance
  unsigned long long *ctr; // [kQueues] tickets, zero before the launch
  uint64_t n_items;
  uint64_t qbeg[kQueues], qend[kQueues];
  uint32_t n_ids, level, rays, nq;
  int persistent, xcd;
};

__device__ inline uint64_t uniform_u64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ inline uint64_t wave_max_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

__device__ inline uint64_t wave_min_u64(uint64_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const uint64_t w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}

__device__ inline ShellArgs shell_args(const BInst &in, int s) {
  ShellArgs sa;
  sa.hs = in.sh[s].hs;
  sa.step = in.sh[s].step;
  sa.d_start = in.sh[s].d_start;
  sa.delta = in.sh[s].delta;
  sa.fd_map = (in.do_fd && s > 0) ? in.fd_cache : nullptr;
  sa.hfd = in.hfd;
  sa.t_dm = sa.t_fd = sa.t_is = sa.t_qs = sa.t_us = nullptr;
  return sa;
}

// Per-ray set-up (pointing, LOS versor, FD cache, culled step range) as an
// out-of-line call: it runs once per ray, and keeping its FP64 sincos and
// HEALPix integer arithmetic out of the march body leaves the body's
// registers to the sample loop.
__device__ __noinline__ void ray_setup_call(const Phys &ph, const ShellArgs &sa, uint64_t ipix,
                                            RayGeo &r) {
  ray_setup(ph, sa, ipix, r);
}

// ------------------------------------------------------------- FP64 ray --

template <bool GRIDS, bool PW1>
__device__ void march_ray_f64(const BInst &in, int s, uint64_t ipix, int lane,
                              const RayGeo &r) {
  const ShellArgs sa = shell_args(in, s);
  Grids gr;
  gr.packed = static_cast<const double4 *>(in.packed);
  gr.g_packed = in.geo;
  const FastConst &fc = in.fc[s];
  const int nf = in.n_freq;
  double dm = 0., I = 0., Q[kMaxF] = {0., 0., 0., 0.}, U[kMaxF] = {0., 0., 0., 0.};
  double carry = 0.;
  for (uint64_t base = r.k_lo; base < r.k_hi; base += 64) {
    const uint64_t k = base + lane;
    FastSample o;
    bool valid = false;
    if (k < r.k_hi)
      valid = eval_fast<GRIDS, PW1>(in.ph, fc, gr, sa, r, k, o);
    const double incl = wave_incl_scan(valid ? o.fdk : 0., lane);
    const double fd_here = carry + incl;
    carry += wave_last(incl);
    if (valid) {
      dm += o.te;
      I += o.jt;
#pragma unroll
      for (int f = 0; f < kMaxF; ++f) {
        if (f < nf) {
          const double X = in.fdin[f] ? r.fd_inner + fd_here : fd_here;
          double sph, cph;
          sincos(in.two_l2[f] * X, &sph, &cph);
          Q[f] += (cph * o.c2i - sph * o.s2i) * o.jp;
          U[f] += (sph * o.c2i + cph * o.s2i) * o.jp;
        }
      }
    }
  }
  dm = wave_sum(dm);
  I = wave_sum(I);
#pragma unroll
  for (int f = 0; f < kMaxF; ++f) {
    if (f < nf) {
      Q[f] = wave_sum(Q[f]);
      U[f] = wave_sum(U[f]);
    }
  }
  if (lane == 0) {
    if (in.do_dm) // a shell thinner than one step has no samples and no finite delta
      in.t_dm[ipix] = sa.step ? dm * sa.delta : 0.;
    in.t_fd[in.fd_off[s] + ipix] = carry;
#pragma unroll
    for (int f = 0; f < kMaxF; ++f) {
      if (f < nf) {
        in.t_i[f][ipix] = I * in.st[f];
        in.t_q[f][ipix] = Q[f] * in.sp[f];
        in.t_u[f][ipix] = U[f] * in.sp[f];
      }
    }
  }
}

// ------------------------------------------------------------- FP32 ray --

// One FP32 sample with the intermediates the reverse mode needs; the
// forward march reads te, fdk, w, c2i and s2i only.
struct S32 {
  double fdk, b_pard; // FD contribution and B.l: FP64 products of the FP32 field
  float te, w, c2i, s2i;
  float bx, by, bz, b_par, bper2, x, y, n2;
  float te_raw;
  bool lsa_in, te_in, cell_ok;
  float cos_p, sin_p, sps, cps, sch, cch, lnr, tnh;
  uint64_t idx;
  float xd, yd, zd;
  int cx, cy, cz; // the cell's lower corner
};

__device__ inline void lerp_plane_f4(const float4 *__restrict__ f, uint64_t i0, uint64_t sy,
                                     float zd, float yd, float4 &w) {
  // FP32 sample arithmetic: a * b + c may fuse (more accurate, fewer instructions)
  #pragma clang fp contract(fast)
  const float4 a0 = f[i0], a1 = f[i0 + 1];
  const float4 b0 = f[i0 + sy], b1 = f[i0 + sy + 1];
  const float z1 = 1.f - zd, y1 = 1.f - yd;
  w.x = (a0.x * z1 + a1.x * zd) * y1 + (b0.x * z1 + b1.x * zd) * yd;
  w.y = (a0.y * z1 + a1.y * zd) * y1 + (b0.y * z1 + b1.y * zd) * yd;
  w.z = (a0.z * z1 + a1.z * zd) * y1 + (b0.z * z1 + b1.z * zd) * yd;
  w.w = (a0.w * z1 + a1.w * zd) * y1 + (b0.w * z1 + b1.w * zd) * yd;
}

// sin and cos of an FP32 angle |r| <= pi/4: the Taylor series to the terms
// r^9 and r^10 (truncation below 2e-9 there; relative accuracy also for
// small angles, where the hardware v_sin_f32's absolute error of ~1e-7 is a
// large relative one, and smooth along a ray, so it would add up in the FD).
__device__ inline void sincos_poly(float r, float &sn, float &cs) {
  // FP32 sample arithmetic: a * b + c may fuse (more accurate, fewer instructions)
  #pragma clang fp contract(fast)
  const float r2 = r * r;
  sn = r * (1.f + r2 * (-1.f / 6.f + r2 * (1.f / 120.f + r2 * (-1.f / 5040.f +
                                                               r2 * (1.f / 362880.f)))));
  cs = 1.f + r2 * (-0.5f + r2 * (1.f / 24.f + r2 * (-1.f / 720.f + r2 * (1.f / 40320.f +
                                                                       r2 * (-1.f / 3628800.f)))));
}

// sin and cos of any FP32 angle: quadrant reduction k = rint(2x/pi) with
// the remainder x - k pi/2 taken in FP64 (exact for any angle a field model
// produces), then sincos_poly on the remainder.
__device__ inline void sincos_f32(float x, float &sn, float &cs) {
  const double xd = static_cast<double>(x);
  const double k = rint(xd * 0.63661977236758134308); // 2/pi
  const float r = static_cast<float>(xd - k * 1.57079632679489661923);
  float s, c;
  sincos_poly(r, s, c);
  const int q = static_cast<int>(static_cast<int64_t>(k) & 3);
  sn = (q == 0) ? s : (q == 1) ? c : (q == 2) ? -s : -c;
  cs = (q == 0) ? c : (q == 1) ? -s : (q == 2) ? -c : s;
}

// grid_cell with the division by the extent replaced by a product with
// (n - 1) / extent (FP32 mode).
__device__ inline bool grid_cell_scaled(const BInst &in, const Vec3 &p, Cell &c, int &cx, int &cy,
                                        int &cz) {
  const GridGeo &g = in.geo;
  double t = (p.x - in.gx0) * in.gsx;
  if (!(t > 0 && t < g.fx))
    return false;
  const uint32_t xl = static_cast<uint32_t>(t); // t > 0: truncation is the floor
  c.xd = t - static_cast<double>(xl);
  t = (p.y - in.gy0) * in.gsy;
  if (!(t > 0 && t < g.fy))
    return false;
  const uint32_t yl = static_cast<uint32_t>(t);
  c.yd = t - static_cast<double>(yl);
  t = (p.z - in.gz0) * in.gsz;
  if (!(t > 0 && t < g.fz))
    return false;
  const uint32_t zl = static_cast<uint32_t>(t);
  c.zd = t - static_cast<double>(zl);
  c.idx = static_cast<uint64_t>(xl) * g.nyz + (yl * static_cast<uint32_t>(g.nz) + zl);
  cx = static_cast<int>(xl);
  cy = static_cast<int>(yl);
  cz = static_cast<int>(zl);
  return true;
}

// The sample's FP64 position and the decisions taken on it (limit tests,
// the LSA window's squared radius, the TE and CRE spheres), evaluated as
// written: this function sits outside the FP32 code's contraction pragma,
// so no product here fuses into an FMA and the decisions round as the FP64
// plain path's do.
struct Geo64 {
  Vec3 pos;
  double rr2;        // squared cylindrical radius
  bool te_in, cre_in; // inside the uniform TE / CRE spheres
};

__device__ inline bool geo_f64(const BInst &in, const BShell &sh, const RayGeo &r, uint64_t k,
                               Geo64 &g) {
  const Phys &ph = in.ph;
  const double dist =
      sh.d_start + (static_cast<double>(static_cast<uint32_t>(k)) + 0.5) * sh.delta;
  g.pos = Vec3{r.l.x * dist + ph.obs.x, r.l.y * dist + ph.obs.y, r.l.z * dist + ph.obs.z};
  const Vec3 &pos = g.pos;
  const double len2 = pos.x * pos.x + pos.y * pos.y + pos.z * pos.z;
  if (len2 < in.r_min2 || len2 > in.r_max2 || pos.z < ph.gc_z_min || pos.z > ph.gc_z_max)
    return false;
  g.rr2 = pos.x * pos.x + pos.y * pos.y;
  const double dx = pos.x - ph.obs.x, dy = pos.y - ph.obs.y, dz = pos.z - ph.obs.z;
  const double d2 = dx * dx + dy * dy + dz * dz;
  g.te_in = !(d2 > in.te_r02);
  g.cre_in = !(d2 > in.cre_r02);
  return true;
}

template <bool GRIDS, bool PW1>
__device__ inline bool eval_f32(const BInst &in, const BShell &sh, const RayGeo &r,
                                const float3 &lf, const float3 &ethf, const float3 &ephf,
                                uint64_t k, S32 &o) {
  // FP32 sample arithmetic: a * b + c may fuse (more accurate, fewer
  // instructions); geo_f64's body keeps the FP64 decisions unfused
  #pragma clang fp contract(fast)
  Geo64 geo;
  if (!geo_f64(in, sh, r, k, geo))
    return false;
  const Phys &ph = in.ph;
  const Vec3 &pos = geo.pos;
  const double inv_kpc = 1.0 / cgs::kpc;
  const double rr2 = geo.rr2;
  const float rk = sqrtf(static_cast<float>(rr2 * (inv_kpc * inv_kpc)));
  const float zk = static_cast<float>(pos.z * inv_kpc);
  const F32Const &c = in.f;
  float bx = 0.f, by = 0.f, bz = 0.f;
  // FD = te B.l accumulates along the ray, so a systematic FP32 rounding of
  // b0, n0 or the direction would be a systematic FD error: B.l is summed
  // in FP64 with the FP64 direction, the LSA part as b0 (FP64) times its
  // unit vector, and te as n0 (FP64) plus the FP32 grid value
  double bpar = 0., te_d = 0.;
  // WMAP-LSA window (breg_lsa.cc:9-26)
  o.lsa_in = !(rr2 > (400. * cgs::kpc * cgs::kpc) || rr2 < (9. * cgs::kpc * cgs::kpc));
  if (o.lsa_in) {
    const float xk = static_cast<float>(pos.x * inv_kpc);
    const float yk = static_cast<float>(pos.y * inv_kpc);
    const float inv_r = __builtin_amdgcn_rcpf(rk);
    o.cos_p = xk * inv_r;
    o.sin_p = yk * inv_r;
    // ln(r / 8 kpc) and exp(-2|z|) from the hardware v_log_f32 (log2) and
    // v_exp_f32 (exp2), within 1 ulp: an absolute error of 1e-7 in ln r
    // moves psi by psi1 * 1e-7, and the rounding of the exp2 argument moves
    // tanh by less than 2e-8
    o.lnr = __builtin_amdgcn_logf(rk * 0.125f) * 0.693147181f;
    const float e2 = __builtin_amdgcn_exp2f(fabsf(zk) * -2.88539008f);
    o.tnh = copysignf((1.f - e2) * __builtin_amdgcn_rcpf(1.f + e2), zk);
    const float psi = c.psi0 + c.psi1 * o.lnr;
    const float chi = c.chi0 * o.tnh;
    if (c.lsa_direct) {
      sincos_poly(psi, o.sps, o.cps);
      sincos_poly(chi, o.sch, o.cch);
    } else {
      sincos_f32(psi, o.sps, o.cps);
      sincos_f32(chi, o.sch, o.cch);
    }
    const float ux = o.cch * (o.sps * o.cos_p - o.cps * o.sin_p);
    const float uy = o.cch * (o.sps * o.sin_p + o.cps * o.cos_p);
    bx = c.b0u * ux;
    by = c.b0u * uy;
    bz = c.b0u * o.sch;
    bpar = c.b0d * (static_cast<double>(ux) * r.l.x + static_cast<double>(uy) * r.l.y +
                    static_cast<double>(o.sch) * r.l.z);
  }
  o.te_in = geo.te_in;
  if (o.te_in)
    te_d = c.n0d;
  o.cell_ok = false;
  if (GRIDS) {
    Cell cell;
    if (grid_cell_scaled(in, pos, cell, o.cx, o.cy, o.cz)) {
      o.cell_ok = true;
      const float4 *g = static_cast<const float4 *>(in.packed);
      const uint64_t sy = in.geo.nz, sx = in.geo.ny * in.geo.nz;
      float4 w1, w2;
      o.idx = cell.idx;
      o.zd = static_cast<float>(cell.zd);
      o.yd = static_cast<float>(cell.yd);
      o.xd = static_cast<float>(cell.xd);
      lerp_plane_f4(g, cell.idx, sy, o.zd, o.yd, w1);
      lerp_plane_f4(g, cell.idx + sx, sy, o.zd, o.yd, w2);
      const float x1 = 1.f - o.xd;
      const float gx = w1.x * x1 + w2.x * o.xd;
      const float gy = w1.y * x1 + w2.y * o.xd;
      const float gz = w1.z * x1 + w2.z * o.xd;
      bx += gx;
      by += gy;
      bz += gz;
      bpar += static_cast<double>(gx) * r.l.x + static_cast<double>(gy) * r.l.y +
              static_cast<double>(gz) * r.l.z;
      te_d += static_cast<double>(w1.w * x1 + w2.w * o.xd);
    }
  }
  o.te_raw = static_cast<float>(te_d); // the clamp decision keeps te_d's sign
  te_d = te_d > 0. ? te_d : 0.;
  const float te = static_cast<float>(te_d);
  o.te = te;
  o.bx = bx;
  o.by = by;
  o.bz = bz;
  o.b_pard = bpar;
  o.b_par = static_cast<float>(bpar);
  o.fdk = te_d * bpar;
  const float cx = by * lf.z - bz * lf.y, cy = bz * lf.x - bx * lf.z,
              cz = bx * lf.y - by * lf.x;
  o.bper2 = cx * cx + cy * cy + cz * cz;
  float prof;
  if (ph.cre_model == HAMX_CRE_UNIF) {
    prof = geo.cre_in ? 1.f : 0.f;
  } else {
    // sech^2(z/z0) = 4 e^2 / (1 + e^2)^2 with e = exp(-|z|/z0) <= 1
    const float e = __expf(-fabsf(zk) * c.inv_z0k);
    const float e2 = e * e, d = 1.f + e2;
    prof = __expf((c.R0k - rk) * c.inv_r0k) * (4.f * e2) * __builtin_amdgcn_rcpf(d * d);
  }
  if (PW1)
    o.w = prof * o.bper2;
  else
    o.w = prof * exp2f(c.hexp * __log2f(o.bper2));
  o.x = -(ephf.x * bx + ephf.y * by + ephf.z * bz);
  o.y = -(ethf.x * bx + ethf.y * by + ethf.z * bz);
  o.n2 = o.x * o.x + o.y * o.y;
  if (o.n2 > 0.f) {
    const float inv = __builtin_amdgcn_rcpf(o.n2); // 1 ulp: scales Q/U only
    o.c2i = (o.x * o.x - o.y * o.y) * inv;
    o.s2i = 2.f * o.x * o.y * inv;
  } else {
    o.c2i = 1.f;
    o.s2i = 0.f;
  }
  return true;
}

template <bool GRIDS, bool PW1>
__device__ void march_ray_f32(const BInst &in, int s, uint64_t ipix, int lane,
                              const RayGeo &r) {
  // FP32 sample arithmetic: a * b + c may fuse (more accurate, fewer instructions)
  #pragma clang fp contract(fast)
  const BShell &sh = in.sh[s];
  const float3 lf = make_float3(static_cast<float>(r.l.x), static_cast<float>(r.l.y),
                                static_cast<float>(r.l.z));
  const float3 ethf = make_float3(static_cast<float>(r.eth.x), static_cast<float>(r.eth.y),
                                  static_cast<float>(r.eth.z));
  const float3 ephf = make_float3(static_cast<float>(r.eph.x), static_cast<float>(r.eph.y),
                                  static_cast<float>(r.eph.z));
  const int nf = in.n_freq;
  float dm = 0.f, I = 0.f, Q[kMaxF] = {0.f, 0.f, 0.f, 0.f}, U[kMaxF] = {0.f, 0.f, 0.f, 0.f};
  double carry = 0.;
  for (uint64_t base = r.k_lo; base < r.k_hi; base += 64) {
    const uint64_t k = base + lane;
    S32 o;
    bool valid = false;
    if (k < r.k_hi)
      valid = eval_f32<GRIDS, PW1>(in, sh, r, lf, ethf, ephf, k, o);
    const double fk = valid ? o.fdk * sh.fdu : 0.;
    const double incl = wave_incl_scan(fk, lane);
    const double fd_here = carry + incl;
    carry += wave_last(incl);
    if (valid) {
      dm += o.te;
      I += o.w;
#pragma unroll
      for (int f = 0; f < kMaxF; ++f) {
        if (f < nf) {
          const double X = in.fdin[f] ? r.fd_inner + fd_here : fd_here;
          double rv = X * in.rev[f];
          rv -= rint(rv);
          const float a = static_cast<float>(rv);
          const float sn = __builtin_amdgcn_sinf(a), cs = __builtin_amdgcn_cosf(a);
          Q[f] += (cs * o.c2i - sn * o.s2i) * o.w;
          U[f] += (sn * o.c2i + cs * o.s2i) * o.w;
        }
      }
    }
  }
  const double dms = wave_sum(static_cast<double>(dm));
  const double Is = wave_sum(static_cast<double>(I));
  double Qs[kMaxF], Us[kMaxF];
#pragma unroll
  for (int f = 0; f < kMaxF; ++f) {
    if (f < nf) {
      Qs[f] = wave_sum(static_cast<double>(Q[f]));
      Us[f] = wave_sum(static_cast<double>(U[f]));
    }
  }
  if (lane == 0) {
    const double dl = sh.step ? sh.delta : 0.; // no samples: no finite delta
    if (in.do_dm)
      in.t_dm[ipix] = dms * dl;
    in.t_fd[in.fd_off[s] + ipix] = carry;
#pragma unroll
    for (int f = 0; f < kMaxF; ++f) {
      if (f < nf) {
        in.t_i[f][ipix] = Is * dl * in.st[f];
        in.t_q[f][ipix] = Qs[f] * dl * in.sp[f];
        in.t_u[f][ipix] = Us[f] * dl * in.sp[f];
      }
    }
  }
}

// The forward with one ray per lane (64-ray items, NESTED 8x8 patches as in
// the reverse sweep): each lane walks its own ray from the observer outward,
// its FD a lane-private running sum (no wave scan), its map sums in FP64.
// The wavefront steps over the union of its lanes' step ranges.  The plan's
// autotuner times this shape against the lanes-per-sample ones.
template <bool GRIDS, bool PW1>
__device__ void march_lanes_f32(const BInst &in, int s, uint64_t ipix, bool live,
                                const RayGeo &r) {
  // FP32 sample arithmetic: a * b + c may fuse (more accurate, fewer instructions)
  #pragma clang fp contract(fast)
  const BShell &sh = in.sh[s];
  const float3 lf = make_float3(static_cast<float>(r.l.x), static_cast<float>(r.l.y),
                                static_cast<float>(r.l.z));
  const float3 ethf = make_float3(static_cast<float>(r.eth.x), static_cast<float>(r.eth.y),
                                  static_cast<float>(r.eth.z));
  const float3 ephf = make_float3(static_cast<float>(r.eph.x), static_cast<float>(r.eph.y),
                                  static_cast<float>(r.eph.z));
  const int nf = in.n_freq;
  double dm = 0., I = 0., fd = 0., Q[kMaxF] = {0., 0., 0., 0.}, U[kMaxF] = {0., 0., 0., 0.};
  const uint64_t k_end = wave_max_u64(live ? r.k_hi : 0ull);
  const uint64_t k_beg = wave_min_u64(live ? r.k_lo : ~0ull);
  for (uint64_t k = k_beg; k < k_end; ++k) {
    S32 o;
    if (live && k >= r.k_lo && k < r.k_hi && eval_f32<GRIDS, PW1>(in, sh, r, lf, ethf, ephf, k, o)) {
      fd += o.fdk * sh.fdu; // the FD through this sample, itself included
      dm += static_cast<double>(o.te);
      I += static_cast<double>(o.w);
#pragma unroll
      for (int f = 0; f < kMaxF; ++f) {
        if (f < nf) {
          const double X = in.fdin[f] ? r.fd_inner + fd : fd;
          double rv = X * in.rev[f];
          rv -= rint(rv);
          const float a = static_cast<float>(rv);
          const float sn = __builtin_amdgcn_sinf(a), cs = __builtin_amdgcn_cosf(a);
          Q[f] += static_cast<double>((cs * o.c2i - sn * o.s2i) * o.w);
          U[f] += static_cast<double>((sn * o.c2i + cs * o.s2i) * o.w);
        }
      }
    }
  }
  if (live) {
    const double dl = sh.step ? sh.delta : 0.; // no samples: no finite delta
    if (in.do_dm)
      in.t_dm[ipix] = dm * dl;
    in.t_fd[in.fd_off[s] + ipix] = fd;
#pragma unroll
    for (int f = 0; f < kMaxF; ++f) {
      if (f < nf) {
        in.t_i[f][ipix] = I * dl * in.st[f];
        in.t_q[f][ipix] = Q[f] * dl * in.sp[f];
        in.t_u[f][ipix] = U[f] * dl * in.sp[f];
      }
    }
  }
}

// -------------------------------------------------------- reverse mode --
//
// One sweep per ray from its far end (as k_los_adj): the running FD is the
// ray's total from the tape minus the samples beyond, and
//   R_k = sum over passes f and samples j >= k of lambda_f^2 dL/dqui_{f,j}
//         + dL/d(shell FD)
// is carried the same way.  Each pass f adds its own phase term; the
// emissivity shape w is shared, so dL/dw sums the passes' I and Q/U terms.

__device__ inline ShellArgs shell_args_rev(const BInst &in, int s) {
  ShellArgs sa = shell_args(in, s);
  sa.fd_map = (in.do_fd && s > 0) ? in.adj.ckpt + static_cast<uint64_t>(s) * in.hfd.npix
                                  : nullptr;
  return sa;
}

template <bool GRIDS, bool PW1>
__device__ void adj_ray_f64(const BInst &in, int s, uint64_t ipix, int lane, void *slab,
                            const RayGeo &r) {
  const ShellArgs sa = shell_args_rev(in, s);
  const BAdj &A = in.adj;
  Grids gr;
  gr.packed = static_cast<const double4 *>(in.packed);
  gr.g_packed = in.geo;
  const FastConst &fc = in.fc[s];
  const int nf = in.n_freq;
  // per-ray constants: dL/d(pass-0 jt) and the Q / U weights of each pass
  double aJt = 0., AIL = 0., aQ[kMaxF] = {0., 0., 0., 0.}, aU[kMaxF] = {0., 0., 0., 0.};
#pragma unroll
  for (int f = 0; f < kMaxF; ++f) {
    if (f < nf) {
      const double gi = A.g_i[f][ipix] * in.st[f];
      aJt += gi;
      AIL += gi * A.lt[f];
      aQ[f] = A.g_q[f][ipix] * in.sp[f];
      aU[f] = A.g_u[f][ipix] * in.sp[f];
    }
  }
  const double gdm = in.do_dm ? A.g_dm[ipix] : 0.;
  const double gfd = in.do_fd ? A.g_fd[ipix] : 0.;
  const double fd_total = in.t_fd[in.fd_off[s] + ipix];
  const uint64_t sy = in.geo.nz, sx = in.geo.ny * in.geo.nz;
  const double alpha1 = 1. - in.ph.idx;
  double p0 = 0., p1 = 0., p2 = 0., p3 = 0., p4 = 0., p5 = 0., q0 = 0.;
  double after = 0., d_after = 0.;
  const uint64_t n = r.k_hi - r.k_lo;
  for (uint64_t c = (n + 63) / 64; c-- > 0;) {
    const uint64_t k = r.k_lo + 64 * c + lane;
    FastSample o;
    bool valid = false;
    if (k < r.k_hi)
      valid = eval_fast<GRIDS, PW1>(in.ph, fc, gr, sa, r, k, o);
    const double incl = wave_incl_scan(valid ? o.fdk : 0., lane);
    const double chunk = wave_last(incl);
    const double fd_here = (fd_total - after - chunk) + incl;
    after += chunk;
    double D = 0., dJp = 0., alpha_p = 0., dipa = 0.;
    if (valid) {
#pragma unroll
      for (int f = 0; f < kMaxF; ++f) {
        if (f < nf) {
          const double X = in.fdin[f] ? r.fd_inner + fd_here : fd_here;
          double sph, cph;
          sincos(in.two_l2[f] * X, &sph, &cph);
          const double c2 = cph * o.c2i - sph * o.s2i, s2 = sph * o.c2i + cph * o.s2i;
          const double P = aQ[f] * c2 + aU[f] * s2;
          dJp += P;
          alpha_p += P * A.lp[f];
          const double dq = 2. * o.jp * (aU[f] * c2 - aQ[f] * s2);
          dipa += dq;
          D += A.l2[f] * dq;
          if (f == 0)
            q0 += dq;
        }
      }
    }
    const double dscan = wave_suffix_scan(D, lane);
    const double R = dscan + d_after + gfd;
    d_after += __shfl(dscan, 0, 64);
    double dte = 0.;
    Vec3 dB{0., 0., 0.};
    if (valid) {
      dte = o.te_raw > 0. ? R * o.b_par * fc.fd_scale + gdm * sa.delta : 0.;
      const double ft = R * o.te * fc.fd_scale;
      dB = Vec3{ft * r.l.x, ft * r.l.y, ft * r.l.z};
      if (o.bper2 > 0.) {
        const double coef = (aJt * o.jt + dJp * o.jp) * alpha1 / (2. * o.bper2);
        dB.x += coef * (o.B.x - o.b_par * r.l.x);
        dB.y += coef * (o.B.y - o.b_par * r.l.y);
        dB.z += coef * (o.B.z - o.b_par * r.l.z);
        const double lnb = 0.25 * log(o.bper2); // (1/2) ln |B_perp|
        p4 += o.jt * (AIL + aJt * lnb) + o.jp * (alpha_p + dJp * lnb);
      }
      if (o.n2 > 0.) {
        const double cc = dipa / o.n2;
        dB.x += cc * (o.y * r.eph.x - o.x * r.eth.x);
        dB.y += cc * (o.y * r.eph.y - o.x * r.eth.y);
        dB.z += cc * (o.y * r.eph.z - o.x * r.eth.z);
      }
      if (o.te_in)
        p5 += dte;
      if (o.lsa_in) {
        const double u = o.sps * o.cos_p - o.cps * o.sin_p;
        const double v = o.sps * o.sin_p + o.cps * o.cos_p;
        p0 += dB.x * (o.cch * u) + dB.y * (o.cch * v) + dB.z * o.sch;
        const double gpsi = A.b0 * o.cch * (dB.x * v - dB.y * u);
        p1 += gpsi;
        p2 += gpsi * o.lnr;
        const double gchi = A.b0 * (o.cch * dB.z - o.sch * (dB.x * u + dB.y * v));
        p3 += gchi * o.tnh;
      }
    }
    if (GRIDS && A.ggrid) { // wave-cooperative: every lane calls it
      const bool has = valid && o.cell_ok;
      scatter_staged(staged_at(slab), static_cast<double *>(A.ggrid), has,
                     has ? o.cell.idx : 0, o.cell.xd, o.cell.yd, o.cell.zd, dB.x, dB.y, dB.z, dte,
                     sx, sy, lane);
    }
  }
  p0 = wave_sum(p0);
  p1 = wave_sum(p1);
  p2 = wave_sum(p2);
  p3 = wave_sum(p3);
  p4 = wave_sum(p4);
  p5 = wave_sum(p5);
  q0 = wave_sum(q0);
  if (lane == 0) {
    double *rp = A.ray_par + 6 * (in.fd_off[s] + ipix);
    rp[0] = p0;
    rp[1] = p1;
    rp[2] = p2;
    rp[3] = p3;
    rp[4] = p4;
    rp[5] = p5;
    if (sa.fd_map && in.fdin[0])
      A.gfdin[ipix] = A.l2[0] * q0;
  }
}

// ------------------------------------------- reverse mode, lane per ray --
//
// The FP32 reverse sweep with one ray per lane: the running FD (total from
// the tape minus the samples beyond) and R_k are lane-private sums, and the
// lane keeps the 32 trilinear corner sums of the grid cell its ray is in,
// adding them to the gradient grid when the ray leaves the cell (a straight
// ray meets each cell in one run of samples).  No cross-lane reduction,
// LDS atomic or wavefront barrier is left in the sample loop.


// Per-wavefront LDS merge of the lanes' gradient flushes.  The 64 rays of a
// NESTED 8x8 patch leave the same grid vertices behind a few steps apart, so
// a flush goes first to a direct-mapped cache of vertices (tag = vertex
// index, 4 FP32 sums) and reaches the grid only when its slot is taken by
// another vertex or at the end of the work item.  One wavefront owns its
// cache, so LDS program order is the only ordering needed.
//
// A lane leaving a cell hands its vertex sums to the cache through the
// wavefront's flush list: in a sample step only a few lanes leave their
// cell, so the lanes that do append their entries to one dense list, and the
// whole wavefront merges the list into the cache 64 entries at a time once
// it is full (or at the end of the work item).
constexpr int kVtxSlotsLog2 = 9;
constexpr int kVtxSlots = 1 << kVtxSlotsLog2;
constexpr uint32_t kFlushCap = 512; // entries of the flush list
constexpr uint32_t kVtxEmpty = 0xffffffffu;

struct VtxCache {
  uint32_t *tag;  // kVtxSlots tags (kVtxEmpty: free); nullptr: no cache
  float4 *val;    // kVtxSlots sums {x, y, z, te}
  uint32_t *ltag; // kFlushCap flush-list vertices
  float4 *lval;   // kFlushCap flush-list sums
};

constexpr int vtx_cache_bytes() { return (kVtxSlots + kFlushCap) * (4 + 16); }

__device__ inline VtxCache vtx_cache_at(void *base) {
  VtxCache c;
  c.val = static_cast<float4 *>(base);
  c.tag = reinterpret_cast<uint32_t *>(c.val + kVtxSlots);
  c.lval = reinterpret_cast<float4 *>(c.tag + kVtxSlots);
  c.ltag = reinterpret_cast<uint32_t *>(c.lval + kFlushCap);
  return c;
}

__device__ inline void vtx_cache_clear(const VtxCache &c, int lane) {
  for (int i = lane; i < kVtxSlots; i += 64) {
    c.tag[i] = kVtxEmpty;
    c.val[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
  wave_lds_sync();
}

__device__ inline void grid_add4(float *__restrict__ g, uint64_t v, float4 a) {
  float *p = g + 4 * v;
  grid_atomic_add(p + 0, a.x);
  grid_atomic_add(p + 1, a.y);
  grid_atomic_add(p + 2, a.z);
  grid_atomic_add(p + 3, a.w);
}

__device__ inline void lds_add4(float4 *p, float4 a) {
  float *q = reinterpret_cast<float *>(p);
  unsafeAtomicAdd(q + 0, a.x);
  unsafeAtomicAdd(q + 1, a.y);
  unsafeAtomicAdd(q + 2, a.z);
  unsafeAtomicAdd(q + 3, a.w);
}

// Adds `a` to vertex v (< kVtxEmpty) for every active lane.  Lanes that read
// the same slot read the same tag, so among the lanes that miss exactly one
// compare-and-swap succeeds: that lane evicts the slot's old sums to the
// grid and the slot then belongs to its vertex.  Missing lanes of that vertex
// add into the slot, the others go to the grid directly.
__device__ inline void vtx_put(const VtxCache &c, float *__restrict__ g, uint32_t key, float4 a) {
  const uint32_t sl = (key * 0x9e3779b1u) >> (32 - kVtxSlotsLog2);
  const uint32_t t = c.tag[sl];
  const bool hit = t == key;
  if (hit)
    lds_add4(c.val + sl, a);
  wave_lds_sync();
  if (!hit) {
    const bool own = atomicCAS(c.tag + sl, t, key) == t;
    float4 old = make_float4(0.f, 0.f, 0.f, 0.f);
    if (own) {
      old = c.val[sl];
      c.val[sl] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
    wave_lds_sync();
    if (own && t != kVtxEmpty)
      grid_add4(g, t, old);
    if (c.tag[sl] == key)
      lds_add4(c.val + sl, a);
    else
      grid_add4(g, key, a);
  }
  wave_lds_sync();
}

// Every held vertex -> the grid; the cache is left empty.
__device__ inline void vtx_cache_flush(const VtxCache &c, float *__restrict__ g, int lane) {
  wave_lds_sync();
  for (int i = lane; i < kVtxSlots; i += 64) {
    const uint32_t t = c.tag[i];
    if (t != kVtxEmpty) {
      grid_add4(g, t, c.val[i]);
      c.tag[i] = kVtxEmpty;
      c.val[i] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
  }
  wave_lds_sync();
}

// The first n (wave-uniform) flush-list entries -> the vertex cache, one
// entry per lane; called by the whole wavefront.
__device__ inline void flush_list_drain(const VtxCache &c, float *__restrict__ g, uint32_t n,
                                        int lane) {
  wave_lds_sync(); // the entries written by other lanes are visible
  for (uint32_t b = 0; b < n; b += 64) {
    const uint32_t e = b + static_cast<uint32_t>(lane);
    if (e < n)
      vtx_put(c, g, c.ltag[e], c.lval[e]);
  }
  wave_lds_sync(); // read before the list is written again
}

// One vertex sum of a lane that takes part in a list step: to the list at
// `pos`, or straight to the grid when there is no cache (vertex indices
// beyond 32 bits).
__device__ inline void list_put(const VtxCache &c, float *__restrict__ g, uint32_t pos,
                                uint64_t v, float4 a) {
  if (c.tag) {
    c.ltag[pos] = static_cast<uint32_t>(v);
    c.lval[pos] = a;
  } else {
    grid_add4(g, v, a);
  }
}

// Called by the whole wavefront: the lanes with `part` each get `per` flush
// list entries; returns the lane's first entry (meaningful where part).
// The list is drained first when the entries would not fit.
__device__ inline uint32_t list_reserve(const VtxCache &c, float *__restrict__ g, bool part,
                                        uint32_t per, uint32_t &nlist, int lane,
                                        uint64_t &mask) {
  mask = __ballot(part);
  if (mask == 0 || c.tag == nullptr)
    return 0;
  const uint32_t total = per * static_cast<uint32_t>(__popcll(mask));
  if (nlist + total > kFlushCap) {
    flush_list_drain(c, g, nlist, lane);
    nlist = 0;
  }
  const uint32_t pos =
      nlist + per * static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)));
  nlist += total;
  return pos;
}

// acc (8 corners x {x, y, z, te}) of the lanes with `part` -> the flush list
// as the 8 vertices of cell `cell`, then acc = 0 there.  Whole wavefront.
__device__ inline void list_cell_f32(float (&acc)[32], bool part, uint64_t cell,
                                     float *__restrict__ g, uint64_t sx, uint64_t sy,
                                     const VtxCache &vc, uint32_t &nlist, int lane) {
  uint64_t mask;
  const uint32_t pos = list_reserve(vc, g, part, 8, nlist, lane, mask);
  if (mask == 0 || !part)
    return;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const uint64_t v = cell + (c >> 2) * sx + ((c >> 1) & 1) * sy + (c & 1);
    list_put(vc, g, pos + c,
             v, make_float4(acc[4 * c], acc[4 * c + 1], acc[4 * c + 2], acc[4 * c + 3]));
#pragma unroll
    for (int m = 0; m < 4; ++m)
      acc[4 * c + m] = 0.f;
  }
}

// The rays of the lanes with `part` move from `cell` to its neighbour along
// one axis (BIT: 4 x, 2 y, 1 z in the corner numbering 4a + 2b + c),
// upwards when `up`: the 4 corners they leave behind go to the flush list,
// the 4 shared with the new cell move to their place in it, the new cell's
// far corners start at 0, and `cell` becomes the new cell.  Whole wavefront.
template <int BIT>
__device__ inline void list_shift_f32(float (&acc)[32], bool part, bool up, uint64_t &cell,
                                      float *__restrict__ g, uint64_t sx, uint64_t sy,
                                      const VtxCache &vc, uint32_t &nlist, int lane) {
  uint64_t mask;
  const uint32_t pos = list_reserve(vc, g, part, 4, nlist, lane, mask);
  if (mask == 0 || !part)
    return;
  int j = 0;
#pragma unroll
  for (int lo = 0; lo < 8; ++lo) {
    if (lo & BIT)
      continue;
    const int hi = lo | BIT;
    const uint64_t vlo = cell + (lo >> 2) * sx + ((lo >> 1) & 1) * sy + (lo & 1);
    const uint64_t vhi = cell + (hi >> 2) * sx + ((hi >> 1) & 1) * sy + (hi & 1);
    float left[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      left[m] = up ? acc[4 * lo + m] : acc[4 * hi + m];
      const float kept = up ? acc[4 * hi + m] : acc[4 * lo + m];
      acc[4 * lo + m] = up ? kept : 0.f;
      acc[4 * hi + m] = up ? 0.f : kept;
    }
    list_put(vc, g, pos + j, up ? vlo : vhi, make_float4(left[0], left[1], left[2], left[3]));
    ++j;
  }
  const uint64_t st = BIT == 4 ? sx : (BIT == 2 ? sy : 1);
  cell = up ? cell + st : cell - st;
}

template <bool GRIDS, bool PW1>
__device__ void adj_lanes_f32(const BInst &in, int s, uint64_t ipix, bool live,
                              const RayGeo &r, VtxCache vc) {
  // FP32 sample arithmetic: a * b + c may fuse (more accurate, fewer instructions)
  #pragma clang fp contract(fast)
  const bool has_cache = in.do_fd && s > 0;
  const BShell &sh = in.sh[s];
  const BAdj &A = in.adj;
  const float3 lf = make_float3(static_cast<float>(r.l.x), static_cast<float>(r.l.y),
                                static_cast<float>(r.l.z));
  const float3 ethf = make_float3(static_cast<float>(r.eth.x), static_cast<float>(r.eth.y),
                                  static_cast<float>(r.eth.z));
  const float3 ephf = make_float3(static_cast<float>(r.eph.x), static_cast<float>(r.eph.y),
                                  static_cast<float>(r.eph.z));
  const int nf = in.n_freq;
  const double delta = sh.delta;
  double aI = 0., AIL = 0.;
  float aQ[kMaxF] = {0.f, 0.f, 0.f, 0.f}, aU[kMaxF] = {0.f, 0.f, 0.f, 0.f};
  float lpf[kMaxF] = {0.f, 0.f, 0.f, 0.f};
  double gdm = 0., gfd = 0., fd_total = 0.;
  if (live) {
#pragma unroll
    for (int f = 0; f < kMaxF; ++f) {
      if (f < nf) {
        const double gi = A.g_i[f][ipix] * in.st[f] * delta;
        aI += gi;
        AIL += gi * A.lt[f];
        aQ[f] = static_cast<float>(A.g_q[f][ipix] * in.sp[f] * delta);
        aU[f] = static_cast<float>(A.g_u[f][ipix] * in.sp[f] * delta);
        lpf[f] = static_cast<float>(A.lp[f]);
      }
    }
    gdm = in.do_dm ? A.g_dm[ipix] : 0.;
    gfd = in.do_fd ? A.g_fd[ipix] : 0.;
    fd_total = in.t_fd[in.fd_off[s] + ipix];
  }
  const uint64_t sy = in.geo.nz, sx = in.geo.ny * in.geo.nz;
  float *ggrid = static_cast<float *>(A.ggrid);
  // vertex indices are the cache's 32-bit tags
  if ((in.geo.fx + 1.) * static_cast<double>(sx) >= static_cast<double>(kVtxEmpty))
    vc.tag = nullptr;
  double p0 = 0., p1 = 0., p2 = 0., p3 = 0., p4 = 0., p5 = 0., q0 = 0.;
  double after = 0., dacc = 0.; // FD and R contributions of the samples beyond
  float acc[32];
#pragma unroll
  for (int i = 0; i < 32; ++i)
    acc[i] = 0.f;
  uint64_t cur = ~0ull; // cell whose corner sums acc holds
  int ccx = 0, ccy = 0, ccz = 0;
  // the union of the lanes' step ranges, far end first
  const uint64_t k_end = wave_max_u64(live ? r.k_hi : 0ull);
  const uint64_t k_beg = wave_min_u64(live ? r.k_lo : ~0ull);
  const int lane = static_cast<int>(threadIdx.x & 63);
  uint32_t nlist = 0; // entries in the wavefront's flush list
  for (uint64_t k = k_end; k-- > k_beg;) {
    S32 o;
    bool valid = false;
    if (live && k >= r.k_lo && k < r.k_hi)
      valid = eval_f32<GRIDS, PW1>(in, sh, r, lf, ethf, ephf, k, o);
    float dbx = 0.f, dby = 0.f, dbz = 0.f;
    double dte = 0.;
    if (valid) {
      const double fd_here = fd_total - after; // FD through this sample
      after += o.fdk * sh.fdu;
      double D = 0.;
      float dLdw = static_cast<float>(aI), alpha_p = 0.f, dipa = 0.f;
#pragma unroll
      for (int f = 0; f < kMaxF; ++f) {
        if (f < nf) {
          const double X = in.fdin[f] ? r.fd_inner + fd_here : fd_here;
          double rv = X * in.rev[f];
          rv -= rint(rv);
          const float a = static_cast<float>(rv);
          const float sn = __builtin_amdgcn_sinf(a), cs = __builtin_amdgcn_cosf(a);
          const float c2 = cs * o.c2i - sn * o.s2i, s2 = sn * o.c2i + cs * o.s2i;
          const float P = aQ[f] * c2 + aU[f] * s2;
          dLdw += P;
          alpha_p += P * lpf[f];
          const float dq = 2.f * o.w * (aU[f] * c2 - aQ[f] * s2);
          dipa += dq;
          D += A.l2[f] * static_cast<double>(dq);
          if (f == 0)
            q0 += dq;
        }
      }
      dacc += D;
      const double R = dacc + gfd; // dL/d(FD contribution of this sample)
      dte = o.te_raw > 0.f ? R * o.b_pard * sh.fdu + gdm * delta : 0.;
      const float ft = static_cast<float>(R * static_cast<double>(o.te) * sh.fdu);
      dbx = ft * lf.x;
      dby = ft * lf.y;
      dbz = ft * lf.z;
      if (o.bper2 > 0.f) {
        const float coef = dLdw * 2.f * in.f.hexp * o.w * __builtin_amdgcn_rcpf(o.bper2);
        dbx += coef * (o.bx - o.b_par * lf.x);
        dby += coef * (o.by - o.b_par * lf.y);
        dbz += coef * (o.bz - o.b_par * lf.z);
        // ln |B_perp| in Gauss: the B grid and b0 are in muG here
        const double lnb = 0.5 * static_cast<double>(__logf(o.bper2)) - 13.815510557964274;
        p4 += static_cast<double>(o.w) *
              (AIL + static_cast<double>(alpha_p) + 0.5 * lnb * static_cast<double>(dLdw));
      }
      if (o.n2 > 0.f) {
        const float cc = dipa * __builtin_amdgcn_rcpf(o.n2);
        dbx += cc * (o.y * ephf.x - o.x * ethf.x);
        dby += cc * (o.y * ephf.y - o.x * ethf.y);
        dbz += cc * (o.y * ephf.z - o.x * ethf.z);
      }
      if (o.te_in)
        p5 += dte;
      if (o.lsa_in) {
        const float u = o.sps * o.cos_p - o.cps * o.sin_p;
        const float v = o.sps * o.sin_p + o.cps * o.cos_p;
        p0 += 1e6 * static_cast<double>(dbx * (o.cch * u) + dby * (o.cch * v) + dbz * o.sch);
        const float gpsi = in.f.b0u * o.cch * (dbx * v - dby * u);
        p1 += gpsi;
        p2 += gpsi * o.lnr;
        const float gchi = in.f.b0u * (o.cch * dbz - o.sch * (dbx * u + dby * v));
        p3 += gchi * o.tnh;
      }
    }
    if (GRIDS && ggrid) {
      // the whole wavefront: lanes whose ray leaves its cell hand on the
      // corner sums of the cell (face steps one axis at a time: every
      // intermediate cell lies between two valid cells, so is valid)
      const bool in_cell = valid && o.cell_ok;
      const bool moves = in_cell && cur != ~0ull && o.idx != cur;
      int dx = 0, dy = 0, dz = 0;
      if (moves) {
        dx = o.cx - ccx;
        dy = o.cy - ccy;
        dz = o.cz - ccz;
      }
      const bool near = dx >= -1 && dx <= 1 && dy >= -1 && dy <= 1 && dz >= -1 && dz <= 1;
      uint64_t c = cur;
      list_shift_f32<4>(acc, moves && near && dx != 0, dx > 0, c, ggrid, sx, sy, vc, nlist,
                        lane);
      list_shift_f32<2>(acc, moves && near && dy != 0, dy > 0, c, ggrid, sx, sy, vc, nlist,
                        lane);
      list_shift_f32<1>(acc, moves && near && dz != 0, dz > 0, c, ggrid, sx, sy, vc, nlist,
                        lane);
      list_cell_f32(acc, moves && !near, cur, ggrid, sx, sy, vc, nlist, lane);
      if (in_cell) {
        if (o.idx != cur) {
          cur = o.idx;
          ccx = o.cx;
          ccy = o.cy;
          ccz = o.cz;
        }
        const float gte = static_cast<float>(dte);
        const float wx[2] = {1.f - o.xd, o.xd}, wy[2] = {1.f - o.yd, o.yd},
                    wz[2] = {1.f - o.zd, o.zd};
#pragma unroll
        for (int cn = 0; cn < 8; ++cn) {
          const float w = wx[cn >> 2] * wy[(cn >> 1) & 1] * wz[cn & 1];
          acc[4 * cn + 0] += w * dbx;
          acc[4 * cn + 1] += w * dby;
          acc[4 * cn + 2] += w * dbz;
          acc[4 * cn + 3] += w * gte;
        }
      }
    }
  }
  if (GRIDS && ggrid) {
    list_cell_f32(acc, cur != ~0ull, cur, ggrid, sx, sy, vc, nlist, lane);
    if (vc.tag) {
      flush_list_drain(vc, ggrid, nlist, lane);
      vtx_cache_flush(vc, ggrid, lane);
    }
  }
  if (live) {
    double *rp = A.ray_par + 6 * (in.fd_off[s] + ipix);
    rp[0] = p0;
    rp[1] = p1;
    rp[2] = p2;
    rp[3] = p3;
    rp[4] = p4;
    rp[5] = p5;
    if (has_cache && in.fdin[0])
      A.gfdin[ipix] = A.l2[0] * q0;
  }
}

// ---------------------------------------------------------- work items --

template <int PREC, bool GRIDS, bool PW1, bool REV>
__device__ void do_item(const LevelArgs &a, uint64_t item, int lane, void *slab, RayGeo *rays) {
  uint32_t lo = 0, hi = a.n_ids;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a.prefix[mid] <= item)
      lo = mid;
    else
      hi = mid;
  }
  const uint32_t id = __builtin_amdgcn_readfirstlane(a.ids[lo]);
  // The instance records are written by the host before the launch and only
  // read here: reading them through the constant address space lets the
  // compiler keep their (wave-uniform) fields in scalar loads and registers
  // instead of reloading them with vector loads in every sample step, where
  // each reload would also wait for the step's gradient atomics.
  using ConstInst = const __attribute__((address_space(4))) BInst;
  const BInst &in = *(const BInst *)((ConstInst *)a.inst + id);
  const int s = static_cast<int>(a.level);
  const uint64_t npix = in.sh[s].hs.npix;
  const uint64_t p0 = (item - a.prefix[lo]) * a.rays;
  if constexpr (!REV && PREC == HAMX_PREC_F32) {
    if (a.rays == 64) { // one ray per lane, as the reverse sweep below
      const Healpix &hs = in.sh[s].hs;
      const uint64_t pl = p0 + static_cast<uint64_t>(lane);
      const bool live = pl < npix;
      const uint64_t ipix = live ? hp_nest2ring(hs, pl) : 0;
      RayGeo rs{};
      if (live)
        ray_setup_call(in.ph, shell_args(in, s), ipix, rs);
      const RayGeo r = rs;
      march_lanes_f32<GRIDS, PW1>(in, s, ipix, live, r);
      return;
    }
  }
  if constexpr (REV && PREC == HAMX_PREC_F32) {
    // one ray per lane (a.rays == 64): a NESTED 8x8 patch of the shell (the
    // lanes' samples at one step then lie within a few grid cells; shell
    // nsides are powers of two, validate_los)
    const Healpix &hs = in.sh[s].hs;
    const uint64_t pl = p0 + static_cast<uint64_t>(lane);
    const bool live = pl < npix;
    const uint64_t ipix = live ? hp_nest2ring(hs, pl) : 0;
    RayGeo rs{};
    if (live)
      ray_setup_call(in.ph, shell_args_rev(in, s), ipix, rs);
    const RayGeo r = rs; // a copy that stays in registers (rs went to the call's stack)
    adj_lanes_f32<GRIDS, PW1>(in, s, ipix, live, r, vtx_cache_at(slab));
    return;
  }
  const uint64_t p1 = (p0 + a.rays < npix) ? p0 + a.rays : npix;
  // the item's rays are set up one per lane (pointing, versors, FD cache,
  // culled step range: scalar FP64 work that every lane would otherwise
  // repeat) into the wavefront's LDS slots, then marched one after another
  const uint64_t nr = p1 - p0; // <= kItemRaysMax (plan)
  if (static_cast<uint64_t>(lane) < nr) {
    const ShellArgs sa = REV ? shell_args_rev(in, s) : shell_args(in, s);
    ray_setup_call(in.ph, sa, p0 + lane, rays[lane]);
  }
  // (from here on REV implies FP64: the FP32 reverse returned above)
  const bool cached = REV && GRIDS && in.adj.ggrid;
  if (cached) // the item's rays share one cell cache, flushed at the end
    staged_init(staged_at(slab), lane);
  wave_lds_sync();
  for (uint64_t j = 0; j < nr; ++j) {
    const RayGeo r = rays[j];
    const uint64_t p = p0 + j;
    if constexpr (REV)
      adj_ray_f64<GRIDS, PW1>(in, s, p, lane, slab, r);
    else if constexpr (PREC == HAMX_PREC_F32)
      march_ray_f32<GRIDS, PW1>(in, s, p, lane, r);
    else
      march_ray_f64<GRIDS, PW1>(in, s, p, lane, r);
  }
  if (cached)
    staged_flush_all(staged_at(slab), static_cast<double *>(in.adj.ggrid),
                     in.geo.ny * in.geo.nz, in.geo.nz, lane);
  wave_lds_sync(); // the slots are rewritten by the next item
}

// One launch covers the instances of one shell level that share the
// template case (random grids present, emissivity exponent -1); REV
// selects the reverse-mode sweep.
template <int PREC, bool GRIDS, bool PW1, bool REV>
__global__ __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(PREC == HAMX_PREC_F32 && !REV ? 3 : 1))) void k_batch_march(
    LevelArgs a) {
  const int lane = threadIdx.x & 63;
  const uint32_t wave = threadIdx.x >> 6, wpb = blockDim.x >> 6;
  __shared__ RayGeo ray_slots[4][kItemRaysMax];
  RayGeo *rays = ray_slots[wave];
  void *slab = nullptr; // reverse mode: one LDS cell cache per wavefront
  if constexpr (REV && GRIDS && PREC == HAMX_PREC_F64) {
    __shared__ __align__(16) unsigned char caches[4][staged_cache_bytes()];
    slab = caches[wave];
  }
  if constexpr (REV && GRIDS && PREC == HAMX_PREC_F32) {
    __shared__ __align__(16) unsigned char vcaches[4][vtx_cache_bytes()];
    slab = vcaches[wave];
    vtx_cache_clear(vtx_cache_at(slab), lane);
  }
  if (!a.persistent) {
    uint64_t b = blockIdx.x;
    if (a.xcd) { // gridDim.x is a multiple of kQueues: b -> XCD-contiguous order
      const uint64_t per = gridDim.x / kQueues;
      b = (b % kQueues) * per + b / kQueues;
    }
    const uint64_t item = b * wpb + wave;
    if (item < a.n_items)
      do_item<PREC, GRIDS, PW1, REV>(a, uniform_u64(item), lane, slab, rays);
    return;
  }
  const uint32_t q0 = a.xcd ? (blockIdx.x % a.nq) : 0;
  for (uint32_t qi = 0; qi < a.nq; ++qi) {
    const uint32_t q = (q0 + qi) % a.nq;
    for (;;) {
      unsigned long long t = 0;
      if (lane == 0)
        t = atomicAdd(&a.ctr[q], 1ull);
      t = uniform_u64(__shfl(t, 0, 64));
      const uint64_t item = a.qbeg[q] + t;
      if (item >= a.qend[q])
        break; // every wave leaves once all queues are drained
      do_item<PREC, GRIDS, PW1, REV>(a, item, lane, slab, rays);
    }
  }
}

// ---------------------------------------------------------- accumulate --

struct AccJob {
  Healpix hd, hs;
  int n, pad;
  double *dst[kAccMaps];
  const double *src[kAccMaps];
  // transposed interpolation (k_batch_acc_t), CSR by pixel of the nside of
  // hs: t_row[j]..t_row[j+1] index the pixels t_col of the nside of hd that
  // interpolate from j, with coefficients t_coef = w / (sum of their 4 w);
  // NULL: scatter with atomics
  const uint32_t *t_row, *t_col;
  const double *t_coef;
};

// dst += src (same nside) or dst += interpolate(src) at dst's pixel
// centres (hamdis.h:698-721), as k_accumulate; one job per blockIdx.y.
__global__ __launch_bounds__(256) void k_batch_acc(const AccJob *jobs) {
  const AccJob &J = jobs[blockIdx.y];
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= J.hd.npix)
    return;
  if (J.hd.npix == J.hs.npix) {
    for (int m = 0; m < J.n; ++m)
      J.dst[m][i] = J.dst[m][i] + J.src[m][i];
    return;
  }
  double theta, phi;
  hp_pix2ang(J.hd, i, theta, phi);
  uint64_t pix[4];
  double wght[4];
  hp_interpol_weights(J.hs, theta, phi, pix, wght);
  for (int m = 0; m < J.n; ++m) {
    double wtot = 0.0, res = 0.0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double val = J.src[m][pix[q]];
      if (val > cgs::undef_threshold) {
        res += val * wght[q];
        wtot += wght[q];
      }
    }
    const double v = (wtot == 0.0) ? cgs::undef : res / wtot;
    J.dst[m][i] = J.dst[m][i] + v;
  }
}

// dst (shell adjoint, nside_s) += A^T(src) (output adjoint): the transpose
// of k_batch_acc.  Interpolating jobs with the transposed matrix gather per
// shell pixel (each row of the transpose holds the output pixels that
// interpolate from it, so no two threads write one value); without it they
// scatter with atomics.  One job per blockIdx.y, whose records are read
// through the constant address space (written before the launch).
__global__ __launch_bounds__(256) void k_batch_acc_t(const AccJob *jobs) {
  using ConstJob = const __attribute__((address_space(4))) AccJob;
  const AccJob &J = *(const AccJob *)((ConstJob *)jobs + blockIdx.y);
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (J.hd.npix == J.hs.npix) {
    if (i >= J.hd.npix)
      return;
    for (int m = 0; m < J.n; ++m)
      J.dst[m][i] += J.src[m][i];
    return;
  }
  if (J.t_row) {
    if (i >= J.hs.npix)
      return;
    double acc[kAccMaps];
#pragma unroll
    for (int m = 0; m < kAccMaps; ++m)
      acc[m] = 0.;
    const uint32_t e1 = J.t_row[i + 1];
    for (uint32_t e = J.t_row[i]; e < e1; ++e) {
      const double c = J.t_coef[e];
      const uint32_t q = J.t_col[e];
#pragma unroll
      for (int m = 0; m < kAccMaps; ++m)
        if (m < J.n)
          acc[m] += c * J.src[m][q];
    }
#pragma unroll
    for (int m = 0; m < kAccMaps; ++m)
      if (m < J.n)
        J.dst[m][i] += acc[m];
    return;
  }
  if (i >= J.hd.npix)
    return;
  double theta, phi;
  hp_pix2ang(J.hd, i, theta, phi);
  uint64_t pix[4];
  double w[4];
  hp_interpol_weights(J.hs, theta, phi, pix, w);
  const double wtot = ((w[0] + w[1]) + w[2]) + w[3];
  for (int m = 0; m < J.n; ++m) {
    const double g = J.src[m][i] / wtot;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      unsafeAtomicAdd(&J.dst[m][pix[q]], w[q] * g);
  }
}

// The interpolation of k_batch_acc from nside(hs) to every pixel of
// nside(hd): the 4 source pixels and w / (sum of the 4 w) per pixel.
__global__ __launch_bounds__(256) void k_interp_coef(Healpix hd, Healpix hs, uint32_t *pix,
                                                     double *coef) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= hd.npix)
    return;
  double theta, phi;
  hp_pix2ang(hd, i, theta, phi);
  uint64_t p[4];
  double w[4];
  hp_interpol_weights(hs, theta, phi, p, w);
  const double wtot = ((w[0] + w[1]) + w[2]) + w[3];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    pix[4 * i + q] = static_cast<uint32_t>(p[q]);
    coef[4 * i + q] = w[q] / wtot;
  }
}

// G (FD-map adjoint) += Interp_s^T(gfdin): the transpose of the FD-cache
// read of level s.
struct FdtJob {
  double *G;
  const double *gfdin;
  Healpix hfd, hs;
};

__global__ __launch_bounds__(256) void k_batch_fdcache_t(const FdtJob *jobs) {
  const FdtJob &J = jobs[blockIdx.y];
  const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= J.hs.npix)
    return;
  const double g = J.gfdin[p];
  if (g == 0.)
    return;
  double theta, phi;
  hp_pix2ang(J.hs, p, theta, phi);
  uint64_t pix[4];
  double w[4];
  hp_interpol_weights(J.hfd, theta, phi, pix, w);
  const double wtot = ((w[0] + w[1]) + w[2]) + w[3];
#pragma unroll
  for (int q = 0; q < 4; ++q)
    unsafeAtomicAdd(&J.G[pix[q]], w[q] * (g / wtot));
}

// Per-instance sums of the 6 per-ray parameter partials into out[0..5].
struct RedJob {
  const double *ray_par;
  uint64_t n; // rays
  double *out;
};

__global__ __launch_bounds__(256) void k_batch_reduce(const RedJob *jobs) {
  __shared__ double sm[6][4];
  const RedJob &J = jobs[blockIdx.y];
  double v[6] = {0., 0., 0., 0., 0., 0.};
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < J.n;
       i += stride)
#pragma unroll
    for (int j = 0; j < 6; ++j)
      v[j] += J.ray_par[6 * i + j];
#pragma unroll
  for (int j = 0; j < 6; ++j) {
    v[j] = wave_sum(v[j]);
    if ((threadIdx.x & 63) == 0)
      sm[j][threadIdx.x >> 6] = v[j];
  }
  __syncthreads();
  if (threadIdx.x < 6)
    unsafeAtomicAdd(&J.out[threadIdx.x], ((sm[threadIdx.x][0] + sm[threadIdx.x][1]) +
                                          sm[threadIdx.x][2]) + sm[threadIdx.x][3]);
}

// Packed {bx, by, bz, te} gradient -> the caller's SoA arrays (written, not
// added); bscale converts a muG-unit FP32 gradient to per-Gauss.
struct UnpackJob {
  const void *g;
  double *x, *y, *z, *t;
  uint64_t n;
  double bscale;
  int f32, pad;
};

__global__ __launch_bounds__(256) void k_batch_unpack(const UnpackJob *jobs) {
  const UnpackJob &J = jobs[blockIdx.y];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < J.n;
       i += stride) {
    double gx, gy, gz, gt;
    if (J.f32) {
      const float4 v = static_cast<const float4 *>(J.g)[i];
      gx = v.x;
      gy = v.y;
      gz = v.z;
      gt = v.w;
    } else {
      const double4 v = static_cast<const double4 *>(J.g)[i];
      gx = v.x;
      gy = v.y;
      gz = v.z;
      gt = v.w;
    }
    if (J.x)
      J.x[i] = gx * J.bscale;
    if (J.y)
      J.y[i] = gy * J.bscale;
    if (J.z)
      J.z[i] = gz * J.bscale;
    if (J.t)
      J.t[i] = gt;
  }
}

struct CopyJob {
  double *dst;
  const double *src;
  uint64_t n;
};

__global__ __launch_bounds__(256) void k_batch_copy(const CopyJob *jobs) {
  const CopyJob &J = jobs[blockIdx.y];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < J.n;
       i += stride)
    J.dst[i] = J.src[i];
}

struct FillJob {
  double *p;
  uint64_t n;
};

__global__ __launch_bounds__(256) void k_batch_zero(const FillJob *jobs,
                                                    unsigned long long *ctr, uint32_t nctr) {
  const FillJob &J = jobs[blockIdx.y];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < J.n;
       i += stride)
    J.p[i] = 0.;
  if (blockIdx.x == 0 && blockIdx.y == 0)
    for (uint32_t i = threadIdx.x; i < nctr; i += blockDim.x)
      ctr[i] = 0ull;
}

struct PackJob {
  void *dst;
  const double *bx, *by, *bz, *te;
  uint64_t n;
  double bscale; // 1 (FP64 records) or 1e6 (FP32 records in muG)
  int f32, pad;
};

__global__ __launch_bounds__(256) void k_batch_pack(const PackJob *jobs) {
  const PackJob &J = jobs[blockIdx.y];
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < J.n;
       i += stride) {
    const double x = J.bx ? J.bx[i] * J.bscale : 0., y = J.by ? J.by[i] * J.bscale : 0.;
    const double z = J.bz ? J.bz[i] * J.bscale : 0., t = J.te ? J.te[i] : 0.;
    if (J.f32)
      static_cast<float4 *>(J.dst)[i] =
          make_float4(static_cast<float>(x), static_cast<float>(y), static_cast<float>(z),
                      static_cast<float>(t));
    else
      static_cast<double4 *>(J.dst)[i] = make_double4(x, y, z, t);
  }
}

// ----------------------------------------------------------------- host --

using MarchFn = void (*)(LevelArgs);

// Kernel of a launch group: case c = 2 * grids + pw1, forward or reverse.
MarchFn march_kernel(uint32_t prec, int c, bool rev = false) {
  static const MarchFn f32[2][4] = {
      {k_batch_march<HAMX_PREC_F32, false, false, false>,
       k_batch_march<HAMX_PREC_F32, false, true, false>,
       k_batch_march<HAMX_PREC_F32, true, false, false>,
       k_batch_march<HAMX_PREC_F32, true, true, false>},
      {k_batch_march<HAMX_PREC_F32, false, false, true>,
       k_batch_march<HAMX_PREC_F32, false, true, true>,
       k_batch_march<HAMX_PREC_F32, true, false, true>,
       k_batch_march<HAMX_PREC_F32, true, true, true>}};
  static const MarchFn f64[2][4] = {
      {k_batch_march<HAMX_PREC_F64, false, false, false>,
       k_batch_march<HAMX_PREC_F64, false, true, false>,
       k_batch_march<HAMX_PREC_F64, true, false, false>,
       k_batch_march<HAMX_PREC_F64, true, true, false>},
      {k_batch_march<HAMX_PREC_F64, false, false, true>,
       k_batch_march<HAMX_PREC_F64, false, true, true>,
       k_batch_march<HAMX_PREC_F64, true, false, true>,
       k_batch_march<HAMX_PREC_F64, true, true, true>}};
  return prec == HAMX_PREC_F32 ? f32[rev][c] : f64[rev][c];
}

struct Variant {
  int persistent, xcd;
  uint32_t rays, waves_per_block;
  const char *name;
};

const Variant kVariants[] = {
    {0, 0, 4, 4, "one-shot, 4 rays/item, 4 waves/WG"},
    {0, 1, 4, 4, "one-shot XCD-ordered, 4 rays/item, 4 waves/WG"},
    {0, 0, 16, 4, "one-shot, 16 rays/item, 4 waves/WG"},
    {0, 1, 16, 4, "one-shot XCD-ordered, 16 rays/item, 4 waves/WG"},
    {0, 1, 16, 1, "one-shot XCD-ordered, 16 rays/item, 1 wave/WG"},
    {1, 0, 16, 4, "persistent global queue, 16 rays/item, 4 waves/WG"},
    {1, 1, 4, 4, "persistent XCD queues, 4 rays/item, 4 waves/WG"},
    {1, 1, 16, 4, "persistent XCD queues, 16 rays/item, 4 waves/WG"},
    // FP32 only: one ray per lane (64-ray items, NESTED 8x8 patches)
    {1, 1, 64, 4, "persistent XCD queues, 64 rays/item (lane per ray)"},
    {0, 1, 64, 4, "one-shot XCD-ordered, 64 rays/item (lane per ray)"},
};
constexpr uint32_t kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

// The FP32 reverse sweep runs one ray per lane: 64-ray items (NESTED 8x8
// patches), persistent XCD queues.
const Variant kLaneRev = {1, 1, 64, 4, "persistent XCD queues, 64 rays/item (lane per ray)"};

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
};

} // namespace
} // namespace hamx

using namespace hamx;

struct hamx_batch {
  hamx_ctx *ctx = nullptr;
  uint32_t prec = 0;
  std::vector<BInst> host_inst;
  std::vector<void *> allocs; // every device allocation of the plan
  BInst *d_inst = nullptr;
  uint32_t levels = 0;
  // per level: instance ids, item prefix (host copies for planning)
  // per launch group g = level * 4 + case (case = 2 * grids + pw1)
  std::vector<std::vector<uint32_t>> ids;
  std::vector<std::vector<uint64_t>> steps; // step of each instance at the level
  std::vector<uint32_t *> d_ids;
  std::vector<AccJob *> d_acc;
  std::vector<uint32_t> n_acc;
  std::vector<uint64_t> acc_max_npix;
  FillJob *d_fill = nullptr;
  uint32_t n_fill = 0;
  uint64_t fill_max = 0;
  PackJob *d_pack = nullptr;
  uint32_t n_pack = 0;
  uint64_t pack_max = 0;
  unsigned long long *d_ctr = nullptr;
  // per variant and level: prefix arrays depend on rays per item
  std::vector<std::vector<uint64_t *>> d_prefix; // [variant][group]
  std::vector<LevelArgs> rev_largs;               // [group]: FP32 reverse plan (kLaneRev)
  std::vector<std::vector<LevelArgs>> largs;     // [variant][group]
  int n_cu = 0;
  int per_cu[4][5] = {}; // resident march workgroups per CU, by case and waves per WG
  uint32_t variant = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipGraph_t adj_graph = nullptr; // reverse mode (HAMX_BATCH_ADJOINT)
  hipGraphExec_t adj_exec = nullptr;
  // reverse-mode job lists
  bool adjoint = false;
  int per_cu_rev[4][5] = {};
  unsigned long long *d_actr = nullptr; // tickets of the reverse launches
  struct Jobs {
    void *d = nullptr;
    uint32_t n = 0;
    uint64_t max = 0; // largest job, elements
  };
  Jobs a_init_fill, a_init_copy, a_red, a_unpack;
  std::vector<Jobs> a_zero, a_acct, a_fdt, f_ckpt; // per level
  // transposed interpolations by (nside of the output, nside of the shell):
  // CSR row pointers, columns and coefficients (see AccJob)
  struct Transpose {
    const uint32_t *row, *col;
    const double *coef;
  };
  std::map<std::pair<uint64_t, uint64_t>, Transpose> transposes;
  hamx_batch_info info{};
};

namespace {

int dev_alloc(hamx_batch *b, size_t bytes, void **p) {
  *p = nullptr;
  if (bytes == 0)
    bytes = 8;
  if (hipMalloc(p, bytes) != hipSuccess) {
    *p = nullptr;
    return set_err(b->ctx, HAMX_ERR_NOMEM, "batch: device allocation of %zu bytes failed",
                   bytes);
  }
  b->allocs.push_back(*p);
  return HAMX_OK;
}

template <typename T>
int upload(hamx_batch *b, const std::vector<T> &v, T **dptr) {
  void *p;
  HAMX_TRY(dev_alloc(b, v.size() * sizeof(T), &p));
  if (!v.empty())
    HAMX_HIP(b->ctx, hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *dptr = static_cast<T *>(p);
  return HAMX_OK;
}

// The transpose of k_batch_acc's interpolation from nside(hs) to nside(hd),
// as CSR by pixel of nside(hs) with the output pixels in increasing order
// (so every replay sums in the same order), built once per nside pair of the
// plan: the 4 pixels and coefficients per output pixel on the device, the
// counting sort on the host.  Output maps beyond 2^26 pixels (nside > 2048)
// keep the atomic scatter.
int interp_transpose(hamx_batch *b, const Healpix &hd, const Healpix &hs, AccJob &j) {
  j.t_row = j.t_col = nullptr;
  j.t_coef = nullptr;
  if (hd.npix == hs.npix || hd.npix > (1ull << 26))
    return HAMX_OK;
  const auto key = std::make_pair(hd.nside, hs.nside);
  auto it = b->transposes.find(key);
  if (it == b->transposes.end()) {
    hamx_ctx *ctx = b->ctx;
    const uint64_t nd = hd.npix, ns = hs.npix;
    uint32_t *dpix = nullptr;
    double *dcoef = nullptr;
    if (hipMalloc(&dpix, 4 * nd * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&dcoef, 4 * nd * sizeof(double)) != hipSuccess) {
      (void)hipFree(dpix);
      return set_err(ctx, HAMX_ERR_NOMEM, "batch: transposed interpolation buffers");
    }
    hipLaunchKernelGGL(k_interp_coef, dim3(static_cast<unsigned>((nd + 255) / 256)), dim3(256),
                       0, ctx->stream, hd, hs, dpix, dcoef);
    std::vector<uint32_t> pix(4 * nd);
    std::vector<double> coef(4 * nd);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
      e = hipStreamSynchronize(ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpy(pix.data(), dpix, pix.size() * sizeof(uint32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess)
      e = hipMemcpy(coef.data(), dcoef, coef.size() * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(dpix);
    (void)hipFree(dcoef);
    HAMX_HIP(ctx, e);
    std::vector<uint32_t> row(ns + 1, 0), col(4 * nd);
    std::vector<double> val(4 * nd);
    for (uint64_t k = 0; k < 4 * nd; ++k) {
      if (pix[k] >= ns)
        return set_err(ctx, HAMX_ERR_DEVICE, "batch: interpolation pixel out of range");
      ++row[pix[k] + 1];
    }
    for (uint64_t q = 0; q < ns; ++q)
      row[q + 1] += row[q];
    std::vector<uint32_t> fill(row.begin(), row.end() - 1);
    for (uint64_t k = 0; k < 4 * nd; ++k) {
      const uint32_t at = fill[pix[k]]++;
      col[at] = static_cast<uint32_t>(k / 4);
      val[at] = coef[k];
    }
    hamx_batch::Transpose t;
    uint32_t *dr, *dc;
    double *dv;
    HAMX_TRY(upload(b, row, &dr));
    HAMX_TRY(upload(b, col, &dc));
    HAMX_TRY(upload(b, val, &dv));
    t.row = dr;
    t.col = dc;
    t.coef = dv;
    it = b->transposes.emplace(key, t).first;
  }
  j.t_row = it->second.row;
  j.t_col = it->second.col;
  j.t_coef = it->second.coef;
  return HAMX_OK;
}

void release(hamx_batch *b) {
  if (b->exec)
    (void)hipGraphExecDestroy(b->exec);
  if (b->graph)
    (void)hipGraphDestroy(b->graph);
  if (b->adj_exec)
    (void)hipGraphExecDestroy(b->adj_exec);
  if (b->adj_graph)
    (void)hipGraphDestroy(b->adj_graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->adj_exec = nullptr;
  b->adj_graph = nullptr;
  for (void *p : b->allocs)
    (void)hipFree(p);
  b->allocs.clear();
}

// Items of `level` for `rays` rays per item, split into kQueues contiguous
// slices of about equal cost (cost of an item = rays * steps of its shell).
void plan_level(const hamx_batch *b, uint32_t group, uint32_t rays, std::vector<uint64_t> &prefix,
                LevelArgs &la) {
  const uint32_t level = group / 4;
  const auto &ids = b->ids[group];
  prefix.assign(ids.size() + 1, 0);
  std::vector<double> cost_prefix(ids.size() + 1, 0.);
  for (size_t j = 0; j < ids.size(); ++j) {
    const BInst &in = b->host_inst[ids[j]];
    const uint64_t npix = in.sh[level].hs.npix;
    const uint64_t items = (npix + rays - 1) / rays;
    prefix[j + 1] = prefix[j] + items;
    cost_prefix[j + 1] =
        cost_prefix[j] + static_cast<double>(items) * rays * static_cast<double>(in.sh[level].step);
  }
  la.n_items = prefix.back();
  la.n_ids = static_cast<uint32_t>(ids.size());
  la.level = level;
  la.rays = rays;
  // cost-balanced contiguous slices
  const double total = cost_prefix.back();
  uint64_t prev = 0;
  for (int q = 0; q < kQueues; ++q) {
    uint64_t end = la.n_items;
    if (q < kQueues - 1) {
      const double target = total * (q + 1) / kQueues;
      // find the item whose cumulative cost reaches target
      size_t j = 0;
      while (j + 1 < cost_prefix.size() && cost_prefix[j + 1] < target)
        ++j;
      if (j >= ids.size()) {
        end = la.n_items;
      } else {
        const BInst &in = b->host_inst[ids[j]];
        const double per_item = static_cast<double>(rays) * in.sh[level].step;
        const double within = per_item > 0 ? (target - cost_prefix[j]) / per_item : 0.;
        uint64_t e = prefix[j] + static_cast<uint64_t>(within);
        if (e > prefix[j + 1])
          e = prefix[j + 1];
        end = e;
      }
    }
    if (end < prev)
      end = prev;
    la.qbeg[q] = prev;
    la.qend[q] = end;
    prev = end;
  }
  la.qend[kQueues - 1] = la.n_items;
}

template <typename K, typename J>
int launch_jobs(hamx_ctx *ctx, K kernel, const hamx_batch::Jobs &jobs, uint64_t cap) {
  if (jobs.n == 0)
    return HAMX_OK;
  uint64_t blocks = (jobs.max + 255) / 256;
  blocks = std::min<uint64_t>(std::max<uint64_t>(blocks, 1), cap);
  hipLaunchKernelGGL(kernel, dim3(static_cast<unsigned>(blocks), jobs.n), dim3(256), 0,
                     ctx->stream, static_cast<const J *>(jobs.d));
  HAMX_HIP(ctx, hipGetLastError());
  return HAMX_OK;
}

unsigned march_blocks(const hamx_batch *b, const Variant &V, const LevelArgs &la, int c,
                      bool rev) {
  if (V.persistent)
    return static_cast<unsigned>(b->n_cu * (rev ? b->per_cu_rev : b->per_cu)[c][V.waves_per_block]);
  uint64_t nb = (la.n_items + V.waves_per_block - 1) / V.waves_per_block;
  if (V.xcd)
    nb = (nb + kQueues - 1) / kQueues * kQueues;
  return static_cast<unsigned>(nb);
}

// The reverse-mode sequence: zero the gradients, then per level from the
// outermost: shell-map adjoints (A^T), the reverse march, and the FD-cache
// transpose into the FD-map adjoint; finally the parameter sums and the
// grid-gradient unpack.
int launch_adjoint(hamx_batch *b, uint32_t v) {
  hamx_ctx *ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const Variant &V = kVariants[v];
  {
    uint64_t blocks = (std::max<uint64_t>(b->a_init_fill.max, 1) + 255) / 256;
    blocks = std::min<uint64_t>(blocks, 1024);
    hipLaunchKernelGGL(k_batch_zero, dim3(static_cast<unsigned>(blocks), b->a_init_fill.n),
                       dim3(256), 0, st, static_cast<const FillJob *>(b->a_init_fill.d),
                       b->d_actr, b->levels * 4 * kQueues);
    HAMX_HIP(ctx, hipGetLastError());
  }
  HAMX_TRY((launch_jobs<decltype(&k_batch_copy), CopyJob>(ctx, k_batch_copy, b->a_init_copy,
                                                           1024)));
  const bool lanes = b->prec == HAMX_PREC_F32;
  const Variant &VR = lanes ? kLaneRev : V;
  const unsigned threads = 64u * VR.waves_per_block;
  for (uint32_t s = b->levels; s-- > 0;) {
    if (b->a_zero[s].n) {
      uint64_t blocks = std::min<uint64_t>((b->a_zero[s].max + 255) / 256, 1024);
      hipLaunchKernelGGL(k_batch_zero, dim3(static_cast<unsigned>(blocks), b->a_zero[s].n),
                         dim3(256), 0, st, static_cast<const FillJob *>(b->a_zero[s].d),
                         static_cast<unsigned long long *>(nullptr), 0u);
      HAMX_HIP(ctx, hipGetLastError());
    }
    if (b->a_acct[s].n) {
      const unsigned ab = static_cast<unsigned>((b->a_acct[s].max + 255) / 256);
      hipLaunchKernelGGL(k_batch_acc_t, dim3(ab, b->a_acct[s].n), dim3(256), 0, st,
                         static_cast<const AccJob *>(b->a_acct[s].d));
      HAMX_HIP(ctx, hipGetLastError());
    }
    for (int c = 0; c < 4; ++c) {
      LevelArgs arg = lanes ? b->rev_largs[s * 4 + c] : b->largs[v][s * 4 + c];
      if (arg.n_items == 0)
        continue;
      arg.ctr = b->d_actr + static_cast<size_t>(s * 4 + c) * kQueues;
      void *args[] = {&arg};
      HAMX_HIP(ctx, hipLaunchKernel(reinterpret_cast<const void *>(march_kernel(b->prec, c, true)),
                                    dim3(march_blocks(b, VR, arg, c, true)), dim3(threads), args,
                                    0, st));
    }
    if (b->a_fdt[s].n) {
      const unsigned fb = static_cast<unsigned>((b->a_fdt[s].max + 255) / 256);
      hipLaunchKernelGGL(k_batch_fdcache_t, dim3(fb, b->a_fdt[s].n), dim3(256), 0, st,
                         static_cast<const FdtJob *>(b->a_fdt[s].d));
      HAMX_HIP(ctx, hipGetLastError());
    }
  }
  HAMX_TRY((launch_jobs<decltype(&k_batch_reduce), RedJob>(ctx, k_batch_reduce, b->a_red, 64)));
  HAMX_TRY((launch_jobs<decltype(&k_batch_unpack), UnpackJob>(ctx, k_batch_unpack, b->a_unpack,
                                                               2048)));
  return HAMX_OK;
}

int launch_all(hamx_batch *b, uint32_t v) {
  hamx_ctx *ctx = b->ctx;
  hipStream_t st = ctx->stream;
  const Variant &V = kVariants[v];
  // zero outputs and tickets
  {
    uint64_t blocks = (b->fill_max + 255) / 256;
    blocks = std::min<uint64_t>(std::max<uint64_t>(blocks, 1), 1024);
    hipLaunchKernelGGL(k_batch_zero, dim3(static_cast<unsigned>(blocks), b->n_fill), dim3(256),
                       0, st, static_cast<const FillJob *>(b->d_fill), b->d_ctr,
                       b->levels * 4 * kQueues);
    HAMX_HIP(ctx, hipGetLastError());
  }
  if (b->n_pack) {
    uint64_t blocks = (b->pack_max + 255) / 256;
    blocks = std::min<uint64_t>(std::max<uint64_t>(blocks, 1), 2048);
    hipLaunchKernelGGL(k_batch_pack, dim3(static_cast<unsigned>(blocks), b->n_pack), dim3(256),
                       0, st, static_cast<const PackJob *>(b->d_pack));
    HAMX_HIP(ctx, hipGetLastError());
  }
  const unsigned threads = 64u * V.waves_per_block;
  for (uint32_t s = 0; s < b->levels; ++s) {
    if (b->adjoint) // the FD map before this level, for the reverse mode
      HAMX_TRY((launch_jobs<decltype(&k_batch_copy), CopyJob>(ctx, k_batch_copy, b->f_ckpt[s],
                                                               1024)));
    for (int c = 0; c < 4; ++c) {
      const LevelArgs &la = b->largs[v][s * 4 + c];
      if (la.n_items == 0)
        continue;
      const unsigned blocks = march_blocks(b, V, la, c, false);
      LevelArgs arg = la;
      void *args[] = {&arg};
      HAMX_HIP(ctx, hipLaunchKernel(reinterpret_cast<const void *>(march_kernel(b->prec, c)),
                                    dim3(blocks), dim3(threads), args, 0, st));
    }
    if (b->n_acc[s]) {
      const unsigned ab = static_cast<unsigned>((b->acc_max_npix[s] + 255) / 256);
      hipLaunchKernelGGL(k_batch_acc, dim3(ab, b->n_acc[s]), dim3(256), 0, st,
                         static_cast<const AccJob *>(b->d_acc[s]));
      HAMX_HIP(ctx, hipGetLastError());
    }
  }
  return HAMX_OK;
}

int check_inst(hamx_ctx *ctx, const hamx_batch_inst &x, uint32_t i) {
  const hamx_los_desc &d = x.los;
  if (x.n_freq < 1 || x.n_freq > HAMX_MAX_FREQ)
    return set_err(ctx, HAMX_ERR_ARG, "batch instance %u: n_freq %u outside [1, %d]", i,
                   x.n_freq, HAMX_MAX_FREQ);
  for (uint32_t f = 0; f < x.n_freq; ++f)
    if (!(x.freq[f] > 0.))
      return set_err(ctx, HAMX_ERR_ARG, "batch instance %u: frequency %u not positive", i, f);
  if (d.breg_model != HAMX_BREG_LSA || d.tereg_model != HAMX_TEREG_UNIF ||
      !(d.cre_model == HAMX_CRE_ANA || d.cre_model == HAMX_CRE_UNIF) ||
      (d.cre_model == HAMX_CRE_ANA && (d.cre[1] != 0. || d.cre[2] != 0.)))
    return set_err(ctx, HAMX_ERR_UNSUPPORTED,
                   "batch instance %u: the batch covers LSA B, uniform TE and a "
                   "constant-index CRE", i);
  const bool any_b = x.brnd_x || x.brnd_y || x.brnd_z;
  if (any_b && !(x.brnd_x && x.brnd_y && x.brnd_z))
    return set_err(ctx, HAMX_ERR_ARG, "batch instance %u: brnd needs all three components", i);
  if (any_b || x.ternd)
    HAMX_TRY(check_box(ctx, x.box, "batch grid"));
  for (uint32_t f = 0; f < x.n_freq; ++f)
    if (!x.is[f] || !x.qs[f] || !x.us[f])
      return set_err(ctx, HAMX_ERR_ARG, "batch instance %u: sync maps of frequency %u missing",
                     i, f);
  return HAMX_OK;
}

} // namespace

extern "C" int hamx_batch_create(hamx_ctx *ctx, const hamx_batch_inst *inst, uint32_t n_inst,
                                 uint32_t precision, uint32_t flags, hamx_batch **out) {
  if (!ctx)
    return HAMX_ERR_ARG;
  if (!out || !inst || n_inst == 0)
    return set_err(ctx, HAMX_ERR_ARG, "hamx_batch_create: empty batch or null output");
  *out = nullptr;
  if (precision != HAMX_PREC_F64 && precision != HAMX_PREC_F32)
    return set_err(ctx, HAMX_ERR_ARG, "hamx_batch_create: unknown precision %u", precision);
  if ((flags & (HAMX_IN_DEVICE | HAMX_OUT_DEVICE)) != (HAMX_IN_DEVICE | HAMX_OUT_DEVICE))
    return set_err(ctx, HAMX_ERR_ARG,
                   "hamx_batch_create: grids and maps must be device buffers "
                   "(HAMX_IN_DEVICE | HAMX_OUT_DEVICE)");
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  for (uint32_t i = 0; i < n_inst; ++i) {
    HAMX_TRY(check_inst(ctx, inst[i], i));
    hamx_los_desc d = inst[i].los;
    d.do_sync = 1;
    d.sync_freq = inst[i].freq[inst[i].n_freq - 1];
    d.schedule = HAMX_SCHED_AUTO;
    HAMX_TRY(validate_los(ctx, &d));
  }
  hamx_batch *b = new (std::nothrow) hamx_batch;
  if (!b)
    return set_err(ctx, HAMX_ERR_NOMEM, "hamx_batch_create: host allocation failed");
  b->ctx = ctx;
  b->prec = precision;
  b->adjoint = (flags & HAMX_BATCH_ADJOINT) != 0;
  int rc = HAMX_OK;
  auto fail = [&](int code) {
    release(b);
    delete b;
    return code;
  };
  {
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ctx->device) != hipSuccess)
      return fail(set_err(ctx, HAMX_ERR_DEVICE, "hipGetDeviceProperties failed"));
    b->n_cu = prop.multiProcessorCount;
    for (int c = 0; c < 4; ++c)
      for (uint32_t w : {1u, 2u, 4u}) {
        int per_cu = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void *>(march_kernel(precision, c)),
                static_cast<int>(64 * w), 0) != hipSuccess)
          return fail(set_err(ctx, HAMX_ERR_DEVICE, "occupancy query failed"));
        b->per_cu[c][w] = per_cu < 1 ? 1 : per_cu;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                &per_cu, reinterpret_cast<const void *>(march_kernel(precision, c, true)),
                static_cast<int>(64 * w), 0) != hipSuccess)
          return fail(set_err(ctx, HAMX_ERR_DEVICE, "occupancy query failed"));
        b->per_cu_rev[c][w] = per_cu < 1 ? 1 : per_cu;
      }
  }
  b->host_inst.resize(n_inst);
  std::vector<FillJob> fills;
  std::vector<PackJob> packs;
  // packed grids, shared by instances with identical grid pointers and box
  struct GridKey {
    const double *x, *y, *z, *t;
    hamx_box box;
    void *packed;
  };
  std::vector<GridKey> grid_keys;
  // reverse mode: grid-gradient buffers shared by instances naming the same
  // output arrays, and the one-off jobs of every replay
  struct GradKey {
    double *x, *y, *z, *t;
    hamx_box box;
    void *ggrid;
  };
  std::vector<GradKey> grad_keys;
  std::vector<FillJob> a_fills;
  std::vector<CopyJob> a_copies;
  std::vector<RedJob> a_reds;
  std::vector<UnpackJob> a_unpacks;
  uint64_t samples_geo = 0, samples_ref = 0;
  for (uint32_t i = 0; i < n_inst; ++i) {
    const hamx_batch_inst &x = inst[i];
    BInst &in = b->host_inst[i];
    std::memset(&in, 0, sizeof(in));
    const uint32_t K = x.n_freq;
    // pass f of the batch = frequency freq[K-1-f] (the reference's pop_back order)
    hamx_los_desc d0 = x.los;
    d0.do_sync = 1;
    d0.sync_freq = x.freq[K - 1];
    in.ph = make_phys(&d0);
    const bool pw1 = 0.5 * (in.ph.idx + 1) == -1.0;
    in.pw1 = pw1 ? 1 : 0;
    in.n_shell = static_cast<int>(x.los.n_shell);
    in.n_freq = static_cast<int>(K);
    in.do_dm = x.los.do_dm ? 1 : 0;
    in.do_fd = x.los.do_fd ? 1 : 0;
    for (uint32_t s = 0; s < x.los.n_shell; ++s) {
      BShell &sh = in.sh[s];
      sh.hs = healpix_make(x.los.nside_shell[s]);
      shell_ref(&d0, s, sh.step, sh.d_start, sh.delta);
      sh.fdu = in.ph.fd_fore * 1e-6 * sh.delta;
      in.fc[s] = make_fast(&d0, in.ph, sh.delta, pw1);
      samples_geo += sh.hs.npix * sh.step;
      samples_ref += sh.hs.npix * sh.step * K;
    }
    // per-pass constants
    const FastConst raw0 = make_fast(&d0, in.ph, 1.0, false);
    const FastConst pw0 = make_fast(&d0, in.ph, 1.0, pw1);
    const double pw_e = raw0.pw_e;
    double a_f[kMaxF] = {0., 0., 0., 0.}, l2_f[kMaxF] = {0., 0., 0., 0.};
    for (uint32_t f = 0; f < K; ++f) {
      hamx_los_desc df = d0;
      df.sync_freq = x.freq[K - 1 - f];
      const Phys pf = make_phys(&df);
      const FastConst raw = make_fast(&df, pf, 1.0, false);
      const FastConst pwf = make_fast(&df, pf, 1.0, pw1);
      const double tc = pf.tc_em1 * pf.tc_em1 / pf.tc_den;
      a_f[f] = raw.a_coef;
      l2_f[f] = pf.lambda_sq;
      in.two_l2[f] = 2. * pf.lambda_sq;
      in.rev[f] = pf.lambda_sq / cgs::pi;
      in.fdin[f] = (f == 0 && x.los.do_fd) ? 1 : 0;
      if (precision == HAMX_PREC_F64) {
        // ratio of this pass's emissivity to pass 0's, as eval_fast forms it
        double rt = pwf.kt / pw0.kt, rp = pwf.kp / pw0.kp;
        if (!pw1) {
          const double ra = std::pow(raw.a_coef / raw0.a_coef, pw_e);
          rt *= ra;
          rp *= ra;
        }
        in.st[f] = rt * tc;
        in.sp[f] = rp * tc;
      } else {
        const double unit = std::pow(1e-6, 1. - pw_e) * std::pow(raw.a_coef, pw_e);
        in.st[f] = raw.kt * unit * tc;
        in.sp[f] = raw.kp * unit * tc;
      }
    }
    // FP32 constants
    in.f.psi0 = static_cast<float>(x.los.breg[1]);
    in.f.psi1 = static_cast<float>(x.los.breg[2]);
    in.f.chi0 = static_cast<float>(x.los.breg[3]);
    // |ln(r / 8 kpc)| <= ln(8 / 3) over the LSA window 3 <= r <= 20 kpc
    in.f.lsa_direct = (std::fabs(x.los.breg[1]) + std::fabs(x.los.breg[2]) * 0.99 <=
                           0.25 * cgs::pi &&
                       std::fabs(x.los.breg[3]) <= 0.25 * cgs::pi)
                          ? 1
                          : 0;
    in.f.b0u = static_cast<float>(x.los.breg[0] * 1e6);
    in.f.n0 = static_cast<float>(x.los.tereg[0]);
    in.f.n0d = x.los.tereg[0];
    in.f.b0d = x.los.breg[0] * 1e6;
    if (x.los.cre_model == HAMX_CRE_ANA) {
      in.f.R0k = static_cast<float>(in.ph.R0 / cgs::kpc);
      in.f.inv_r0k = static_cast<float>(cgs::kpc / x.los.cre[3]);
      in.f.inv_z0k = static_cast<float>(cgs::kpc / x.los.cre[4]);
    }
    in.f.hexp = static_cast<float>(0.5 * (1. - pw_e));
    in.r_min2 = x.los.gc_r_min * x.los.gc_r_min;
    in.r_max2 = x.los.gc_r_max * x.los.gc_r_max;
    in.te_r02 = x.los.tereg[1] * x.los.tereg[1];
    in.cre_r02 = x.los.cre[1] * x.los.cre[1];
    // grids
    const bool any_grid = x.brnd_x || x.ternd;
    in.grids = any_grid ? 1 : 0;
    if (any_grid) {
      in.geo = grid_geo(x.box);
      in.gx0 = x.box.x_min;
      in.gy0 = x.box.y_min;
      in.gz0 = x.box.z_min;
      in.gsx = in.geo.fx / (x.box.x_max - x.box.x_min);
      in.gsy = in.geo.fy / (x.box.y_max - x.box.y_min);
      in.gsz = in.geo.fz / (x.box.z_max - x.box.z_min);
      void *packed = nullptr;
      for (const GridKey &g : grid_keys)
        if (g.x == x.brnd_x && g.y == x.brnd_y && g.z == x.brnd_z && g.t == x.ternd &&
            std::memcmp(&g.box, &x.box, sizeof(hamx_box)) == 0)
          packed = g.packed;
      if (!packed) {
        const uint64_t n = box_size(x.box);
        const size_t rec = precision == HAMX_PREC_F32 ? sizeof(float4) : sizeof(double4);
        if ((rc = dev_alloc(b, n * rec, &packed)) != HAMX_OK)
          return fail(rc);
        grid_keys.push_back(GridKey{x.brnd_x, x.brnd_y, x.brnd_z, x.ternd, x.box, packed});
        PackJob pj;
        std::memset(&pj, 0, sizeof(pj));
        pj.dst = packed;
        pj.bx = x.brnd_x;
        pj.by = x.brnd_y;
        pj.bz = x.brnd_z;
        pj.te = x.ternd;
        pj.n = n;
        pj.f32 = precision == HAMX_PREC_F32;
        pj.bscale = pj.f32 ? 1e6 : 1.;
        packs.push_back(pj);
        b->pack_max = std::max(b->pack_max, n);
      }
      in.packed = packed;
    }
    // output and shell maps
    uint64_t max_npix = 0;
    for (uint32_t s = 0; s < x.los.n_shell; ++s)
      max_npix = std::max<uint64_t>(max_npix, in.sh[s].hs.npix);
    const Healpix h_sync = healpix_make(x.los.nside_sync);
    in.hfd = healpix_make(x.los.do_fd ? x.los.nside_fd : 1);
    void *p;
    if (in.do_dm) {
      if ((rc = dev_alloc(b, max_npix * 8, &p)) != HAMX_OK)
        return fail(rc);
      in.t_dm = static_cast<double *>(p);
      if (x.dm)
        fills.push_back(FillJob{x.dm, healpix_make(x.los.nside_dm).npix});
    }
    {
      uint64_t rays = 0; // every ray's FD of every level (the reverse mode's tape)
      for (uint32_t s = 0; s < x.los.n_shell; ++s) {
        in.fd_off[s] = rays;
        rays += in.sh[s].hs.npix;
      }
      if ((rc = dev_alloc(b, rays * 8, &p)) != HAMX_OK)
        return fail(rc);
      in.t_fd = static_cast<double *>(p);
    }
    if (in.do_fd) {
      double *fdm = x.fd;
      if (!fdm) { // the FD cache needs the accumulated map even when not requested
        if ((rc = dev_alloc(b, in.hfd.npix * 8, &p)) != HAMX_OK)
          return fail(rc);
        fdm = static_cast<double *>(p);
      }
      in.fd_cache = fdm;
      fills.push_back(FillJob{fdm, in.hfd.npix});
    }
    for (uint32_t f = 0; f < K; ++f) {
      double **t[3] = {&in.t_i[f], &in.t_q[f], &in.t_u[f]};
      for (int m = 0; m < 3; ++m) {
        if ((rc = dev_alloc(b, max_npix * 8, &p)) != HAMX_OK)
          return fail(rc);
        *t[m] = static_cast<double *>(p);
      }
      // pass f writes the maps of frequency freq[K-1-f]
      fills.push_back(FillJob{x.is[K - 1 - f], h_sync.npix});
      fills.push_back(FillJob{x.qs[K - 1 - f], h_sync.npix});
      fills.push_back(FillJob{x.us[K - 1 - f], h_sync.npix});
    }
    if (b->adjoint) {
      BAdj &A = in.adj;
      auto alloc_d = [&](uint64_t n, double **dst) -> int {
        void *q;
        const int r2 = dev_alloc(b, n * 8, &q);
        *dst = static_cast<double *>(q);
        return r2;
      };
      for (uint32_t f = 0; f < K; ++f)
        if ((rc = alloc_d(max_npix, &A.g_i[f])) != HAMX_OK ||
            (rc = alloc_d(max_npix, &A.g_q[f])) != HAMX_OK ||
            (rc = alloc_d(max_npix, &A.g_u[f])) != HAMX_OK)
          return fail(rc);
      if (in.do_dm && (rc = alloc_d(max_npix, &A.g_dm)) != HAMX_OK)
        return fail(rc);
      if (in.do_fd) {
        if ((rc = alloc_d(max_npix, &A.g_fd)) != HAMX_OK ||
            (rc = alloc_d(in.hfd.npix, &A.G)) != HAMX_OK ||
            (rc = alloc_d(max_npix, &A.gfdin)) != HAMX_OK ||
            (rc = alloc_d(x.los.n_shell * in.hfd.npix, &A.ckpt)) != HAMX_OK)
          return fail(rc);
        if (x.adj_fd)
          a_copies.push_back(CopyJob{A.G, x.adj_fd, in.hfd.npix});
        else
          a_fills.push_back(FillJob{A.G, in.hfd.npix});
      }
      const uint32_t last = x.los.n_shell - 1;
      const uint64_t rays = in.fd_off[last] + in.sh[last].hs.npix;
      if ((rc = alloc_d(6 * rays, &A.ray_par)) != HAMX_OK)
        return fail(rc);
      double *dpar = x.dparams;
      if (!dpar && (rc = alloc_d(HAMX_NPARAM, &dpar)) != HAMX_OK)
        return fail(rc);
      a_fills.push_back(FillJob{dpar, HAMX_NPARAM});
      a_reds.push_back(RedJob{A.ray_par, rays, dpar});
      // d ln(j_f)/d alpha: ln gamma0 - ln(a_f)/2 plus the Gamma factors' digammas
      const double alpha = -in.ph.idx;
      const double E0 = x.los.cre_model == HAMX_CRE_UNIF ? x.los.cre[2] : x.los.cre[5];
      const double ln_g0 = std::log(E0 / cgs::mec2);
      const double psi_t = 0.25 * digamma(alpha / 4. + 19. / 12.);
      const double psi_p = 0.25 * digamma(alpha / 4. + 7. / 12.);
      const double psi_c = 0.25 * digamma(alpha / 4. - 1. / 12.);
      for (uint32_t f = 0; f < K; ++f) {
        const double base = ln_g0 - 0.5 * std::log(a_f[f]);
        A.lt[f] = base - 1. / (1. + alpha) + psi_t + psi_c;
        A.lp[f] = base + psi_p + psi_c;
        A.l2[f] = l2_f[f];
      }
      A.b0 = x.los.breg[0];
      if (in.grids && (x.dbrnd_x || x.dbrnd_y || x.dbrnd_z || x.dternd)) {
        void *gg = nullptr;
        for (const GradKey &g : grad_keys)
          if (g.x == x.dbrnd_x && g.y == x.dbrnd_y && g.z == x.dbrnd_z && g.t == x.dternd) {
            if (std::memcmp(&g.box, &x.box, sizeof(hamx_box)) != 0)
              return fail(set_err(ctx, HAMX_ERR_ARG,
                                  "batch instance %u: shared grid-gradient arrays on "
                                  "different boxes", i));
            gg = g.ggrid;
          }
        if (!gg) {
          const uint64_t n = box_size(x.box);
          const bool f32 = precision == HAMX_PREC_F32;
          if ((rc = dev_alloc(b, n * (f32 ? sizeof(float4) : sizeof(double4)), &gg)) != HAMX_OK)
            return fail(rc);
          grad_keys.push_back(GradKey{x.dbrnd_x, x.dbrnd_y, x.dbrnd_z, x.dternd, x.box, gg});
          a_fills.push_back(FillJob{static_cast<double *>(gg), n * (f32 ? 2 : 4)});
          UnpackJob u;
          std::memset(&u, 0, sizeof(u));
          u.g = gg;
          u.x = x.dbrnd_x;
          u.y = x.dbrnd_y;
          u.z = x.dbrnd_z;
          u.t = x.dternd;
          u.n = n;
          u.f32 = f32 ? 1 : 0;
          u.bscale = f32 ? 1e6 : 1.;
          a_unpacks.push_back(u);
        }
        A.ggrid = gg;
      }
    }
    b->levels = std::max<uint32_t>(b->levels, x.los.n_shell);
  }
  if (fills.size() > 65535 || n_inst > 65535 / 3)
    return fail(set_err(ctx, HAMX_ERR_ARG, "hamx_batch_create: too many instances"));
  for (const FillJob &f : fills)
    b->fill_max = std::max(b->fill_max, f.n);
  b->n_fill = static_cast<uint32_t>(fills.size());
  b->n_pack = static_cast<uint32_t>(packs.size());
  if ((rc = upload(b, fills, &b->d_fill)) != HAMX_OK)
    return fail(rc);
  if (!packs.empty() && (rc = upload(b, packs, &b->d_pack)) != HAMX_OK)
    return fail(rc);
  {
    void *p;
    if ((rc = dev_alloc(b, sizeof(unsigned long long) * kQueues * 4 * b->levels, &p)) !=
        HAMX_OK)
      return fail(rc);
    b->d_ctr = static_cast<unsigned long long *>(p);
  }
  if ((rc = upload(b, b->host_inst, &b->d_inst)) != HAMX_OK)
    return fail(rc);
  // levels: instance lists (instance-major item order) and accumulate jobs
  b->ids.resize(b->levels * 4);
  b->d_ids.resize(b->levels * 4, nullptr);
  b->d_acc.resize(b->levels);
  b->n_acc.assign(b->levels, 0);
  b->acc_max_npix.assign(b->levels, 0);
  b->a_zero.assign(b->levels, hamx_batch::Jobs{});
  b->a_acct.assign(b->levels, hamx_batch::Jobs{});
  b->a_fdt.assign(b->levels, hamx_batch::Jobs{});
  b->f_ckpt.assign(b->levels, hamx_batch::Jobs{});
  auto put = [&](auto &vec, hamx_batch::Jobs &jobs, uint64_t mx) -> int {
    jobs.n = static_cast<uint32_t>(vec.size());
    jobs.max = mx;
    if (vec.empty())
      return HAMX_OK;
    using T = typename std::decay_t<decltype(vec)>::value_type;
    T *dptr;
    HAMX_TRY(upload(b, vec, &dptr));
    jobs.d = dptr;
    return HAMX_OK;
  };
  for (uint32_t s = 0; s < b->levels; ++s) {
    std::vector<AccJob> jobs;
    std::vector<FillJob> zs;
    std::vector<AccJob> ats;
    std::vector<FdtJob> fts;
    std::vector<CopyJob> cks;
    uint64_t z_max = 0, at_max = 0, ft_max = 0, ck_max = 0;
    for (uint32_t i = 0; i < n_inst; ++i) {
      const hamx_batch_inst &x = inst[i];
      const BInst &in = b->host_inst[i];
      if (s >= x.los.n_shell)
        continue;
      b->ids[s * 4 + 2 * in.grids + in.pw1].push_back(i);
      const Healpix hs = in.sh[s].hs;
      const uint32_t K = x.n_freq;
      AccJob sj;
      std::memset(&sj, 0, sizeof(sj));
      sj.hd = healpix_make(x.los.nside_sync);
      sj.hs = hs;
      for (uint32_t f = 0; f < K; ++f) {
        sj.dst[sj.n] = x.is[K - 1 - f];
        sj.src[sj.n++] = in.t_i[f];
        sj.dst[sj.n] = x.qs[K - 1 - f];
        sj.src[sj.n++] = in.t_q[f];
        sj.dst[sj.n] = x.us[K - 1 - f];
        sj.src[sj.n++] = in.t_u[f];
      }
      jobs.push_back(sj);
      if (in.do_dm && x.dm) {
        AccJob j;
        std::memset(&j, 0, sizeof(j));
        j.hd = healpix_make(x.los.nside_dm);
        j.hs = hs;
        j.n = 1;
        j.dst[0] = x.dm;
        j.src[0] = in.t_dm;
        jobs.push_back(j);
      }
      if (in.do_fd) {
        AccJob j;
        std::memset(&j, 0, sizeof(j));
        j.hd = in.hfd;
        j.hs = hs;
        j.n = 1;
        j.dst[0] = const_cast<double *>(in.fd_cache);
        j.src[0] = in.t_fd + in.fd_off[s];
        jobs.push_back(j);
      }
      if (b->adjoint) {
        const BAdj &A = in.adj;
        AccJob t;
        std::memset(&t, 0, sizeof(t));
        t.hd = healpix_make(x.los.nside_sync);
        t.hs = hs;
        for (uint32_t f = 0; f < K; ++f) {
          const double *src[3] = {x.adj_is[K - 1 - f], x.adj_qs[K - 1 - f], x.adj_us[K - 1 - f]};
          double *dst[3] = {A.g_i[f], A.g_q[f], A.g_u[f]};
          for (int m = 0; m < 3; ++m) {
            zs.push_back(FillJob{dst[m], hs.npix});
            if (src[m]) { // a NULL adjoint map is zero
              t.dst[t.n] = dst[m];
              t.src[t.n++] = src[m];
            }
          }
        }
        if (t.n) {
          if ((rc = interp_transpose(b, t.hd, t.hs, t)) != HAMX_OK)
            return fail(rc);
          ats.push_back(t);
          at_max = std::max({at_max, t.hd.npix, t.hs.npix});
        }
        if (in.do_dm) {
          zs.push_back(FillJob{A.g_dm, hs.npix});
          if (x.adj_dm) {
            AccJob j;
            std::memset(&j, 0, sizeof(j));
            j.hd = healpix_make(x.los.nside_dm);
            j.hs = hs;
            j.n = 1;
            j.dst[0] = A.g_dm;
            j.src[0] = x.adj_dm;
            if ((rc = interp_transpose(b, j.hd, j.hs, j)) != HAMX_OK)
              return fail(rc);
            ats.push_back(j);
            at_max = std::max({at_max, j.hd.npix, j.hs.npix});
          }
        }
        if (in.do_fd) {
          zs.push_back(FillJob{A.g_fd, hs.npix});
          AccJob j;
          std::memset(&j, 0, sizeof(j));
          j.hd = in.hfd;
          j.hs = hs;
          j.n = 1;
          j.dst[0] = A.g_fd;
          j.src[0] = A.G;
          if ((rc = interp_transpose(b, j.hd, j.hs, j)) != HAMX_OK)
            return fail(rc);
          ats.push_back(j);
          at_max = std::max({at_max, j.hd.npix, j.hs.npix});
          if (s > 0) {
            fts.push_back(FdtJob{A.G, A.gfdin, in.hfd, hs});
            ft_max = std::max(ft_max, hs.npix);
            cks.push_back(CopyJob{A.ckpt + static_cast<uint64_t>(s) * in.hfd.npix,
                                  in.fd_cache, in.hfd.npix});
            ck_max = std::max(ck_max, in.hfd.npix);
          }
        }
        z_max = std::max(z_max, hs.npix);
      }
    }
    if ((rc = put(zs, b->a_zero[s], z_max)) != HAMX_OK ||
        (rc = put(ats, b->a_acct[s], at_max)) != HAMX_OK ||
        (rc = put(fts, b->a_fdt[s], ft_max)) != HAMX_OK ||
        (rc = put(cks, b->f_ckpt[s], ck_max)) != HAMX_OK)
      return fail(rc);
    for (const AccJob &j : jobs)
      b->acc_max_npix[s] = std::max(b->acc_max_npix[s], j.hd.npix);
    b->n_acc[s] = static_cast<uint32_t>(jobs.size());
    if ((rc = upload(b, jobs, &b->d_acc[s])) != HAMX_OK)
      return fail(rc);
    for (int c = 0; c < 4; ++c)
      if ((rc = upload(b, b->ids[s * 4 + c], &b->d_ids[s * 4 + c])) != HAMX_OK)
        return fail(rc);
  }
  if (b->adjoint) {
    uint64_t mx = 0;
    for (const FillJob &f : a_fills)
      mx = std::max(mx, f.n);
    if ((rc = put(a_fills, b->a_init_fill, mx)) != HAMX_OK)
      return fail(rc);
    mx = 0;
    for (const CopyJob &c : a_copies)
      mx = std::max(mx, c.n);
    if ((rc = put(a_copies, b->a_init_copy, mx)) != HAMX_OK)
      return fail(rc);
    mx = 0;
    for (const RedJob &r2 : a_reds)
      mx = std::max(mx, r2.n);
    if ((rc = put(a_reds, b->a_red, mx)) != HAMX_OK)
      return fail(rc);
    mx = 0;
    for (const UnpackJob &u : a_unpacks)
      mx = std::max(mx, u.n);
    if ((rc = put(a_unpacks, b->a_unpack, mx)) != HAMX_OK)
      return fail(rc);
    if (b->a_init_fill.n > 65535 || b->a_red.n > 65535)
      return fail(set_err(ctx, HAMX_ERR_ARG, "hamx_batch_create: too many instances"));
    void *p;
    if ((rc = dev_alloc(b, sizeof(unsigned long long) * kQueues * 4 * b->levels, &p)) !=
        HAMX_OK)
      return fail(rc);
    b->d_actr = static_cast<unsigned long long *>(p);
  }
  // item plans of every variant
  b->d_prefix.assign(kNumVariants, std::vector<uint64_t *>(b->levels * 4, nullptr));
  b->largs.assign(kNumVariants, std::vector<LevelArgs>(b->levels * 4));
  for (uint32_t v = 0; v < kNumVariants; ++v) {
    for (uint32_t s = 0; s < b->levels * 4; ++s) {
      LevelArgs &la = b->largs[v][s];
      std::memset(&la, 0, sizeof(la));
      std::vector<uint64_t> prefix;
      plan_level(b, s, kVariants[v].rays, prefix, la);
      if ((rc = upload(b, prefix, &b->d_prefix[v][s])) != HAMX_OK)
        return fail(rc);
      la.inst = b->d_inst;
      la.ids = b->d_ids[s];
      la.prefix = b->d_prefix[v][s];
      la.ctr = b->d_ctr + static_cast<size_t>(s) * kQueues;
      la.persistent = kVariants[v].persistent;
      la.xcd = kVariants[v].xcd;
      la.nq = kVariants[v].xcd ? kQueues : 1;
      if (!la.xcd) { // one global queue
        la.qbeg[0] = 0;
        la.qend[0] = la.n_items;
      }
    }
  }
  // the FP32 reverse plan
  if (b->adjoint && b->prec == HAMX_PREC_F32) {
    b->rev_largs.assign(b->levels * 4, LevelArgs{});
    for (uint32_t s = 0; s < b->levels * 4; ++s) {
      LevelArgs &la = b->rev_largs[s];
      std::memset(&la, 0, sizeof(la));
      std::vector<uint64_t> prefix;
      plan_level(b, s, kLaneRev.rays, prefix, la);
      uint64_t *d = nullptr;
      if ((rc = upload(b, prefix, &d)) != HAMX_OK)
        return fail(rc);
      la.inst = b->d_inst;
      la.ids = b->d_ids[s];
      la.prefix = d;
      la.persistent = kLaneRev.persistent;
      la.xcd = kLaneRev.xcd;
      la.nq = kQueues;
    }
  }
  // autotune: one warm-up and one timed eager run per candidate
  b->info.samples_geo = samples_geo;
  b->info.samples_ref = samples_ref;
  b->info.levels = b->levels;
  const bool lanes = (flags & HAMX_BATCH_LANES) != 0;
  const bool tune = (flags & HAMX_BATCH_NO_TUNE) == 0 && !lanes;
  uint32_t best = 0;
  if (lanes) { // the first one-ray-per-lane shape
    if (b->prec != HAMX_PREC_F32)
      return fail(set_err(ctx, HAMX_ERR_ARG,
                          "hamx_batch_create: HAMX_BATCH_LANES needs HAMX_PREC_F32"));
    while (kVariants[best].rays <= static_cast<uint32_t>(kItemRaysMax))
      ++best;
  }
  if (tune) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
      return fail(set_err(ctx, HAMX_ERR_DEVICE, "hipEventCreate failed"));
    double best_ms = 1e300;
    for (uint32_t v = 0; v < kNumVariants && rc == HAMX_OK; ++v) {
      // lane-per-ray items exist only in the FP32 march (the FP64 one sets
      // up at most kItemRaysMax rays per item in LDS)
      if (kVariants[v].rays > static_cast<uint32_t>(kItemRaysMax) &&
          b->prec != HAMX_PREC_F32) {
        b->info.variant_ms[v] = -1.;
        continue;
      }
      rc = launch_all(b, v);
      if (rc == HAMX_OK)
        (void)hipEventRecord(e0, ctx->stream);
      if (rc == HAMX_OK)
        rc = launch_all(b, v);
      if (rc == HAMX_OK) {
        (void)hipEventRecord(e1, ctx->stream);
        if (hipEventSynchronize(e1) != hipSuccess)
          rc = set_err(ctx, HAMX_ERR_DEVICE, "batch autotune: %s",
                       hipGetErrorString(hipGetLastError()));
      }
      if (rc == HAMX_OK) {
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        b->info.variant_ms[v] = ms;
        if (ms < best_ms) {
          best_ms = ms;
          best = v;
        }
      }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != HAMX_OK)
      return fail(rc);
    b->info.n_variants = kNumVariants;
  }
  b->variant = best;
  b->info.variant = best;
  std::snprintf(b->info.shape, sizeof(b->info.shape), "%s", kVariants[best].name);
  // capture the chosen variant
  if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(set_err(ctx, HAMX_ERR_DEVICE, "hipStreamBeginCapture failed"));
  rc = launch_all(b, best);
  hipGraph_t g = nullptr;
  const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
  if (rc != HAMX_OK) {
    if (g)
      (void)hipGraphDestroy(g);
    return fail(rc);
  }
  if (ec != hipSuccess || !g)
    return fail(set_err(ctx, HAMX_ERR_DEVICE, "hipStreamEndCapture failed"));
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess)
    return fail(set_err(ctx, HAMX_ERR_DEVICE, "hipGraphInstantiate failed"));
  if (b->adjoint) {
    if (hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return fail(set_err(ctx, HAMX_ERR_DEVICE, "hipStreamBeginCapture failed"));
    rc = launch_adjoint(b, best);
    hipGraph_t ga = nullptr;
    const hipError_t ea = hipStreamEndCapture(ctx->stream, &ga);
    if (rc != HAMX_OK) {
      if (ga)
        (void)hipGraphDestroy(ga);
      return fail(rc);
    }
    if (ea != hipSuccess || !ga)
      return fail(set_err(ctx, HAMX_ERR_DEVICE, "hipStreamEndCapture failed"));
    b->adj_graph = ga;
    if (hipGraphInstantiate(&b->adj_exec, ga, nullptr, nullptr, 0) != hipSuccess)
      return fail(set_err(ctx, HAMX_ERR_DEVICE, "hipGraphInstantiate failed"));
  }
  *out = b;
  return HAMX_OK;
}

extern "C" int hamx_batch_run(hamx_batch *b) {
  if (!b || !b->exec)
    return HAMX_ERR_ARG;
  hamx_ctx *ctx = b->ctx;
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  HAMX_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}

extern "C" int hamx_batch_run_adjoint(hamx_batch *b) {
  if (!b)
    return HAMX_ERR_ARG;
  if (!b->adj_exec)
    return set_err(b->ctx, HAMX_ERR_ARG,
                   "hamx_batch_run_adjoint: the batch was created without HAMX_BATCH_ADJOINT");
  hamx_ctx *ctx = b->ctx;
  HAMX_HIP(ctx, hipSetDevice(ctx->device));
  HAMX_HIP(ctx, hipGraphLaunch(b->adj_exec, ctx->stream));
  HAMX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return HAMX_OK;
}

extern "C" int hamx_batch_get_info(const hamx_batch *b, hamx_batch_info *info) {
  if (!b || !info)
    return HAMX_ERR_ARG;
  *info = b->info;
  return HAMX_OK;
}

extern "C" void hamx_batch_destroy(hamx_batch *b) {
  if (!b)
    return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  release(b);
  delete b;
}
