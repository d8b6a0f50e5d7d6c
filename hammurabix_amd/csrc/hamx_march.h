// Shared pieces of the LOS march (forward in hamx_los.hip, adjoint in
// hamx_adj.hip): pass constants, grid and shell descriptors, per-ray set-up
// with the conservative step interval, wavefront reductions, and the
// host-side descriptor checks and constant folding.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "hamx.h"
#include "hamx_cgs.h"

#include "hamx_fields.h"
#include "hamx_math.h"
#include "hamx_healpix.h"
#include "hamx_internal.h"

namespace hamx {

// Constants of one write_grid pass (host-computed in the reference's
// expression order).
struct Phys {
  Vec3 obs;
  double gc_r_min, gc_r_max, gc_z_min, gc_z_max;
  double gc_r_min2, gc_r_max2; // squared radial limits (eval_fast's decisions)
  int do_dm, do_fd, do_sync;
  int breg_model, tereg_model, cre_model;
  double breg[HAMX_BREG_NPARAM], tereg[HAMX_TEREG_NPARAM], cre[8];
  double fd_fore;       // -(qe^3) / (2 pi mec2^2), integrator.cc:160-161
  double lambda_sq;     // (c/nu)^2, integrator.cc:164-165
  double i2bt;          // c^2 / (2 kB nu^2), integrator.cc:168-170
  double A_num;         // 2 pi nu mec, integrator.cc:341-342
  double three_qe;      // 3 qe
  double qe3;           // qe^3
  double fourpi_mec2;   // 4 pi mec2
  double sixteenpi_mec2;// 16 pi mec2
  double cre_nu;        // j0 * gamma0^(-idx(obs)) * unit, cre_ana.cc:221-229
  double R0;            // observer cylindrical radius, cre_ana.cc:196-197
  int idx_const;        // spectral index constant along rays
  double idx, g_t1, g_t2, g_p1; // index and its Gamma factors when constant
  double tc_em1, tc_den; // temp_convert: (e^p - 1) and p^2 e^p, integrator.cc:275-279
  // CRE grid emissivities (integrator.cc:296-298,325-327,365-367,379-381):
  // the x_fact numerators of J_tot and J_pol in their own operation orders,
  // and the fore factor's 1.73205081 qe^3 and denominator.
  double xft_num, xfp_num, fore_num, fore_den;
};

struct Grids {
  const double4 *packed; // brnd (x,y,z) + ternd (w), one box
  GridGeo g_packed;
  const double *brnd_x, *brnd_y, *brnd_z;
  GridGeo g_brnd;
  const double *ternd;
  GridGeo g_ternd;
  const double *breg_x, *breg_y, *breg_z;
  GridGeo g_breg;
  const double *tereg;
  GridGeo g_tereg;
  // CRE_num grid [x][y][z][E] and its energy tables: E_i^2, beta_i,
  // |E_{i+1} - E_i| and ln((1/E_{i+1}^2 + 1/E_i^2) / 2), cre_nE doubles
  // each from cre_tab
  const double *cre;
  GridGeo g_cre;
  uint64_t cre_nE;
  const double *cre_tab;
};

struct ShellArgs {
  Healpix hs;
  uint64_t step;
  double d_start, delta;
  const double *mask;   // shell-nside mask: integrate where == 1.0 (NULL: all)
  const double *fd_map; // accumulated FD of inner shells (NULL: zero cache)
  Healpix hfd;
  double *t_dm, *t_fd, *t_is, *t_qs, *t_us; // shell maps, nside_s
};

struct RayGeo {
  double theta, phi;
  Vec3 l;         // LOS versor
  Vec3 eth, eph;  // spherical unit vectors for the IAU angle
  double fd_inner;
  uint64_t k_lo, k_hi;
};

// Trilinear gathers of separate component grids (0 outside the box).
__device__ inline Vec3 gather3(const double *fx, const double *fy,
                               const double *fz, const GridGeo &g,
                               const Vec3 &pos) {
  Cell c;
  if (!grid_cell(g, pos, c))
    return Vec3{0., 0., 0.};
  const uint64_t sy = g.nz, sx = g.ny * g.nz;
  return Vec3{lerp_cell(fx, c, sy, sx), lerp_cell(fy, c, sy, sx),
              lerp_cell(fz, c, sy, sx)};
}

__device__ inline double gather1(const double *f, const GridGeo &g,
                                 const Vec3 &pos) {
  Cell c;
  if (!grid_cell(g, pos, c))
    return 0.;
  return lerp_cell(f, c, g.nz, g.ny * g.nz);
}

// ---------------------------------------------------------------- per ray --

__device__ inline void ray_interval(const Phys &ph, const ShellArgs &sa,
                                    RayGeo &r) {
  // Conservative d-interval satisfying |pos| <= gc_r_max and
  // gc_z_min <= z <= gc_z_max, widened by 2 steps each side.
  r.k_lo = 0;
  r.k_hi = sa.step;
  if (sa.step == 0)
    return;
  const double slack = 2.0 * sa.delta;
  double lo = -1e300, hi = 1e300;
  {
    const double R = ph.gc_r_max * (1.0 + 1e-9) + slack;
    const double b = r.l.x * ph.obs.x + r.l.y * ph.obs.y + r.l.z * ph.obs.z;
    const double o2 = ph.obs.x * ph.obs.x + ph.obs.y * ph.obs.y + ph.obs.z * ph.obs.z;
    const double disc = b * b - (o2 - R * R);
    if (!(disc >= 0.)) {
      r.k_hi = 0;
      return;
    }
    const double sq = sqrt(disc);
    lo = fmax(lo, -b - sq);
    hi = fmin(hi, -b + sq);
  }
  {
    const double lz = r.l.z;
    const double zlo = ph.gc_z_min - 1e-9 * fabs(ph.gc_z_min) - slack;
    const double zhi = ph.gc_z_max + 1e-9 * fabs(ph.gc_z_max) + slack;
    if (fabs(lz) < 1e-12) {
      if (ph.obs.z < zlo || ph.obs.z > zhi) {
        r.k_hi = 0;
        return;
      }
    } else {
      double t1 = (zlo - ph.obs.z) / lz, t2 = (zhi - ph.obs.z) / lz;
      if (t1 > t2) {
        const double t = t1;
        t1 = t2;
        t2 = t;
      }
      lo = fmax(lo, t1 - slack);
      hi = fmin(hi, t2 + slack);
    }
  }
  if (!(hi >= lo)) {
    r.k_hi = 0;
    return;
  }
  // dist_k = d0 + (k + 1/2) delta  ->  k = (d - d0)/delta - 1/2
  const double kf_lo = (lo - sa.d_start) / sa.delta - 0.5 - 2.0;
  const double kf_hi = (hi - sa.d_start) / sa.delta - 0.5 + 3.0;
  const double stepd = static_cast<double>(sa.step);
  const double clo = fmin(fmax(floor(kf_lo), 0.), stepd);
  const double chi = fmin(fmax(ceil(kf_hi), 0.), stepd);
  r.k_lo = static_cast<uint64_t>(clo);
  r.k_hi = static_cast<uint64_t>(chi);
  if (r.k_hi < r.k_lo)
    r.k_hi = r.k_lo;
}

__device__ inline void ray_setup(const Phys &ph, const ShellArgs &sa,
                                 uint64_t ipix, RayGeo &r) {
  hp_pix2ang(sa.hs, ipix, r.theta, r.phi);
  double st, ct, sp, cp;
  sincos(r.theta, &st, &ct);
  sincos(r.phi, &sp, &cp);
  // los_versor (integrator.cc:268-272)
  r.l = Vec3{cp * st, sp * st, ct};
  // sync_ipa unit vectors (integrator.cc:260-262)
  r.eth = Vec3{ct * cp, ct * sp, -st};
  r.eph = Vec3{-sp, cp, 0.};
  r.fd_inner = 0.;
  if (sa.fd_map)
    r.fd_inner = hp_interpolate(sa.hfd, sa.fd_map, r.theta, r.phi);
  ray_interval(ph, sa, r);
}

// ------------------------------------------------------------------ waves --

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    v += __shfl_xor(v, o, 64);
  return v;
}

// v moved across lanes by the DPP control CTRL (both 32-bit halves), 0 in
// the lanes whose row ROW_MASK leaves out or whose source lies outside the
// row.
template <int CTRL, int ROW_MASK>
__device__ inline double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, static_cast<int>(b), CTRL, ROW_MASK, 0xf, true);
  const int hi =
      __builtin_amdgcn_update_dpp(0, static_cast<int>(b >> 32), CTRL, ROW_MASK, 0xf, true);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}

// Lane L's value in every lane (scalar reads; the wavefront is converged).
__device__ inline double wave_lane(double v, int L) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(static_cast<int>(b), L);
  const int hi = __builtin_amdgcn_readlane(static_cast<int>(b >> 32), L);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<long long>(static_cast<unsigned>(lo)));
}

__device__ inline double wave_last(double v) { return wave_lane(v, 63); }

// Inclusive prefix sum over the 64 lanes: shifts by 1, 2, 4 and 8 within
// each 16-lane row, then lane 15 of rows 0 and 2 into rows 1 and 3, then
// lane 31 into rows 2 and 3 (DPP moves, no LDS traffic).
__device__ inline double wave_incl_scan(double v, int) {
  v += dpp_f64<0x111, 0xf>(v); // row_shr:1
  v += dpp_f64<0x112, 0xf>(v); // row_shr:2
  v += dpp_f64<0x114, 0xf>(v); // row_shr:4
  v += dpp_f64<0x118, 0xf>(v); // row_shr:8
  v += dpp_f64<0x142, 0xa>(v); // row_bcast:15
  v += dpp_f64<0x143, 0xc>(v); // row_bcast:31
  return v;
}

// Sum over lanes >= lane (a reverse inclusive scan).
// Shifts by 1, 2, 4 and 8 within each 16-lane row (DPP), then each row adds
// the totals of the rows above it, read from lanes 16, 32 and 48.
__device__ inline double wave_suffix_scan(double v, int lane) {
  v += dpp_f64<0x101, 0xf>(v); // row_shl:1
  v += dpp_f64<0x102, 0xf>(v); // row_shl:2
  v += dpp_f64<0x104, 0xf>(v); // row_shl:4
  v += dpp_f64<0x108, 0xf>(v); // row_shl:8
  const double t3 = wave_lane(v, 48);
  const double t23 = wave_lane(v, 32) + t3;
  const double t123 = wave_lane(v, 16) + t23;
  const int row = lane >> 4;
  return v + (row == 0 ? t123 : row == 1 ? t23 : row == 2 ? t3 : 0.);
}

// Orders the LDS accesses of one wavefront around this point (the lanes of
// a wavefront exchange data through its own LDS slab; waves of a workgroup
// never share a slab, so no workgroup barrier is involved).
__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Atomic add to a gradient grid in device memory, issued as a global (not
// flat) atomic: a flat instruction also counts against the LDS wait counter,
// so each LDS round trip after it would wait for the atomic's round trip to
// memory as well.
template <typename T>
__device__ inline void grid_atomic_add(T *p, T v) {
  auto *gp = (__attribute__((address_space(1))) T *)p;
  [[clang::atomic(ignore_denormal_mode)]] {
    __hip_atomic_fetch_add(gp, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Rays per wavefront in the specialised march (kFastRays) and its adjoint
// (kAdjRays): their set-up runs one ray per lane into LDS, then they are
// marched one after another.  Measured on the 256x256x64 three-shell
// workload (march / adjoint ms): 1 ray 14.0 / 44.8, 2 rays 13.4 / 42.1,
// 4 rays 14.5, 8 rays 21.8 / 66.1: more rays per wavefront leave too few
// wavefronts to balance the uneven ray lengths (rays leaving the z range
// early) across the SIMDs.
constexpr int kFastRays = 2;
constexpr int kAdjRays = 2;

// ------------------------------------------------ staged grid scatter --
// The grid-gradient scatter of the FP64 reverse sweeps (the plain adjoint
// and the FP64 batch), lanes = consecutive samples.  A contended FP64 LDS atomic
// costs about 30 times a plain 64-lane LDS store or load (measured on
// gfx950), so the 32 corner values of a chunk's samples are merged without
// atomics: every lane stores its values into a per-wavefront staging area
// (16 values at a time), and the lanes then sum each run of samples that
// share a cell transposed, one lane per (value, run), 4 runs at a time.
// Each run's sums go once into the run's slot of a per-wavefront cell
// cache (plain stores: one wavefront owns its cache and, within a chunk,
// each slot serves one run), or straight to the grid when an earlier run
// of the chunk holds the slot.  A cell reaches the grid (32 global atomics)
// when its slot is taken by another cell and at the end of the wavefront.
constexpr int kStageSlots = 40; // the most that keeps two 4-wave workgroups per CU
constexpr int kStageStride = 65; // doubles per staged row: 64 lanes + 1, so rows start in different banks

struct StagedCache {
  unsigned long long *key; // cell held by each slot (~0: empty)
  double *val;             // kStageSlots x 32 values
  double *stage;           // 16 rows x kStageStride
};

constexpr int staged_cache_bytes() {
  return kStageSlots * (8 + 32 * 8) + 16 * kStageStride * 8;
}

__device__ inline StagedCache staged_at(void *base) {
  StagedCache c;
  c.key = static_cast<unsigned long long *>(base);
  c.val = reinterpret_cast<double *>(c.key + kStageSlots);
  c.stage = c.val + 32 * kStageSlots;
  return c;
}

__device__ inline void staged_init(const StagedCache &c, int lane) {
  if (lane < kStageSlots)
    c.key[lane] = ~0ull;
  wave_lds_sync();
}

__device__ inline int staged_hash(unsigned long long idx) {
  // Fibonacci hash to 32 bits, then multiply-shift onto the slots
  return static_cast<int>(((idx * 0x9E3779B97F4A7C15ull) >> 32) * kStageSlots >> 32);
}

__device__ inline unsigned long long lane_u64(unsigned long long v, int L) {
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(v), L);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(v >> 32), L);
  return (static_cast<unsigned long long>(hi) << 32) | lo;
}

// Value v (corner v >> 2, component v & 3) of cell `cell` into the grid.
__device__ inline void staged_grid_add(double *__restrict__ g, unsigned long long cell, int v,
                                       double x, uint64_t sx, uint64_t sy) {
  const int corner = v >> 2;
  const uint64_t vx = cell + (corner >> 2) * sx + ((corner >> 1) & 1) * sy + (corner & 1);
  grid_atomic_add(g + 4 * vx + (v & 3), x);
}

// The slots in `slots` (bits), holding the cells the wavefront reads from
// c.key, emptied into the grid: two slots per pass, lanes 0-31 and 32-63.
__device__ inline void staged_flush(const StagedCache &c, double *__restrict__ g,
                                    unsigned long long slots, uint64_t sx, uint64_t sy, int lane) {
  while (slots) { // uniform
    const int s0 = __ffsll(slots) - 1;
    slots &= slots - 1;
    const int s1 = slots ? __ffsll(slots) - 1 : -1;
    if (s1 >= 0)
      slots &= slots - 1;
    const int s = lane < 32 ? s0 : s1;
    if (s >= 0) {
      const unsigned long long cell = c.key[s];
      staged_grid_add(g, cell, lane & 31, c.val[32 * s + (lane & 31)], sx, sy);
    }
  }
}

__device__ inline void staged_flush_all(const StagedCache &c, double *__restrict__ g,
                                        uint64_t sx, uint64_t sy, int lane) {
  const unsigned long long k = lane < kStageSlots ? c.key[lane] : ~0ull;
  const unsigned long long held = __ballot(k != ~0ull);
  staged_flush(c, g, held, sx, sy, lane);
  wave_lds_sync();
  if (lane < kStageSlots)
    c.key[lane] = ~0ull;
  wave_lds_sync();
}

// Scatter-add of the trilinear corner contributions of every lane with
// `has` (the transpose of lerp_cell_packed, lanes = consecutive samples of
// one ray; a straight ray meets each cell in one run of them).  All 64
// lanes must call it.
__device__ inline void scatter_staged(const StagedCache &c, double *__restrict__ g, bool has,
                                      uint64_t idx, double xd, double yd, double zd, double g0,
                                      double g1, double g2, double g3, uint64_t sx, uint64_t sy,
                                      int lane) {
  const unsigned long long key = has ? idx : ~0ull;
  const unsigned long long prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || key != prev;
  const unsigned long long heads = __ballot(head);
  const unsigned long long runs = __ballot(head && has);
  if (!runs)
    return; // uniform
  // slots: each run's head looks up its cell's slot; a slot an earlier
  // run of this chunk took sends the later run to the grid
  const int s_own = staged_hash(key);
  const unsigned long long held = (head && has) ? c.key[s_own] : ~0ull;
  unsigned long long direct = 0; // runs (by head lane) that go to the grid
  unsigned long long taken = 0;
  for (unsigned long long t = runs; t; t &= t - 1) { // uniform, scalar
    const int h = __ffsll(t) - 1;
    const int sh = __builtin_amdgcn_readlane(s_own, h);
    if ((taken >> sh) & 1ull)
      direct |= 1ull << h;
    else
      taken |= 1ull << sh;
  }
  const bool mine = head && has && !((direct >> lane) & 1ull);
  const bool fresh = mine && held != key;           // the run claims its slot
  const bool evict = fresh && held != ~0ull;        // ... from another cell
  const unsigned long long fresh_runs = __ballot(fresh);
  unsigned long long ev_slots = 0;
  for (unsigned long long t = __ballot(evict); t; t &= t - 1)
    ev_slots |= 1ull << __builtin_amdgcn_readlane(s_own, __ffsll(t) - 1);
  staged_flush(c, g, ev_slots, sx, sy, lane);
  wave_lds_sync();
  if (fresh)
    c.key[s_own] = key;
  // this lane's corner values (0 without a sample), 16 per half: corners
  // 4 half .. 4 half + 3 (the x = half plane), formed just before they are
  // staged so that only half of them hold registers
  const double wy[2] = {1. - yd, yd}, wz[2] = {1. - zd, zd};
  const int vi = lane & 15, grp = lane >> 4;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    const double wxh = has ? (half ? xd : 1. - xd) : 0.;
    wave_lds_sync(); // the previous half's loads are done before the rows are rewritten
#pragma unroll
    for (int cq = 0; cq < 4; ++cq) {
      const double w = wxh * wy[cq >> 1] * wz[cq & 1];
      c.stage[(4 * cq + 0) * kStageStride + lane] = w * g0;
      c.stage[(4 * cq + 1) * kStageStride + lane] = w * g1;
      c.stage[(4 * cq + 2) * kStageStride + lane] = w * g2;
      c.stage[(4 * cq + 3) * kStageStride + lane] = w * g3;
    }
    wave_lds_sync();
    for (unsigned long long t = runs; t;) { // uniform: 4 runs per round
      int a[4], b[4], len = 0;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        a[j] = b[j] = 0;
        if (t) {
          a[j] = __ffsll(t) - 1;
          t &= t - 1;
          const unsigned long long later = heads & ~((2ull << a[j]) - 1ull);
          b[j] = later ? __ffsll(later) - 1 : 64;
          len = max(len, b[j] - a[j]);
        }
      }
      // the run of this lane's 16-lane group (run heads are uniform)
      unsigned long long cl[4];
#pragma unroll
      for (int j = 0; j < 4; ++j)
        cl[j] = lane_u64(key, a[j]);
      const int ra = grp == 0 ? a[0] : grp == 1 ? a[1] : grp == 2 ? a[2] : a[3];
      const int rb = grp == 0 ? b[0] : grp == 1 ? b[1] : grp == 2 ? b[2] : b[3];
      const unsigned long long cell = grp == 0 ? cl[0] : grp == 1 ? cl[1] : grp == 2 ? cl[2] : cl[3];
      double sum = 0.;
      for (int k = 0; k < len; ++k) {
        const int i = ra + k;
        if (i < rb)
          sum += c.stage[vi * kStageStride + i];
      }
      if (rb > ra) {
        const int val = 16 * half + vi;
        if ((direct >> ra) & 1ull) {
          staged_grid_add(g, cell, val, sum, sx, sy);
        } else {
          double *q = c.val + 32 * staged_hash(cell) + val;
          *q = ((fresh_runs >> ra) & 1ull) ? sum : *q + sum;
        }
      }
    }
  }
  wave_lds_sync();
}

// Per-shell constants of the specialised march (see make_fast).
struct FastConst {
  double psi0, psi1, chi0, b0;
  // psi(r) and chi(z) stay within [-pi/4, pi/4] over the LSA window
  // (checked on the host): their sines and cosines are the series alone
  int lsa_direct;
  double inv_8kpc, inv_kpc;
  double n0, te_r0, te_r02;
  int cre_unif; // 0: analytic profile, 1: uniform sphere
  double R0, inv_cre_r0, inv_cre_z0, cre_r0, cre_r02;
  int cre_z_kpc; // analytic CRE with z0 = 1 kpc: its sech^2 shares tanh's exponential
  double kt, kp;    // folded emissivity factors (include delta and i2bt)
  double pw_e;      // (index + 1) / 2
  double a_coef;    // A = a_coef / |B_perp|
  double fd_scale;  // fd_fore * delta
};

// ------------------------------------------------- specialised sample --
// One sample of the specialised march (LSA + uniform TE + constant-index
// CRE + packed grids), with every intermediate the adjoint needs.  The
// forward kernel reads te, fdk, jt, jp, c2i, s2i; the rest is dead code
// there.  Rewrites relative to the reference (each exact up to rounding):
// cos/sin(2 ipa) from the two projections instead of atan2, folded
// emissivity constants, A^-1 as a product when (index+1)/2 = -1, products by
// reciprocals for divisions by constants that feed no decision.  The
// galactocentric limits, the LSA window, the TE sphere test, the te > 0
// clamp and the grid-cell lookup decide as in the reference, the radial
// limits and the TE / CRE spheres on squared lengths (no square roots: the
// decisions differ only for a sample within 1 ulp of a boundary).
struct FastSample {
  Vec3 pos, B;
  double rr;                 // cylindrical radius
  bool lsa_in;               // 3 kpc <= rr <= 20 kpc
  double cos_p, sin_p, sps, cps, sch, cch, lnr, tnh;
  bool te_in;                // inside the uniform TE sphere
  double te_raw, te;         // before / after the te > 0 clamp
  bool cell_ok;              // inside the random-grid box
  Cell cell;
  double b_par, bper2, bper, prof;
  double x, y, n2;           // IAU-angle projections
  double fdk, jt, jp, c2i, s2i;
};

// MODELS: the regular B and TE come from the descriptor's models (any of
// them, an LSA field with its parameter partials; a uniform TE sphere with
// n0's) instead of being LSA and the uniform sphere.
template <bool GRIDS, bool PW1, bool MODELS = false>
__device__ inline bool eval_fast(const Phys &ph, const FastConst &fc, const Grids &gr,
                                 const ShellArgs &sa, const RayGeo &r, uint64_t k,
                                 FastSample &o) {
  // k < 2^31 (validate_los): one 32-bit conversion
  const double dist =
      sa.d_start + (static_cast<double>(static_cast<uint32_t>(k)) + 0.5) * sa.delta;
  o.pos = Vec3{r.l.x * dist + ph.obs.x, r.l.y * dist + ph.obs.y, r.l.z * dist + ph.obs.z};
  const Vec3 &pos = o.pos;
  const double len2 = pos.x * pos.x + pos.y * pos.y + pos.z * pos.z;
  if (len2 < ph.gc_r_min2 || len2 > ph.gc_r_max2 || pos.z < ph.gc_z_min ||
      pos.z > ph.gc_z_max)
    return false;
  // WMAP-LSA (breg_lsa.cc:9-26)
  Vec3 B{0., 0., 0.};
  // r and 1/r from one reciprocal square root; the LSA window decides on r^2
  const double rr2 = pos.x * pos.x + pos.y * pos.y;
  const double inv_r = rr2 > 0. ? rsqrt_f64(rr2) : 0.;
  o.rr = rr2 * inv_r;
  const bool lsa = !MODELS || ph.breg_model == HAMX_BREG_LSA;
  o.lsa_in = lsa && !(rr2 > (400. * cgs::kpc * cgs::kpc) || rr2 < (9. * cgs::kpc * cgs::kpc));
  if (MODELS && !lsa) {
    if (ph.breg_model == HAMX_BREG_GRID)
      B = gather3(gr.breg_x, gr.breg_y, gr.breg_z, gr.g_breg, pos);
    else
      B = breg_eval<true>(ph.breg_model, ph.breg, pos);
  }
  // exp(-2|z| / kpc): the LSA's tanh(z / kpc) and, when the CRE scale
  // height is 1 kpc, the CRE's sech^2(z / z0)
  const double u = pos.z * fc.inv_kpc;
  const double e_kpc = (o.lsa_in || fc.cre_z_kpc) ? exp_fast(-2. * fabs(u)) : 0.;
  if (o.lsa_in) {
    o.cos_p = pos.x * inv_r;
    o.sin_p = pos.y * inv_r;
    o.lnr = log_pos(o.rr * fc.inv_8kpc);
    // tanh(u) = (1 - e) / (1 + e) with e = exp(-2|u|): the absolute error
    // stays at the rounding of 1 - e (1e-16), which is what chi needs
    o.tnh = copysign((1. - e_kpc) * rcp_f64(1. + e_kpc), u);
    const double psi = fc.psi0 + fc.psi1 * o.lnr;
    const double chi = fc.chi0 * o.tnh;
    if (fc.lsa_direct) {
      sincos_series(psi, o.sps, o.cps);
      sincos_series(chi, o.sch, o.cch);
    } else {
      sincos(psi, &o.sps, &o.cps);
      sincos(chi, &o.sch, &o.cch);
    }
    const double bc = fc.b0 * o.cch;
    B.x = bc * (o.sps * o.cos_p - o.cps * o.sin_p);
    B.y = bc * (o.sps * o.sin_p + o.cps * o.cos_p);
    B.z = fc.b0 * o.sch;
  }
  // uniform TE sphere (tereg_unif.cc:6-12)
  double te = 0.;
  if (!MODELS || ph.tereg_model == HAMX_TEREG_UNIF) {
    const Vec3 d{pos.x - ph.obs.x, pos.y - ph.obs.y, pos.z - ph.obs.z};
    o.te_in = !(d.x * d.x + d.y * d.y + d.z * d.z > fc.te_r02);
    if (o.te_in)
      te = fc.n0;
  } else {
    o.te_in = false; // no n0 partial
    if (ph.tereg_model == HAMX_TEREG_YMW16)
      te = tereg_ymw16_eval(ph.tereg, pos.x, pos.y, pos.z);
    else if (ph.tereg_model == HAMX_TEREG_GRID)
      te = gather1(gr.tereg, gr.g_tereg, pos);
  }
  o.cell_ok = false;
  if (GRIDS) {
    if (grid_cell(gr.g_packed, pos, o.cell)) {
      o.cell_ok = true;
      Vec3 br;
      double tr;
      lerp_cell_packed(gr.packed, o.cell, gr.g_packed.nz, gr.g_packed.ny * gr.g_packed.nz,
                       br, tr);
      B.x += br.x;
      B.y += br.y;
      B.z += br.z;
      te += tr;
    }
  }
  o.te_raw = te;
  te *= static_cast<double>(te > 0.);
  o.te = te;
  o.B = B;
  o.b_par = dot0(B, r.l);
  o.fdk = te * o.b_par * fc.fd_scale;
  // perpendicular field and emissivities
  const Vec3 cr{B.y * r.l.z - B.z * r.l.y, B.z * r.l.x - B.x * r.l.z,
                B.x * r.l.y - B.y * r.l.x};
  o.bper2 = cr.x * cr.x + cr.y * cr.y + cr.z * cr.z;
  o.bper = sqrt(o.bper2);
  if (fc.cre_unif) {
    const Vec3 d{pos.x - ph.obs.x, pos.y - ph.obs.y, pos.z - ph.obs.z};
    o.prof = (d.x * d.x + d.y * d.y + d.z * d.z > fc.cre_r02) ? 0. : 1.;
  } else {
    // sech^2(z / z0) = 4 e / (1 + e)^2 with e = exp(-2 |z| / z0) <= 1
    const double e = fc.cre_z_kpc ? e_kpc : exp_fast(-2. * fabs(pos.z * fc.inv_cre_z0));
    const double d = 1. + e;
    o.prof = exp_fast((fc.R0 - o.rr) * fc.inv_cre_r0) * (4. * e) * rcp_f64(d * d);
  }
  double w;
  if (PW1)
    w = o.prof * o.bper2;
  else
    w = o.prof * o.bper * pow(fc.a_coef / o.bper, fc.pw_e);
  o.jt = w * fc.kt;
  o.jp = w * fc.kp;
  // cos / sin of twice the IAU angle (integrator.cc:257-265)
  o.x = -dot0(r.eph, B);
  o.y = -dot0(r.eth, B);
  o.n2 = o.x * o.x + o.y * o.y;
  if (o.n2 > 0.) {
    const double inv = rcp_f64(o.n2);
    o.c2i = (o.x * o.x - o.y * o.y) * inv;
    o.s2i = 2. * o.x * o.y * inv;
  } else {
    o.c2i = 1.;
    o.s2i = 0.;
  }
  return true;
}

// ------------------------------------------------------------ host side --

// digamma(x): recurrence up to x >= 6, then the asymptotic series in 1/x^2
// (Abramowitz & Stegun 6.3.18); reflection for x <= 0.  d/dx ln Gamma(x),
// for the spectral-index derivative of the emissivity Gamma factors.
inline double digamma(double x) {
  if (x <= 0.)
    return digamma(1. - x) - cgs::pi / std::tan(cgs::pi * x);
  double r = 0.;
  while (x < 6.) {
    r -= 1. / x;
    x += 1.;
  }
  const double f = 1. / (x * x);
  const double series =
      f * (1. / 12. - f * (1. / 120. - f * (1. / 252. - f * (1. / 240. - f * (1. / 132.)))));
  return r + std::log(x) - 0.5 / x - series;
}
bool same_box(const hamx_box &a, const hamx_box &b);
int check_box(hamx_ctx *ctx, const hamx_box &b, const char *what);
inline uint64_t box_size(const hamx_box &b) { return b.nx * b.ny * b.nz; }
// Integrator::assemble_shell_ref (integrator.cc:231-244)
void shell_ref(const hamx_los_desc *d, uint32_t s, uint64_t &step, double &d_start,
               double &delta);
int validate_los(hamx_ctx *ctx, const hamx_los_desc *d);
Phys make_phys(const hamx_los_desc *d);
FastConst make_fast(const hamx_los_desc *d, const Phys &ph, double delta, bool pw1);
// A fixed device buffer handed out in call order at 256-byte alignment
// (the scratch of hamx_los_forward_static).  Nothing is allocated: a request
// beyond `cap` fails.
struct Arena {
  char *base;
  size_t cap, used;
};
// Device buffer of `bytes`: carved from `ar` when it is non-null, else the
// context's grow-only workspace slot `slot` (which may allocate).
int scratch_get(hamx_ctx *ctx, Arena *ar, int slot, size_t bytes, void **out);
// Stage the field grids of a call on the device (uploading host inputs into
// the workspace) and pack brnd + ternd into one double4 record per vertex
// when they share a box.  With an arena (device inputs only) the grids are
// gathered where they lie and nothing is packed.
int stage_fields(hamx_ctx *ctx, const hamx_los_desc *d, const hamx_fields &f,
                 bool in_dev, Grids &gr, bool &packed, bool &grids_any,
                 Arena *ar = nullptr);
// First pixel of shell s in a buffer holding one value per ray of every
// shell: sum over t < s of 12 nside_t^2.
inline uint64_t shell_offset(const hamx_los_desc *d, uint32_t s) {
  uint64_t off = 0;
  for (uint32_t t = 0; t < s; ++t)
    off += 12ull * d->nside_shell[t] * d->nside_shell[t];
  return off;
}
// One write_grid pass (hamx_los_forward); when ckpt is non-null and FD is
// carried, the accumulated FD map before shell s is copied to
// ckpt + s * npix(fd map); when shell_fd is non-null each ray's FD over
// shell s is written to shell_fd + shell_offset(d, s) (device memory).
// With an arena every scratch buffer (shell maps, the mask at the shell's
// nside, output maps the caller left NULL) is carved from it before the
// first launch, so an arena too small fails before any work is queued.
int los_forward_impl(hamx_ctx *ctx, const hamx_los_desc *d, const hamx_fields *fields,
                     hamx_maps *out, uint32_t flags, double *ckpt, double *shell_fd,
                     Arena *ar = nullptr);
// The caller's mask map (nside of hp) at the nside of ht, into out
// (Hampisk::duplicate), queued on the context stream.
int mask_at_nside(hamx_ctx *ctx, double *out, const Healpix &ht, const double *piv,
                  const Healpix &hp);
// hamx_los_forward recording the adjoint's tape in the context workspace.
int los_forward_taped(hamx_ctx *ctx, const hamx_los_desc *d, const hamx_fields *fields,
                      hamx_maps *out, uint32_t flags);

} // namespace hamx
