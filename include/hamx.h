/*
 * hamx.h -- C-ABI of the MI355X (gfx950) line-of-sight integrator and
 * Gaussian random-field generators of hammurabiX.
 *
 * Plain C, POD structs, pointers and sizes only.  Every entry point returns
 * HAMX_OK (0) or a negative HAMX_ERR_* code; hamx_last_error() gives the
 * message.  The host adapters (C++ Integrator/Brnd_es/TErnd_dft mirrors and
 * the Python hampyx mirror) turn a negative code into the exception the
 * reference throws (std::runtime_error / RuntimeError).
 *
 * Units are the reference's CGS units after Param parsing: lengths in cm,
 * magnetic fields in Gauss, frequencies in Hz, energies in erg, angles in rad,
 * thermal-electron densities in cm^-3, wave numbers in 1/kpc.
 *
 * Grids are the reference's vertex grids: x slowest, z fastest,
 * flat index i*ny*nz + j*nz + l (toolkit.h:29-34).
 * Maps are HEALPix RING ordered, npix = 12*nside^2 doubles.
 */
#ifndef HAMX_H
#define HAMX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HAMX_ABI_VERSION 2
#define HAMX_MAX_SHELL 16

/* ---- error codes ------------------------------------------------------ */
#define HAMX_OK 0
#define HAMX_ERR_ARG (-1)         /* null pointer, bad size, inconsistent descriptor */
#define HAMX_ERR_UNSUPPORTED (-2) /* model or option without a device implementation */
#define HAMX_ERR_DEVICE (-3)      /* HIP runtime error */
#define HAMX_ERR_NOMEM (-4)       /* device allocation failed */
#define HAMX_ERR_NSIDE (-5)       /* HEALPix nside not a power of two */

/* ---- buffer-residency flags ------------------------------------------- */
#define HAMX_IN_DEVICE 1u  /* every input array pointer is a device pointer */
#define HAMX_OUT_DEVICE 2u /* every output array pointer is a device pointer */
#define HAMX_KEEP_TAPE 8u  /* hamx_los_forward: keep this pass's tape (the FD map
                              before each shell and each ray's FD per shell) in the
                              context for hamx_los_adjoint */
#define HAMX_USE_TAPE 16u  /* hamx_los_adjoint: use the kept tape instead of
                              re-running the forward; the descriptor and field
                              pointers must equal the taped pass's */

/* ---- model selectors (the Param type strings) ------------------------- */
/* regular magnetic field, Param::breg_type (param.cc:161-398) */
#define HAMX_BREG_NONE 0 /* <regular cue="0">: Breg base class, zero field */
#define HAMX_BREG_LSA 1  /* "lsa":  breg[0..3] = b0 [G], psi0, psi1, chi0 [rad] */
#define HAMX_BREG_CART 2 /* "cart": breg[0..2] = bx, by, bz [G] */
#define HAMX_BREG_UNIF 3 /* "unif": breg[0..2] = bp, bv [G], l0 [rad] */
#define HAMX_BREG_GRID 4 /* <fieldio><breg read="1">: trilinear gather of fields.breg_* */
#define HAMX_BREG_HELIX 5 /* "helix": breg[0..4] = bx, by, bz [G], r_min, r_max [cm]
                             (breg_helix.cc:13-30) */
#define HAMX_BREG_JF12 6  /* "jf12": breg[0..6] = b_arm1..7, [7] b_ring [G], [8] h_disk,
                             [9] w_disk [cm], [10] Bn, [11] Bs [G], [12] rn, [13] rs,
                             [14] wh, [15] z0 [cm], [16] B0_X [G], [17] Xtheta [rad],
                             [18] rpc_X, [19] r0_X [cm] (breg_jf12.cc:8-189) */
#define HAMX_BREG_JAFFE 7 /* "jaffe" (breg_jaffe.cc:11-243): breg[0] quadruple, [1] bss,
                             [2] disk_amp [G], [3] disk_z0 [cm], [4] halo_amp [G],
                             [5] halo_z0, [6] r_inner, [7] r_scale, [8] r_peak [cm],
                             [9] ring, [10] bar (0 / 1), [11] ring_amp [G], [12] ring_r [cm],
                             [13] bar_amp [G], [14] bar_a, [15] bar_b [cm], [16] bar_phi0
                             [rad], [17] arm_num (<= 4), [18] arm_r0, [19] arm_z0 [cm],
                             [20..23] arm_phi0 [rad], [24..27] arm_amp [G], [28] arm_pitch
                             [rad], [29] comp_c, [30] comp_d, [31] comp_r [cm], [32] comp_p */
#define HAMX_BREG_NPARAM 40 /* parameters per regular-field model */
/* regular thermal electrons, Param::tereg_type (param.cc:462-667) */
#define HAMX_TEREG_NONE 0 /* zero density */
#define HAMX_TEREG_UNIF 1 /* "unif": tereg[0..1] = n0 [cm^-3], r0 [cm] */
#define HAMX_TEREG_GRID 2 /* <fieldio><tereg read="1">: gather of fields.tereg */
#define HAMX_TEREG_YMW16 3 /* "ymw16" (tereg_ymw16.cc:10-333, param.cc:480-641): tereg[0..72] =
     [0] r_warp [cm], [1] r0 [cm], [2] gamma_w;  thick disk [3] ad, [4] bd [cm], [5] n1 [cm^-3], [6] h1 [cm];
     thin disk [7] a2, [8] b2 [cm], [9] n2, [10] k2;  spiral arms [11] b2s [cm], [12..16] ele_arm_0..4,
     [17..21] wid_arm_0..4 [cm], [22..26] rref_arm_0..4 [cm], [27..31] phiref_arm_0..4 [rad],
     [32..36] tan(pitch_arm_0..4), [37..41] cos(pitch_arm_0..4), [42] aa [cm], [43] ka, [44] ncn,
     [45] thetacn [deg], [46] wcn [deg], [47] nsg, [48] thetasg [deg], [49] wsg [deg];
     Galactic centre [50] ngc, [51] agc, [52] hgc [cm];  Gum [53] ngn, [54] wgn, [55] agn [cm], [56] kgn;
     Local Bubble [57] j_lb, [58] nlb1, [59] thetalb1 [deg], [60] detlb1 [deg], [61] wlb1, [62] hlb1 [cm],
     [63] nlb2, [64] thetalb2, [65] detlb2 [deg], [66] wlb2, [67] hlb2 [cm];
     Loop I [68] nli, [69] rli, [70] wli [cm], [71] detthetali, [72] thetali [deg] */
#define HAMX_TEREG_NPARAM 80
/* cosmic-ray electrons, Param::cre_type (param.cc:731-778) */
#define HAMX_CRE_NONE 0 /* no CRE model: sync output is rejected */
#define HAMX_CRE_ANA 1  /* "analytic": cre[0..6] = alpha, beta, theta, r0 [cm], z0 [cm], E0 [erg], j0 */
#define HAMX_CRE_UNIF 2 /* "unif":     cre[0..3] = alpha, r0 [cm], E0 [erg], j0 */
#define HAMX_CRE_GRID 3 /* <fieldio><cre read="1">: CRE_num over hamx_fields.cre (cre.cc:144-199);
                           the synchrotron emissivities integrate the flux over the energy axis
                           with the kernels F and G (integrator.cc:281-332,350-401) */

/* ---- LOS march schedules ---------------------------------------------- */
#define HAMX_SCHED_AUTO 0   /* wavefront per ray */
#define HAMX_SCHED_THREAD 1 /* one thread per ray, samples summed in the reference order */
#define HAMX_SCHED_WAVE 2   /* one wavefront per ray, lanes over consecutive steps, FD by wave scan */

/* ---- descriptors ------------------------------------------------------- */

/* Cartesian vertex box (Param::grid_* nx..z_max, param.h:32-34). */
typedef struct hamx_box {
  uint64_t nx, ny, nz;
  double x_min, x_max, y_min, y_max, z_min, z_max; /* cm */
} hamx_box;

/* Field grids gathered by the march.  NULL arrays mean "absent".
 * brnd_* is Grid_brnd::{bx,by,bz} (grid.h:94), ternd is Grid_ternd::te
 * (grid.h:144), breg_* / tereg are the imported regular-field grids used by
 * HAMX_BREG_GRID / HAMX_TEREG_GRID. */
typedef struct hamx_fields {
  const double *brnd_x, *brnd_y, *brnd_z;
  hamx_box brnd_box;
  const double *ternd;
  hamx_box ternd_box;
  const double *breg_x, *breg_y, *breg_z;
  hamx_box breg_box;
  const double *tereg;
  hamx_box tereg_box;
  /* Grid_obs mask map (grid_obs.cc:44-50) at nside_mask, or NULL.  Each
   * shell uses it duplicated to its own nside (Hampisk::duplicate,
   * hamsk.h:55-103), and a pixel is integrated only where that map equals
   * 1.0; the others yield zero (integrator.cc:63-70). */
  const double *mask;
  uint32_t nside_mask;
  uint32_t reserved;
  /* Grid_cre::cre_flux (grid.h:167) for HAMX_CRE_GRID: differential flux in
   * [GeV m^2 s sr]^-1 on the vertex box cre_box and the energy axis
   * E_i = cre_E_min * exp(i * cre_E_fact) [erg], i < cre_nE, stored
   * [x][y][z][E] with E fastest (toolkit::index4d as cre.cc:152-199 uses it). */
  const double *cre;
  hamx_box cre_box;
  double cre_E_min, cre_E_fact;
  uint32_t cre_nE;
  uint32_t reserved2;
} hamx_fields;

/* One Integrator::write_grid pass (integrator.cc:23-134) for one <sync>
 * entry (the one Param::grid_obs.*.back() names). */
typedef struct hamx_los_desc {
  double observer[3]; /* Param::observer, cm */
  uint32_t n_shell;   /* Param::grid_obs.total_shell */
  uint32_t nside_shell[HAMX_MAX_SHELL];
  double radii_shell[HAMX_MAX_SHELL + 1]; /* cm, n_shell+1 entries */
  double oc_r_res;                        /* cm */
  double gc_r_min, gc_r_max, gc_z_min, gc_z_max; /* cm */
  int32_t do_dm, do_fd, do_sync;
  uint32_t nside_dm, nside_fd, nside_sync;
  double sync_freq; /* Hz */
  int32_t breg_model;
  double breg[HAMX_BREG_NPARAM];
  int32_t tereg_model;
  double tereg[HAMX_TEREG_NPARAM];
  int32_t cre_model;
  double cre[8];
  int32_t schedule; /* HAMX_SCHED_* */
  int32_t reserved;
} hamx_los_desc;

/* Output maps (Grid_obs::{dm,fd,is,qs,us}_map after write_grid, grid.h:188).
 * Sizes: dm npix(nside_dm); is/qs/us npix(nside_sync); fd npix(nside_fd) when
 * do_fd, else npix(nside_sync).  Unused maps may be NULL.  Values are the
 * raw write_grid results: DM in cm^-3 cm and FD in rad cm^-2 (the export
 * rescale of grid_obs.cc:57,78 is the caller's), I/Q/U in K (CMB).
 * Every map is overwritten. */
typedef struct hamx_maps {
  double *dm, *fd, *is, *qs, *us;
} hamx_maps;

/* Brnd_es generator (brnd_es.cc:76-308) on box `box`.  The anisotropy
 * direction uses the regular field model breg_model/breg (or the grid in
 * `breg_grid_*` for HAMX_BREG_GRID). */
typedef struct hamx_es_desc {
  hamx_box box;
  double observer[3];
  double rms;          /* Gauss */
  double k0, k1;       /* 1/kpc */
  double a0, a1, rho;  /* spectral indices, anisotropy ratio */
  double r0, z0;       /* energy profile scales, cm */
  int32_t breg_model;
  int32_t reserved;
  double breg[HAMX_BREG_NPARAM];
  const double *breg_x, *breg_y, *breg_z;
  hamx_box breg_box;
  /* Brnd_jf12 energy profile (hamx_brnd_jf12 only; brnd_jf12.cc:54-114):
   * b0_1..b0_8, b0_int [G], z0_spiral [cm], b0_halo [G], r0_halo, z0_halo [cm] */
  double jf12[13];
} hamx_es_desc;

/* Brnd_mhd generator (brnd_mhd.cc:16-262) on box `box`: Alfven, fast and
 * slow modes about the regular field at the observer (breg_model / breg, or
 * the grid in breg_* for HAMX_BREG_GRID), Param::brnd_mhd (param.cc:424-440). */
typedef struct hamx_mhd_desc {
  hamx_box box;
  double observer[3];
  double pa0, pf0, ps0;     /* mode power norms, G^2 */
  double k0, k1;            /* 1/kpc */
  double aa0, af0, as0, a1; /* spectral indices */
  double beta, ma;          /* plasma beta, Alfven Mach number */
  int32_t breg_model;
  int32_t reserved;
  double breg[HAMX_BREG_NPARAM];
  const double *breg_x, *breg_y, *breg_z;
  hamx_box breg_box;
} hamx_mhd_desc;

/* TErnd_dft generator (ternd_dft.cc:35-128) on box `box`. */
typedef struct hamx_dft_desc {
  hamx_box box;
  double observer[3];
  double rms; /* cm^-3 */
  double k0;  /* 1/kpc */
  double a0;
  double r0, z0; /* cm */
} hamx_dft_desc;

/* ---- context ----------------------------------------------------------- */
typedef struct hamx_ctx hamx_ctx;

/* Open a context on HIP device `device`: one stream of default priority and
 * a grow-only device workspace.  One context per host thread. */
int hamx_open(int device, hamx_ctx **out);
void hamx_close(hamx_ctx *ctx);
const char *hamx_last_error(const hamx_ctx *ctx);
int hamx_abi_version(void);
/* Synchronise the context stream. */
int hamx_sync(hamx_ctx *ctx);

/* Kernel timing on the context stream (HIP events around each launch of the
 * march and accumulate kernels).  hamx_profile(ctx, 1) clears the record
 * and starts recording; hamx_profile_read copies up to `max` records. */
typedef struct hamx_kernel_time {
  char kernel[32]; /* "los_wave", "los_thread", "accumulate", ... */
  uint32_t shell;  /* shell index of a march / accumulate launch */
  uint32_t reserved;
  uint64_t work;   /* geometric samples of a march, pixels otherwise */
  double ms;       /* launch duration */
} hamx_kernel_time;
int hamx_profile(hamx_ctx *ctx, int enable);
int hamx_profile_read(hamx_ctx *ctx, hamx_kernel_time *out, int max, int *count);

/* Device memory helpers for callers without their own allocator (the C++
 * host adapter keeps grids resident across frequency passes with these). */
int hamx_dev_alloc(hamx_ctx *ctx, size_t bytes, void **dptr);
int hamx_dev_free(hamx_ctx *ctx, void *dptr);
int hamx_memcpy_h2d(hamx_ctx *ctx, void *dst, const void *src, size_t bytes);
int hamx_memcpy_d2h(hamx_ctx *ctx, void *dst, const void *src, size_t bytes);

/* ---- LOS integrator ----------------------------------------------------- */

/* One Integrator::write_grid pass: all shells in order, FD carried from the
 * inner shells through fd_map (integrator.cc:78-86,130-132), shell maps
 * temp_convert-ed (integrator.cc:102-112) and accumulated into the output
 * maps with HEALPix interpolation across nside (hamdis.h:698-721).
 * Replaces Integrator::write_grid (integrator.h:25-28, integrator.cc:23-134).
 * flags: HAMX_IN_DEVICE / HAMX_OUT_DEVICE.  Synchronous. */
int hamx_los_forward(hamx_ctx *ctx, const hamx_los_desc *d,
                     const hamx_fields *fields, hamx_maps *out, uint32_t flags);

/* hamx_los_forward without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Inputs and
 * outputs are device pointers (flags must be exactly HAMX_IN_DEVICE |
 * HAMX_OUT_DEVICE) and every output map of the pass must be given.  The shell
 * maps and the mask at the shell's nside live in one fixed device array of
 * HAMX_STATIC_SCRATCH_BYTES per GPU (held until the process ends), so shells
 * are limited to nside HAMX_STATIC_MAX_NSIDE; a CRE grid and the tape are not
 * supported, and random B and TE grids are gathered where they lie (not
 * packed).  Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (from several contexts or threads) race.  Synchronous. */
#define HAMX_STATIC_MAX_NSIDE 256
#define HAMX_STATIC_SCRATCH_BYTES (40 * 1024 * 1024)
int hamx_los_forward_static(hamx_ctx *ctx, const hamx_los_desc *d,
                            const hamx_fields *fields, hamx_maps *out, uint32_t flags);

/* Reverse mode of hamx_los_forward.  Given adj = dL/d(maps) for a scalar L
 * of the pass's output maps (same sizes as the forward maps; NULL maps are
 * zero), returns
 *   dparams[HAMX_NPARAM]: dL/d(b0, psi0, psi1, chi0) of the LSA field,
 *                         dL/d(alpha) of the CRE model, dL/d(n0) of the
 *                         uniform thermal electrons (CGS units of Param);
 *   dbrnd_x/y/z         : dL/d(random-B grid), grid layout (may be NULL);
 *   dternd              : dL/d(random-TE grid) (may be NULL).
 * Every output is overwritten.  The adjoint needs the forward's tape (the
 * accumulated FD map before each shell and every ray's FD over each
 * shell): with HAMX_USE_TAPE it takes the one a hamx_los_forward call with
 * HAMX_KEEP_TAPE left in the context, otherwise it re-runs the forward to
 * record it.  Shells are then processed outermost first, each ray in one
 * sweep from its far end (running FD = ray total minus the samples beyond).
 * Supported: every regular B model and TE model of hamx_los_forward
 * (dL/d(b0, psi0, psi1, chi0) is zero unless B is LSA, dL/d(n0) unless TE
 * is the uniform sphere), a mask (masked-out rays contribute nothing), the
 * analytic or uniform CRE with a constant spectral index, synchrotron on,
 * random grids (if any) sharing one box.  No counterpart in the reference,
 * which has no gradients.  flags as for hamx_los_forward (inputs: fields
 * and adj; outputs: dparams and the grid gradients). */
#define HAMX_NPARAM 6
int hamx_los_adjoint(hamx_ctx *ctx, const hamx_los_desc *d,
                     const hamx_fields *fields, const hamx_maps *adj,
                     double *dparams, double *dbrnd_x, double *dbrnd_y,
                     double *dbrnd_z, double *dternd, uint32_t flags);

/* Geometric LOS samples of one pass, S_geo = sum_s npix_s * step_s
 * (integrator.cc:62,173 trip counts); host-side arithmetic only. */
int hamx_los_samples(const hamx_los_desc *d, uint64_t *samples);

/* ---- batched LOS -------------------------------------------------------- */

/* Many instances (random-field realisations and / or parameter sets, each
 * with its own shells, output nsides and frequency list) integrated by one
 * replay of a captured HIP graph.  Each instance reproduces
 * Pipeline::assemble_obs (pipeline.cc:165-190): one write_grid pass per
 * synchrotron frequency, the last-listed frequency first with DM and FD
 * (FD carried through the shells when los.do_fd), the other frequencies
 * without DM / FD and with a zero FD cache.  The passes of one instance
 * share every per-sample field evaluation; only the polarisation phase and
 * the emissivity constants differ per frequency.  No counterpart in the
 * reference, which runs one pass per process call. */
#define HAMX_MAX_FREQ 4
#define HAMX_PREC_F64 0u /* FP64 samples: the plain path's arithmetic */
#define HAMX_PREC_F32 1u /* FP32 field / emissivity arithmetic and FP32 grids;
                            sample positions, the limit tests, FD accumulation,
                            the polarisation phase and map sums in FP64 */

typedef struct hamx_batch_inst {
  hamx_los_desc los;          /* LSA B, uniform TE, constant-index CRE; sync_freq and
                                 schedule are ignored (freq[] below) */
  uint32_t n_freq;            /* synchrotron frequencies, 1..HAMX_MAX_FREQ */
  uint32_t reserved;
  double freq[HAMX_MAX_FREQ]; /* Hz, in the order of the <sync> list */
  const double *brnd_x, *brnd_y, *brnd_z; /* device random-B grid on box, or NULL */
  const double *ternd;                    /* device random-TE grid on box, or NULL */
  hamx_box box;
  double *dm, *fd;            /* device output maps (nside_dm / nside_fd), or NULL */
  double *is[HAMX_MAX_FREQ], *qs[HAMX_MAX_FREQ], *us[HAMX_MAX_FREQ]; /* nside_sync */
  /* Reverse mode (batches created with HAMX_BATCH_ADJOINT), device arrays:
   * inputs dL/d(maps) of the same sizes as the maps above (NULL: zero) ... */
  const double *adj_dm, *adj_fd;
  const double *adj_is[HAMX_MAX_FREQ], *adj_qs[HAMX_MAX_FREQ], *adj_us[HAMX_MAX_FREQ];
  /* ... and outputs: dL/d(b0, psi0, psi1, chi0, alpha, n0) (HAMX_NPARAM values)
   * and dL/d(random grids) on box (NULL: not wanted).  Instances naming the
   * same grid-gradient arrays get the sum of their gradients. */
  double *dparams;
  double *dbrnd_x, *dbrnd_y, *dbrnd_z, *dternd;
} hamx_batch_inst;

typedef struct hamx_batch hamx_batch;

typedef struct hamx_batch_info {
  uint64_t samples_geo;  /* geometric samples of one pass, summed over instances */
  uint64_t samples_ref;  /* samples_geo of each instance times its n_freq */
  uint32_t levels;       /* shell levels (march launches per replay) */
  uint32_t variant;      /* launch shape chosen by the autotuner */
  uint32_t n_variants;   /* candidates timed */
  uint32_t reserved;
  double variant_ms[16]; /* one untimed-graph replay per candidate, ms */
  char shape[64];        /* the chosen launch shape, readable */
} hamx_batch_info;

/* Plan a batch: validates every instance, allocates the packed grids and
 * shell maps, times the candidate launch shapes (persistent or one-shot,
 * rays per work item, XCD-aware placement) on the batch itself, and
 * captures the chosen one in a HIP graph.  Input grids are re-read and
 * output maps overwritten by every hamx_batch_run; their addresses are
 * fixed at creation.  flags: HAMX_IN_DEVICE | HAMX_OUT_DEVICE (required),
 * plus HAMX_BATCH_NO_TUNE to skip the timing and take the first shape,
 * HAMX_BATCH_LANES to skip it and take the forward with one ray per lane
 * (HAMX_PREC_F32 only), and HAMX_BATCH_ADJOINT to record the forward's tape
 * on every replay and capture the reverse-mode graph as well. */
#define HAMX_BATCH_NO_TUNE 4u
#define HAMX_BATCH_ADJOINT 32u
#define HAMX_BATCH_LANES 64u
int hamx_batch_create(hamx_ctx *ctx, const hamx_batch_inst *inst, uint32_t n_inst,
                      uint32_t precision, uint32_t flags, hamx_batch **out);
/* Replay the captured graph once on the context's stream.  Synchronous. */
int hamx_batch_run(hamx_batch *b);
/* Replay the reverse-mode graph: the gradients of every instance from its
 * adj_* maps, on the tape of the last hamx_batch_run.  Shells run outermost
 * first, each ray in one sweep from its far end, as hamx_los_adjoint.
 * Synchronous.  No counterpart in the reference, which has no gradients. */
int hamx_batch_run_adjoint(hamx_batch *b);
int hamx_batch_get_info(const hamx_batch *b, hamx_batch_info *info);
void hamx_batch_destroy(hamx_batch *b);

/* ---- random fields ------------------------------------------------------ */

/* Brnd_es::write_grid (bfield.h:279-280, brnd_es.cc:76-308): spectrum draw,
 * backward FFT, energy-profile rescale with LSA anisotropy, forward FFT,
 * Gram-Schmidt divergence cleaning, backward FFT, Hermitian extraction /N.
 * Gaussian draws come from Philox4x32-10 keyed by `seed` (a counter-based
 * stream: one realisation per seed, independent of thread count).
 * Outputs bx, by, bz: nx*ny*nz doubles each. */
int hamx_brnd_es(hamx_ctx *ctx, const hamx_es_desc *d, uint64_t seed,
                 double *bx, double *by, double *bz, uint32_t flags);

/* Brnd_mhd::write_grid (bfield.h:380-381, brnd_mhd.cc:16-167): k-space
 * modes with one backward transform per complex array and the real parts
 * extracted; Gaussian and uniform variates from the same counter-based
 * stream as hamx_brnd_es. */
int hamx_brnd_mhd(hamx_ctx *ctx, const hamx_mhd_desc *d, uint64_t seed,
                  double *bx, double *by, double *bz, uint32_t flags);

/* CREfield::write_grid (crefield.h:53, cre.cc:204-225) for the analytic and
 * uniform models: out[((i*ny + j)*nz + k)*nE + m] = write_field(pos_ijk, E_m)
 * (cre_ana.cc:52-63, cre_unif.cc:48-59) on the vertex box `box` with
 * E_m = E_min * exp(m * E_fact).  The model and observer come from `d`
 * (cre_model HAMX_CRE_ANA or HAMX_CRE_UNIF).  flags: HAMX_OUT_DEVICE. */
int hamx_cre_write_grid(hamx_ctx *ctx, const hamx_los_desc *d, const hamx_box *box,
                        double E_min, double E_fact, uint32_t nE, double *out,
                        uint32_t flags);

/* The synchrotron kernels of the grid emissivities, as the march evaluates
 * them: F[i] = x int_x^inf K_5/3(t) dt and G[i] = x K_2/3(x) at x[i] in
 * (0, 100] (replacing gsl_sf_synchrotron_1 / _2, integrator.cc:321,390).
 * flags: HAMX_IN_DEVICE | HAMX_OUT_DEVICE as for the other calls. */
int hamx_sync_kernels(hamx_ctx *ctx, const double *x, uint64_t n, double *F, double *G,
                      uint32_t flags);

/* Brnd_jf12::write_grid (bfield.h:329-330, brnd_jf12.cc:127-385): the same steps as Brnd_es with
 * the JF12 energy profile (d->jf12; d->r0 / z0 unused) and the reference's
 * literal constants of that file (0.33333333, 0.66666667, 1.73205081). */
int hamx_brnd_jf12(hamx_ctx *ctx, const hamx_es_desc *d, uint64_t seed,
                   double *bx, double *by, double *bz, uint32_t flags);

/* TErnd_dft::write_grid (tefield.h:159-160, ternd_dft.cc:35-128).
 * Output te: nx*ny*nz doubles. */
int hamx_ternd_dft(hamx_ctx *ctx, const hamx_dft_desc *d, uint64_t seed,
                   double *te, uint32_t flags);

/* In-place unnormalised 3D c2c FFT of interleaved complex128 data
 * (re,im pairs, x slowest), sign = +1 (FFTW_BACKWARD) or -1 (FFTW_FORWARD),
 * the transform of the reference's plans (grid_brnd.cc:42-54).
 * Each of nx, ny, nz must be a power of two in [1, 4096]. */
int hamx_fft3d(hamx_ctx *ctx, double *data, uint64_t nx, uint64_t ny,
               uint64_t nz, int sign, uint32_t flags);

/* The Gaussian draws of the spectrum step: out[2*m], out[2*m+1] are the two
 * N(0,1) variates of draw pair m of stream `stream` (0: Brnd_es c0/c1,
 * 1: TErnd_dft te_k) for mode index `mode` = m / pairs_per_mode.  Exposed so
 * that a CPU restatement can be checked element-wise.  count = pairs. */
int hamx_gauss_pairs(hamx_ctx *ctx, uint64_t seed, uint32_t stream,
                     uint64_t first_pair, uint64_t count, double *out,
                     uint32_t flags);

#ifdef __cplusplus
}
#endif

#endif /* HAMX_H */
