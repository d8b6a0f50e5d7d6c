/*
 * hamx_oracle.h -- CPU restatement of the hammurabiX LOS integrator and
 * random-field generators.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against (tests/, __graft_entry__.smoke(), bench.py's CPU
 * baseline).  Nothing in the product links or calls it.
 *
 * Each function names the reference file:line it restates.  Descriptor
 * types come from include/hamx.h so the tests feed identical inputs to the
 * oracle and to the device path.
 */
#ifndef HAMX_ORACLE_H
#define HAMX_ORACLE_H

#include <stdint.h>

#include "hamx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- HEALPix (hamdis.h) ---- */
void orc_pix2ang(uint64_t nside, uint64_t ipix, double *theta, double *phi);
void orc_pix2ang_map(uint64_t nside, double *theta_phi /* 2*npix */);
/* Hampix::rpixxyf / xyfrpix (hamdis.h:515-605) */
void orc_ring2xyf(uint64_t nside, uint64_t pix, int64_t *ix, int64_t *iy, int *face);
uint64_t orc_xyf2ring(uint64_t nside, int64_t ix, int64_t iy, int face);
/* Hampisk::duplicate (hamsk.h:55-103) */
void orc_mask_duplicate(const double *piv, uint64_t piv_nside, uint64_t nside, double *out);
double orc_interpolate(const double *map, uint64_t nside, double theta,
                       double phi);
void orc_accumulate(double *dst, uint64_t nside_dst, const double *src,
                    uint64_t nside_src);

/* ---- integrator helpers (integrator.cc / integrator.h) ---- */
void orc_shell_ref(double d_start, double d_stop, double res, uint64_t *step,
                   double *delta);
void orc_los_versor(double theta, double phi, double out[3]);
double orc_los_parproj(const double b[3], const double l[3]);
double orc_los_perproj(const double b[3], const double l[3]);
double orc_sync_ipa(const double b[3], double theta, double phi);
double orc_temp_convert(double t, double freq);
int orc_check_upper(double value, double limit);
int orc_check_lower(double value, double limit);

/* ---- field models ---- */
void orc_breg(int model, const double *par, const double pos[3], double out[3]);
double orc_tereg_unif(const double *par, const double obs[3], const double pos[3]);
/* trilinear gathers (brnd.cc:25-88, ternd.cc:21-71) */
void orc_read_grid3(const double *bx, const double *by, const double *bz,
                    const hamx_box *box, const double pos[3], double out[3]);
double orc_read_grid1(const double *f, const hamx_box *box, const double pos[3]);
/* analytic synchrotron emissivities per sample (integrator.cc:333-346,402-415) */
void orc_sync_emissivity(const hamx_los_desc *d, const double pos[3],
                         double b_per, double *j_tot, double *j_pol);

/* ---- one full write_grid pass (integrator.cc:23-134) ----
 * nthreads <= 0: OpenMP default.  Returns 0 or a negative HAMX_ERR_* code. */
int orc_write_grid(const hamx_los_desc *d, const hamx_fields *f,
                   hamx_maps *out, int nthreads);
/* The march of one ray (radial_integration, integrator.cc:136-228) on shell
 * s, with fd_inner given: out[0..4] = dm, fd, I, Q, U (before temp_convert). */
void orc_radial_integration(const hamx_los_desc *d, const hamx_fields *f,
                            uint32_t shell, double theta, double phi,
                            double fd_inner, double out[5]);

/* ---- random fields ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
void orc_gauss_pairs(uint64_t seed, uint32_t stream, uint64_t first_pair,
                     uint64_t count, double *out);
void orc_fft3d(double *data, uint64_t nx, uint64_t ny, uint64_t nz, int sign);
double orc_variance(const double *arr, uint64_t n);
void orc_sync_kernels(const double *x, uint64_t n, double *F, double *G);
void orc_sync_emissivity_grid(const hamx_los_desc *d, const hamx_fields *f,
                              const double pos[3], double b_per, double *j_tot,
                              double *j_pol);
double orc_tereg_ymw16(const double *p, const double pos[3]);
double orc_cre_read_grid_num(const hamx_fields *f, const double pos[3], uint64_t e);
int orc_cre_write_grid(const hamx_los_desc *d, const hamx_box *box, double E_min,
                       double E_fact, uint32_t nE, double *out);
int orc_brnd_es(const hamx_es_desc *d, uint64_t seed, double *bx, double *by,
                double *bz);
int orc_brnd_jf12(const hamx_es_desc *d, uint64_t seed, double *bx, double *by,
                  double *bz);
double orc_jf12_profile(const double *q, const double pos[3]);
int orc_brnd_mhd(const hamx_mhd_desc *d, uint64_t seed, double *bx, double *by,
                 double *bz);
void orc_mhd_terms(double beta, double ma, double cosa, double out[3]);
double orc_mhd_spectrum(const hamx_mhd_desc *d, double k, double a0, double p0);
int orc_ternd_dft(const hamx_dft_desc *d, uint64_t seed, double *te);
double orc_es_spectrum(const hamx_es_desc *d, double k);
double orc_dft_spectrum(const hamx_dft_desc *d, double k);

#ifdef __cplusplus
}
#endif

#endif
