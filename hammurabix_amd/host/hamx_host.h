// C++ host side of the device path: the reference's Param, field-plugin,
// Grid_*, Integrator and Pipeline interfaces (include/param.h,
// include/bfield.h, include/tefield.h, include/crefield.h, include/grid.h,
// include/integrator.h, include/pipeline.h), implemented over the C-ABI of
// include/hamx.h.  Class and member names, argument order and meaning, and
// the std::runtime_error behaviour follow the reference, so a program
// written against the reference's classes compiles against these.
//
// Field plugins: a model with a device evaluator (LSA, cartesian and
// uniform regular B; uniform TE; analytic and uniform CRE) runs on the GPU.
// A user-derived model overriding write_field() has none; the integrator
// then bakes it onto its vertex grid (box_breg / box_tereg) with the
// reference's own write_grid sampling and gathers it trilinearly on the
// device -- an approximation of the analytic model at grid resolution.
#pragma once

#include <array>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "hamx.h"

namespace hamx_host {

// -------------------------------------------------------------- vectors --
struct Vec3 {
  std::array<double, 3> e{{0., 0., 0.}};
  Vec3() = default;
  Vec3(double x, double y, double z) : e{{x, y, z}} {}
  double operator[](int i) const { return e[i]; }
  double &operator[](int i) { return e[i]; }
};

// ---------------------------------------------------------------- Param --
class Param {
public:
  explicit Param(const std::string &xml_file);
  Param() = default;
  Param(const Param &) = delete;
  Param &operator=(const Param &) = delete;
  virtual ~Param() = default;

  Vec3 observer;
  struct param_grid {
    std::string filename;
    bool build_permission = false, read_permission = false, write_permission = false;
    double x_max = 0, x_min = 0, y_max = 0, y_min = 0, z_max = 0, z_min = 0;
    unsigned long long nx = 0, ny = 0, nz = 0, full_size = 0;
    // energy axis of the CRE grid (param.cc:760-776)
    double E_min = 0, E_max = 0, E_fact = 0;
    unsigned long long nE = 0, cre_size = 0;
  } grid_breg, grid_brnd, grid_tereg, grid_ternd, grid_cre;
  struct param_obs_grid {
    bool write_permission = false;
    unsigned long long nside_dm = 0, nside_fd = 0, nside_mask = 0, total_shell = 0;
    std::vector<unsigned long long> nside_sync, nside_shell;
    std::vector<double> cut_shell, radii_shell;
    double gc_r_min = 0, gc_r_max = 0, gc_z_min = 0, gc_z_max = 0;
    double oc_r_min = 0, oc_r_max = 0, oc_r_res = 0;
    bool do_dm = false, do_fd = false;
    std::vector<bool> do_sync;
    std::string sim_fd_name, sim_dm_name;
    std::vector<std::string> sim_sync_name;
    std::vector<double> sim_sync_freq;
    bool do_mask = false;
    std::string mask_name;
  } grid_obs;
  std::string breg_type, brnd_type, brnd_method;
  struct { double b0 = 0, psi0 = 0, psi1 = 0, chi0 = 0; } breg_lsa;
  struct { double bp = 0, bv = 0, l0 = 0; } breg_unif;
  struct { double bx = 0, by = 0, bz = 0; } breg_cart;
  struct { double bx = 0, by = 0, bz = 0, r_min = 0, r_max = 0; } breg_helix;
  // in the HAMX_BREG_JF12 parameter order (b_arm1..7, b_ring, h_disk, w_disk,
  // Bn, Bs, rn, rs, wh, z0, B0_X, Xtheta, rpc_X, r0_X)
  double breg_jf12[20] = {};
  // in the HAMX_BREG_JAFFE parameter order (include/hamx.h)
  double breg_jaffe[33] = {};
  unsigned long long brnd_seed = 0;
  struct { double rms = 0, k0 = 0, k1 = 0, a0 = 0, a1 = 0, rho = 0, r0 = 0, z0 = 0; } brnd_es;
  // Brnd_jf12: spectrum in brnd_es (r0, z0 unused); energy profile b0_1..b0_8,
  // b0_int, z0_spiral, b0_halo, r0_halo, z0_halo (param.cc:375-409)
  double brnd_jf12[13] = {};
  // Brnd_mhd (param.cc:424-440); power norms in G^2
  struct {
    double pa0 = 0, pf0 = 0, ps0 = 0, k0 = 0, k1 = 0, aa0 = 0, af0 = 0, as0 = 0, a1 = 0,
           beta = 0, ma = 0;
  } brnd_mhd;
  std::string tereg_type, ternd_type, ternd_method;
  struct { double n0 = 0, r0 = 0; } tereg_unif;
  // YMW16 in HAMX_TEREG_YMW16 order and CGS units (param.cc:480-641)
  double tereg_ymw16[73] = {};
  unsigned long long ternd_seed = 0;
  struct { double rms = 0, k0 = 0, a0 = 0, r0 = 0, z0 = 0; } ternd_dft;
  std::string cre_type;
  struct { double alpha = 0, beta = 0, theta = 0, r0 = 0, z0 = 0, E0 = 0, j0 = 0; } cre_ana;
  struct { double alpha = 0, r0 = 0, E0 = 0, j0 = 0; } cre_unif;
};

// ---------------------------------------------------------------- grids --
class Grid {
public:
  virtual ~Grid() = default;
  virtual void build_grid(const Param *) {}
  virtual void export_grid(const Param *) {}
  virtual void import_grid(const Param *) {}
};

// Host arrays plus device copies kept across write_grid passes.
class Grid_breg : public Grid {
public:
  Grid_breg() = default;
  explicit Grid_breg(const Param *par);
  void build_grid(const Param *) override;
  void export_grid(const Param *) override;
  void import_grid(const Param *) override;
  std::unique_ptr<double[]> bx, by, bz;
};

class Grid_brnd : public Grid {
public:
  Grid_brnd() = default;
  explicit Grid_brnd(const Param *par);
  ~Grid_brnd() override;
  void build_grid(const Param *) override;
  void export_grid(const Param *) override;
  void import_grid(const Param *) override;
  std::unique_ptr<double[]> bx, by, bz;
  // device copies, uploaded on first use by the integrator
  mutable double *dev_bx = nullptr, *dev_by = nullptr, *dev_bz = nullptr;
  mutable bool dev_valid = false;
  void invalidate_device() const { dev_valid = false; }
};

class Grid_tereg : public Grid {
public:
  Grid_tereg() = default;
  explicit Grid_tereg(const Param *par);
  void build_grid(const Param *) override;
  void export_grid(const Param *) override;
  void import_grid(const Param *) override;
  std::unique_ptr<double[]> te;
};

class Grid_ternd : public Grid {
public:
  Grid_ternd() = default;
  explicit Grid_ternd(const Param *par);
  ~Grid_ternd() override;
  void build_grid(const Param *) override;
  void export_grid(const Param *) override;
  void import_grid(const Param *) override;
  std::unique_ptr<double[]> te;
  mutable double *dev_te = nullptr;
  mutable bool dev_valid = false;
  void invalidate_device() const { dev_valid = false; }
};

// Grid_cre (grid_cre.cc): the CRE flux [x][y][z][E], kept on the device
// once uploaded.
class Grid_cre : public Grid {
public:
  Grid_cre() = default;
  explicit Grid_cre(const Param *par);
  ~Grid_cre() override;
  void build_grid(const Param *) override;
  void export_grid(const Param *) override;
  void import_grid(const Param *) override;
  std::unique_ptr<double[]> cre_flux;
  mutable double *dev_flux = nullptr;
  mutable bool dev_valid = false;
};

// A RING-ordered HEALPix map: nside and npix doubles.
struct Hampix {
  unsigned long long nside = 0;
  std::vector<double> data;
  Hampix() = default;
  explicit Hampix(unsigned long long n) : nside(n), data(12 * n * n, 0.0) {}
  unsigned long long npix() const { return data.size(); }
};

class Grid_obs : public Grid {
public:
  Grid_obs() = default;
  explicit Grid_obs(const Param *par);
  void build_grid(const Param *) override;
  void export_grid(const Param *) override;
  std::unique_ptr<Hampix> dm_map, is_map, qs_map, us_map, fd_map;
  std::unique_ptr<Hampix> mask_map; // <mask> map at nside_mask (grid_obs.cc:44-50)
};

// --------------------------------------------------------- field plugins --
class Breg {
public:
  virtual ~Breg() = default;
  // field at a galactocentric position [cm] in Gauss; base class: zero
  virtual Vec3 write_field(const Vec3 &pos, const Param *par) const;
  // dispatch of breg.cc:11-21 (host evaluation)
  Vec3 read_field(const Vec3 &pos, const Param *par, const Grid_breg *grid) const;
  Vec3 read_grid(const Vec3 &pos, const Param *par, const Grid_breg *grid) const;
  // sample write_field on the box_breg vertices (breg.cc:99-121)
  void write_grid(const Param *par, Grid_breg *grid) const;
  // device evaluator: HAMX_BREG_* and its parameters, or -1 (none)
  virtual int device_model(const Param *par, double p[HAMX_BREG_NPARAM]) const;
};
class Breg_lsa : public Breg {
public:
  Vec3 write_field(const Vec3 &pos, const Param *par) const override;
  int device_model(const Param *par, double p[HAMX_BREG_NPARAM]) const override;
};
class Breg_cart : public Breg {
public:
  Vec3 write_field(const Vec3 &pos, const Param *par) const override;
  int device_model(const Param *par, double p[HAMX_BREG_NPARAM]) const override;
};
class Breg_unif : public Breg {
public:
  Vec3 write_field(const Vec3 &pos, const Param *par) const override;
  int device_model(const Param *par, double p[HAMX_BREG_NPARAM]) const override;
};
class Breg_helix : public Breg {
public:
  Vec3 write_field(const Vec3 &pos, const Param *par) const override;
  int device_model(const Param *par, double p[HAMX_BREG_NPARAM]) const override;
};
class Breg_jf12 : public Breg {
public:
  Vec3 write_field(const Vec3 &pos, const Param *par) const override;
  int device_model(const Param *par, double p[HAMX_BREG_NPARAM]) const override;
};
class Breg_jaffe : public Breg {
public:
  Vec3 write_field(const Vec3 &pos, const Param *par) const override;
  int device_model(const Param *par, double p[HAMX_BREG_NPARAM]) const override;
};

class Brnd {
public:
  virtual ~Brnd() = default;
  virtual void write_grid(const Param *, const Breg *, const Grid_breg *, Grid_brnd *) const;
};
class Brnd_es : public Brnd {
public:
  void write_grid(const Param *par, const Breg *breg, const Grid_breg *gbreg,
                  Grid_brnd *grid) const override;

protected:
  virtual int generate(hamx_ctx *ctx, const hamx_es_desc *d, uint64_t seed, Grid_brnd *grid) const;
};
// Brnd_mhd (brnd_mhd.cc): Alfven, fast and slow modes
class Brnd_mhd : public Brnd {
public:
  void write_grid(const Param *par, const Breg *breg, const Grid_breg *gbreg,
                  Grid_brnd *grid) const override;
};
// Brnd_jf12 (brnd_jf12.cc): Brnd_es with the JF12 energy profile
class Brnd_jf12 : public Brnd_es {
protected:
  int generate(hamx_ctx *ctx, const hamx_es_desc *d, uint64_t seed,
               Grid_brnd *grid) const override;
};

class TEreg {
public:
  virtual ~TEreg() = default;
  virtual double write_field(const Vec3 &pos, const Param *par) const;
  double read_field(const Vec3 &pos, const Param *par, const Grid_tereg *grid) const;
  void write_grid(const Param *par, Grid_tereg *grid) const;
  // device evaluator: HAMX_TEREG_* and its HAMX_TEREG_NPARAM parameters, or -1
  virtual int device_model(const Param *par, double *p) const;
};
class TEreg_unif : public TEreg {
public:
  double write_field(const Vec3 &pos, const Param *par) const override;
  int device_model(const Param *par, double *p) const override;
};
// TEreg_ymw16 (tereg_ymw16.cc): the shared host/device evaluator
class TEreg_ymw16 : public TEreg {
public:
  double write_field(const Vec3 &pos, const Param *par) const override;
  int device_model(const Param *par, double *p) const override;
};

class TErnd {
public:
  virtual ~TErnd() = default;
  virtual void write_grid(const Param *, const TEreg *, const Grid_tereg *, Grid_ternd *) const;
};
class TErnd_dft : public TErnd {
public:
  void write_grid(const Param *par, const TEreg *tereg, const Grid_tereg *gtereg,
                  Grid_ternd *grid) const override;
};

class CREfield {
public:
  virtual ~CREfield() = default;
  // device evaluator: HAMX_CRE_* and its parameters, or -1 (none)
  virtual int device_model(const Param *par, double p[8]) const;
  // CREfield::write_grid (cre.cc:204-225) on the device
  virtual void write_grid(const Param *par, Grid_cre *grid) const;
};
// CRE_num (crefield.h:132): the imported grid
class CRE_num : public CREfield {
public:
  int device_model(const Param *par, double p[8]) const override;
};
class CRE_ana : public CREfield {
public:
  int device_model(const Param *par, double p[8]) const override;
};
class CRE_unif : public CREfield {
public:
  int device_model(const Param *par, double p[8]) const override;
};

// ----------------------------------------------------------- integrator --
class Integrator {
public:
  Integrator() = default;
  Integrator(const Integrator &) = delete;
  Integrator &operator=(const Integrator &) = delete;
  virtual ~Integrator() = default;
  // integrator.h:25-28: one pass for the sync entry grid_obs.*.back()
  void write_grid(const Breg *, const Brnd *, const TEreg *, const TErnd *,
                  const CREfield *, const Grid_breg *, const Grid_brnd *,
                  const Grid_tereg *, const Grid_ternd *, const Grid_cre *,
                  Grid_obs *, const Param *) const;
};

// ------------------------------------------------------------- pipeline --
class Pipeline {
public:
  explicit Pipeline(const std::string &xml_file);
  void assemble_grid();
  void assemble_tereg();
  void assemble_breg();
  void assemble_ternd();
  void assemble_brnd();
  void assemble_cre();
  void assemble_obs();
  std::unique_ptr<Param> par;
  std::unique_ptr<Grid_breg> grid_breg;
  std::unique_ptr<Grid_brnd> grid_brnd;
  std::unique_ptr<Grid_tereg> grid_tereg;
  std::unique_ptr<Grid_ternd> grid_ternd;
  std::unique_ptr<Grid_cre> grid_cre;
  std::unique_ptr<Grid_obs> grid_obs;
  std::unique_ptr<Breg> breg;
  std::unique_ptr<Brnd> brnd;
  std::unique_ptr<TEreg> tereg;
  std::unique_ptr<TErnd> ternd;
  std::unique_ptr<CREfield> cre;
  std::unique_ptr<Integrator> intobj;
};

// The per-thread device context used by the adapters.
hamx_ctx *context();
// toolkit::random_seed (toolkit.h:122-136): 0 -> clock and thread based
unsigned long long random_seed(unsigned long long s);

} // namespace hamx_host
