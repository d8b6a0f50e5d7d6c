// C++ host side of the device path: the reference's Param, field-plugin,
// Grid_*, Integrator and Pipeline interfaces (include/param.h,
// include/bfield.h, include/tefield.h, include/crefield.h, include/grid.h,
// include/integrator.h, include/pipeline.h), implemented over the C-ABI of
// include/hamx.h.  Class and member names, argument order and meaning, and
// the std::runtime_error behaviour follow the reference, so a program
// written against the reference's classes compiles against these.
//
// Field plugins: the reference's virtual interface (bfield.h:44-97,
// tefield.h:40-81, crefield.h:24-84) with its signatures over
// Hamvec<3, ham_float>.  A library model with a device evaluator (LSA,
// cartesian, uniform, helix, JF12 and Jaffe regular B; uniform and YMW16 TE;
// analytic and uniform CRE) runs on the GPU.  Any other model -- a user class
// derived from Breg / TEreg / CREfield, or from one of the library's models
// -- has none: the integrator bakes what its read_field() returns at the
// vertices of box_breg / box_tereg (write_field() by default, the reference's
// write_grid sampling, breg.cc:99-121, tereg.cc:76-95), and a user CRE model
// is written onto box_cre by CREfield::write_grid (cre.cc:202-227) and
// integrated as an imported CRE grid (CRE_num, integrator.cc:281-401).  The
// device then gathers those grids trilinearly: an approximation of the
// user's model at grid resolution (and, for a CRE model, the energy
// integral over the box_cre energies in place of the analytic power law).
#pragma once

#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <ostream>
#include <string>
#include <vector>

#include "hamx.h"

namespace hamx_host {

// ---------------------------------------------------------------- types --
// the reference's scalar types (hamtype.h)
typedef std::uint_fast32_t ham_uint;
typedef std::int_fast32_t ham_int;
typedef std::int_fast8_t ham_sint;
typedef double ham_float;

// -------------------------------------------------------------- vectors --
// A small fixed-size vector with the member functions of the reference's
// Hamvec (hamvec.h): element access, +, -, scalar * and /, comparisons,
// length, lengthsq, flip, versor, dotprod, crossprod and stream output.
template <ham_sint dim, typename T> class Hamvec {
public:
  Hamvec() { e.fill(T(0)); }
  Hamvec(const T &x) {
    e.fill(T(0));
    e[0] = x;
  }
  Hamvec(const T &x, const T &y) {
    e.fill(T(0));
    e[0] = x;
    if (dim > 1)
      e[1] = y;
  }
  Hamvec(const T &x, const T &y, const T &z) {
    e.fill(T(0));
    e[0] = x;
    if (dim > 1)
      e[1] = y;
    if (dim > 2)
      e[2] = z;
  }
  T operator[](const ham_sint &i) const { return e[i]; }
  T &operator[](const ham_sint &i) { return e[i]; }
  Hamvec operator+(const Hamvec &v) const {
    Hamvec r(*this);
    return r += v;
  }
  Hamvec &operator+=(const Hamvec &v) {
    for (int i = 0; i < dim; ++i)
      e[i] += v.e[i];
    return *this;
  }
  Hamvec operator-(const Hamvec &v) const {
    Hamvec r(*this);
    return r -= v;
  }
  Hamvec &operator-=(const Hamvec &v) {
    for (int i = 0; i < dim; ++i)
      e[i] -= v.e[i];
    return *this;
  }
  Hamvec operator*(const T &s) const {
    Hamvec r(*this);
    return r *= s;
  }
  Hamvec &operator*=(const T &s) {
    for (int i = 0; i < dim; ++i)
      e[i] *= s;
    return *this;
  }
  Hamvec operator/(const T &s) const {
    Hamvec r(*this);
    return r /= s;
  }
  Hamvec &operator/=(const T &s) {
    for (int i = 0; i < dim; ++i)
      e[i] /= s;
    return *this;
  }
  bool operator==(const Hamvec &v) const { return e == v.e; }
  bool operator!=(const Hamvec &v) const { return e != v.e; }
  ham_float lengthsq() const {
    ham_float s = 0;
    for (int i = 0; i < dim; ++i)
      s += static_cast<ham_float>(e[i]) * static_cast<ham_float>(e[i]);
    return s;
  }
  ham_float length() const { return std::sqrt(lengthsq()); }
  void flip() {
    for (int i = 0; i < dim; ++i)
      e[i] = -e[i];
  }
  Hamvec<dim, ham_float> versor() const {
    Hamvec<dim, ham_float> r;
    const ham_float l = length();
    if (l == 0)
      return r;
    for (int i = 0; i < dim; ++i)
      r[i] = static_cast<ham_float>(e[i]) / l;
    return r;
  }
  ham_float dotprod(const Hamvec &v) const {
    ham_float s = 0;
    for (int i = 0; i < dim; ++i)
      s += static_cast<ham_float>(e[i]) * static_cast<ham_float>(v.e[i]);
    return s;
  }
  Hamvec crossprod(const Hamvec &v) const {
    static_assert(dim == 3, "the cross product is defined in three dimensions");
    return Hamvec(e[1] * v.e[2] - e[2] * v.e[1], e[2] * v.e[0] - e[0] * v.e[2],
                  e[0] * v.e[1] - e[1] * v.e[0]);
  }
  friend std::ostream &operator<<(std::ostream &os, const Hamvec &v) {
    os << "(";
    for (int i = 0; i < dim; ++i)
      os << (i ? "," : "") << v.e[i];
    return os << ")";
  }

private:
  std::array<T, static_cast<size_t>(dim)> e;
};

using Vec3 = Hamvec<3, ham_float>;

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
// Every class below keeps the reference's virtual members.  device_model()
// is this build's addition: the HAMX_* model of the library's own classes
// (and its parameters), -1 for anything else -- a user class, including one
// derived from a library model, so that what the user class computes is
// what gets integrated.
class Breg {
public:
  Breg() = default;
  Breg(const Breg &) = delete;
  Breg &operator=(const Breg &) = delete;
  virtual ~Breg() = default;
  // breg.cc:11-21: the grid when imported, else write_field when built
  virtual Vec3 read_field(const Vec3 &pos, const Param *par, const Grid_breg *grid) const;
  // field at a galactocentric position [cm] in Gauss; base class: zero
  virtual Vec3 write_field(const Vec3 &pos, const Param *par) const;
  // trilinear gather of the box_breg grid (breg.cc:23-97)
  virtual Vec3 read_grid(const Vec3 &pos, const Param *par, const Grid_breg *grid) const;
  // sample write_field on the box_breg vertices (breg.cc:99-121)
  virtual void write_grid(const Param *par, Grid_breg *grid) const;
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
  Brnd() = default;
  Brnd(const Brnd &) = delete;
  Brnd &operator=(const Brnd &) = delete;
  virtual ~Brnd() = default;
  // brnd.cc:12-19: the random field is read from its grid
  virtual Vec3 read_field(const Vec3 &pos, const Param *par, const Grid_brnd *grid) const;
  // trilinear gather of the box_brnd grid (brnd.cc:21-88): zero on and
  // outside the faces, z then y then x
  virtual Vec3 read_grid(const Vec3 &pos, const Param *par, const Grid_brnd *grid) const;
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
  TEreg() = default;
  TEreg(const TEreg &) = delete;
  TEreg &operator=(const TEreg &) = delete;
  virtual ~TEreg() = default;
  // tereg.cc:11-20: the grid when imported, else write_field when built
  virtual double read_field(const Vec3 &pos, const Param *par, const Grid_tereg *grid) const;
  // density at a galactocentric position [cm] in cm^-3; base class: zero
  virtual double write_field(const Vec3 &pos, const Param *par) const;
  // trilinear gather of the box_tereg grid (tereg.cc:22-74)
  virtual double read_grid(const Vec3 &pos, const Param *par, const Grid_tereg *grid) const;
  // sample write_field on the box_tereg vertices (tereg.cc:76-95)
  virtual void write_grid(const Param *par, Grid_tereg *grid) const;
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
  TErnd() = default;
  TErnd(const TErnd &) = delete;
  TErnd &operator=(const TErnd &) = delete;
  virtual ~TErnd() = default;
  // ternd.cc:12-19: read from the grid
  virtual double read_field(const Vec3 &pos, const Param *par, const Grid_ternd *grid) const;
  // trilinear gather of the box_ternd grid (ternd.cc:21-71)
  virtual double read_grid(const Vec3 &pos, const Param *par, const Grid_ternd *grid) const;
  virtual void write_grid(const Param *, const TEreg *, const Grid_tereg *, Grid_ternd *) const;
};
class TErnd_dft : public TErnd {
public:
  void write_grid(const Param *par, const TEreg *tereg, const Grid_tereg *gtereg,
                  Grid_ternd *grid) const override;
};

class CREfield {
public:
  CREfield() = default;
  CREfield(const CREfield &) = delete;
  CREfield &operator=(const CREfield &) = delete;
  virtual ~CREfield() = default;
  // flux at position and energy E [erg] from the box_cre grid, linear in
  // log E and trilinear in space (cre.cc:49-137) [GeV m^2 s sr]^-1
  virtual double read_grid(const Vec3 &pos, const double &E, const Param *par,
                           const Grid_cre *grid) const;
  // flux at energy index Eidx of the grid, trilinear in space (cre.cc:139-199)
  double read_grid_num(const Vec3 &pos, const ham_uint &Eidx, const Param *par,
                       const Grid_cre *grid) const;
  // CREfield::write_grid (cre.cc:202-227): write_field on the box_cre
  // vertices and energies (on the device for the library's analytic and
  // uniform models)
  virtual void write_grid(const Param *par, Grid_cre *grid) const;
  // cre.cc:12-22: the grid when imported, else write_field when built
  virtual double read_field(const Vec3 &pos, const double &E, const Param *par,
                            const Grid_cre *grid) const;
  // the flux PHI(E) [GeV m^2 s sr]^-1 at a position; base class: zero
  virtual double write_field(const Vec3 &pos, const double &E, const Param *par) const;
  // N(gamma) = flux_norm * gamma^flux_idx; the base class throws
  // (cre.cc:28-44), as the reference's does
  virtual double flux_norm(const Vec3 &pos, const Param *par) const;
  virtual double flux_idx(const Vec3 &pos, const Param *par) const;
  virtual double spatial_profile(const Vec3 &pos, const Param *par) const;
  virtual int device_model(const Param *par, double p[8]) const;
};
// CRE_num (crefield.h:132): the imported grid
class CRE_num : public CREfield {
public:
  int device_model(const Param *par, double p[8]) const override;
};
// CRE_ana (cre_ana.cc:11-63)
class CRE_ana : public CREfield {
public:
  double write_field(const Vec3 &pos, const double &E, const Param *par) const override;
  double flux_norm(const Vec3 &pos, const Param *par) const override;
  double flux_idx(const Vec3 &pos, const Param *par) const override;
  double spatial_profile(const Vec3 &pos, const Param *par) const override;
  int device_model(const Param *par, double p[8]) const override;
};
// CRE_unif (cre_unif.cc:10-59)
class CRE_unif : public CREfield {
public:
  double write_field(const Vec3 &pos, const double &E, const Param *par) const override;
  double flux_norm(const Vec3 &pos, const Param *par) const override;
  double flux_idx(const Vec3 &pos, const Param *par) const override;
  double spatial_profile(const Vec3 &pos, const Param *par) const override;
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
// pipeline.h:14-45: a user model is added the reference's way, by a
// Pipeline subclass whose assemble_* installs it (the members are public
// here, protected there, so a derived class reaches them either way).
class Pipeline {
public:
  Pipeline() = default;
  explicit Pipeline(const std::string &xml_file);
  Pipeline(const Pipeline &) = delete;
  Pipeline &operator=(const Pipeline &) = delete;
  virtual ~Pipeline() = default;
  virtual void assemble_grid();
  virtual void assemble_tereg();
  virtual void assemble_breg();
  virtual void assemble_ternd();
  virtual void assemble_brnd();
  virtual void assemble_cre();
  virtual void assemble_obs();
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
