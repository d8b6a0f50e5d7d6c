// Field plugins, grids, integrator and pipeline over the C-ABI.
#include "hamx_host.h"
#include "hamx_cgs.h"
#include "hamx_models.h"

#include <chrono>
#include <cstring>
#include <fstream>
#include <functional>
#include <stdexcept>
#include <thread>
#include <typeinfo>

namespace hamx_host {

namespace {

const double kKpc = 3.0856775806e+21;
const double kPc = 3.0856775806e+18;
const double kM = 1.e2;
// Is `obj` exactly of the library class T (not a user class derived from it)?
template <class T, class U> bool exactly(const U &obj) { return typeid(obj) == typeid(T); }

// Throw the library's message as the reference's std::runtime_error.
void check(int rc, hamx_ctx *ctx) {
  if (rc != HAMX_OK)
    throw std::runtime_error(std::string("hamx: ") + hamx_last_error(ctx));
}

hamx_box to_box(const Param::param_grid &g) {
  hamx_box b;
  b.nx = g.nx;
  b.ny = g.ny;
  b.nz = g.nz;
  b.x_min = g.x_min;
  b.x_max = g.x_max;
  b.y_min = g.y_min;
  b.y_max = g.y_max;
  b.z_min = g.z_min;
  b.z_max = g.z_max;
  return b;
}

double length(const Vec3 &v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

// Raw little-endian FP64 arrays, the reference's grid and map formats.
void write_doubles(const std::string &path, const std::vector<const double *> &comps,
                   unsigned long long n) {
  std::ofstream out(path, std::ios::binary);
  if (!out)
    throw std::runtime_error("unable to open file " + path);
  for (unsigned long long i = 0; i < n; ++i)
    for (const double *c : comps)
      out.write(reinterpret_cast<const char *>(&c[i]), sizeof(double));
}

void read_doubles(const std::string &path, const std::vector<double *> &comps,
                  unsigned long long n) {
  std::ifstream in(path, std::ios::binary | std::ios::ate);
  if (!in)
    throw std::runtime_error("unable to open file " + path);
  const auto size = static_cast<unsigned long long>(in.tellg());
  if (size != n * comps.size() * sizeof(double))
    throw std::runtime_error("grid file " + path + " has an unexpected size");
  in.seekg(0);
  for (unsigned long long i = 0; i < n; ++i)
    for (double *c : comps)
      in.read(reinterpret_cast<char *>(&c[i]), sizeof(double));
}

// Trilinear vertex-grid gather on the host (brnd.cc:25-88), used only to
// evaluate imported regular grids when baking other models.
bool host_cell(const Param::param_grid &g, const Vec3 &p, unsigned long long &idx,
               double &xd, double &yd, double &zd) {
  double t = (g.nx - 1) * (p[0] - g.x_min) / (g.x_max - g.x_min);
  if (!(t > 0 && t < g.nx - 1))
    return false;
  const auto xl = static_cast<unsigned long long>(std::floor(t));
  xd = t - xl;
  t = (g.ny - 1) * (p[1] - g.y_min) / (g.y_max - g.y_min);
  if (!(t > 0 && t < g.ny - 1))
    return false;
  const auto yl = static_cast<unsigned long long>(std::floor(t));
  yd = t - yl;
  t = (g.nz - 1) * (p[2] - g.z_min) / (g.z_max - g.z_min);
  if (!(t > 0 && t < g.nz - 1))
    return false;
  const auto zl = static_cast<unsigned long long>(std::floor(t));
  zd = t - zl;
  idx = xl * g.ny * g.nz + yl * g.nz + zl;
  return true;
}

double host_lerp(const double *f, const Param::param_grid &g, unsigned long long idx,
                 double xd, double yd, double zd) {
  const unsigned long long sy = g.nz, sx = g.ny * g.nz;
  const double i1 = f[idx] * (1. - zd) + f[idx + 1] * zd;
  const double i2 = f[idx + sy] * (1. - zd) + f[idx + sy + 1] * zd;
  const double j1 = f[idx + sx] * (1. - zd) + f[idx + sx + 1] * zd;
  const double j2 = f[idx + sx + sy] * (1. - zd) + f[idx + sx + sy + 1] * zd;
  const double w1 = i1 * (1. - yd) + i2 * yd;
  const double w2 = j1 * (1. - yd) + j2 * yd;
  return w1 * (1. - xd) + w2 * xd;
}

// sample a field model on the vertices of a box (breg.cc:99-121)
void sample_vertices(const Param::param_grid &g,
                     const std::function<void(const Vec3 &, unsigned long long)> &f) {
  const double lx = g.x_max - g.x_min, ly = g.y_max - g.y_min, lz = g.z_max - g.z_min;
  Vec3 p;
  for (unsigned long long i = 0; i != g.nx; ++i) {
    p[0] = i * lx / (g.nx - 1) + g.x_min;
    for (unsigned long long j = 0; j != g.ny; ++j) {
      p[1] = j * ly / (g.ny - 1) + g.y_min;
      for (unsigned long long k = 0; k != g.nz; ++k) {
        p[2] = k * lz / (g.nz - 1) + g.z_min;
        f(p, i * g.ny * g.nz + j * g.nz + k);
      }
    }
  }
}

} // namespace

hamx_ctx *context() {
  thread_local struct Holder {
    hamx_ctx *ctx = nullptr;
    ~Holder() {
      if (ctx)
        hamx_close(ctx);
    }
  } h;
  if (!h.ctx) {
    const int rc = hamx_open(0, &h.ctx);
    if (rc != HAMX_OK)
      throw std::runtime_error("hamx: cannot open a context on device 0 (no usable GPU?)");
  }
  return h.ctx;
}

unsigned long long random_seed(unsigned long long s) {
  if (s != 0)
    return s;
  const auto now = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
  return static_cast<unsigned long long>(now) +
         std::hash<std::thread::id>()(std::this_thread::get_id());
}

// ----------------------------------------------------------------- grids --
Grid_breg::Grid_breg(const Param *par) {
  if (par->grid_breg.read_permission || par->grid_breg.write_permission)
    build_grid(par);
}
void Grid_breg::build_grid(const Param *par) {
  const auto n = par->grid_breg.full_size;
  bx = std::make_unique<double[]>(n);
  by = std::make_unique<double[]>(n);
  bz = std::make_unique<double[]>(n);
}
void Grid_breg::export_grid(const Param *par) {
  write_doubles(par->grid_breg.filename, {bx.get(), by.get(), bz.get()},
                par->grid_breg.full_size);
}
void Grid_breg::import_grid(const Param *par) {
  read_doubles(par->grid_breg.filename, {bx.get(), by.get(), bz.get()},
               par->grid_breg.full_size);
}

Grid_brnd::Grid_brnd(const Param *par) {
  if (par->grid_brnd.build_permission || par->grid_brnd.read_permission)
    build_grid(par);
}
Grid_brnd::~Grid_brnd() {
  if (dev_bx) {
    hamx_ctx *c = context();
    hamx_dev_free(c, dev_bx);
    hamx_dev_free(c, dev_by);
    hamx_dev_free(c, dev_bz);
  }
}
void Grid_brnd::build_grid(const Param *par) {
  const auto n = par->grid_brnd.full_size;
  bx = std::make_unique<double[]>(n);
  by = std::make_unique<double[]>(n);
  bz = std::make_unique<double[]>(n);
  invalidate_device();
}
void Grid_brnd::export_grid(const Param *par) {
  write_doubles(par->grid_brnd.filename, {bx.get(), by.get(), bz.get()},
                par->grid_brnd.full_size);
}
void Grid_brnd::import_grid(const Param *par) {
  read_doubles(par->grid_brnd.filename, {bx.get(), by.get(), bz.get()},
               par->grid_brnd.full_size);
  invalidate_device();
}

Grid_tereg::Grid_tereg(const Param *par) {
  if (par->grid_tereg.read_permission || par->grid_tereg.write_permission)
    build_grid(par);
}
void Grid_tereg::build_grid(const Param *par) {
  te = std::make_unique<double[]>(par->grid_tereg.full_size);
}
void Grid_tereg::export_grid(const Param *par) {
  write_doubles(par->grid_tereg.filename, {te.get()}, par->grid_tereg.full_size);
}
void Grid_tereg::import_grid(const Param *par) {
  read_doubles(par->grid_tereg.filename, {te.get()}, par->grid_tereg.full_size);
}

Grid_ternd::Grid_ternd(const Param *par) {
  if (par->grid_ternd.build_permission || par->grid_ternd.read_permission)
    build_grid(par);
}
Grid_ternd::~Grid_ternd() {
  if (dev_te)
    hamx_dev_free(context(), dev_te);
}
void Grid_ternd::build_grid(const Param *par) {
  te = std::make_unique<double[]>(par->grid_ternd.full_size);
  invalidate_device();
}
void Grid_ternd::export_grid(const Param *par) {
  write_doubles(par->grid_ternd.filename, {te.get()}, par->grid_ternd.full_size);
}
void Grid_ternd::import_grid(const Param *par) {
  read_doubles(par->grid_ternd.filename, {te.get()}, par->grid_ternd.full_size);
  invalidate_device();
}

Grid_obs::Grid_obs(const Param *par) { build_grid(par); }

// grid_obs.cc:21-43
void Grid_obs::build_grid(const Param *par) {
  const auto &g = par->grid_obs;
  if (g.do_dm)
    dm_map = std::make_unique<Hampix>(g.nside_dm);
  if (g.do_sync.back()) {
    is_map = std::make_unique<Hampix>(g.nside_sync.back());
    qs_map = std::make_unique<Hampix>(g.nside_sync.back());
    us_map = std::make_unique<Hampix>(g.nside_sync.back());
    fd_map = std::make_unique<Hampix>(g.nside_sync.back());
  }
  if (g.do_fd)
    fd_map = std::make_unique<Hampix>(g.nside_fd);
  if (g.do_mask) {
    mask_map = std::make_unique<Hampix>(g.nside_mask);
    read_doubles(g.mask_name, {mask_map->data.data()}, mask_map->npix());
  }
}

// grid_obs.cc:53-83: DM in pc cm^-3, FD in rad m^-2, _I/_Q/_U file names
void Grid_obs::export_grid(const Param *par) {
  const auto &g = par->grid_obs;
  auto dump = [](const std::string &path, const Hampix &m, double scale) {
    std::vector<double> v(m.data);
    for (double &x : v)
      x *= scale;
    write_doubles(path, {v.data()}, v.size());
  };
  if (g.do_dm)
    dump(g.sim_dm_name, *dm_map, 1. / kPc);
  if (g.do_sync.back()) {
    const std::string &name = g.sim_sync_name.back();
    if (name.size() < 4)
      throw std::runtime_error("sync output file name too short: " + name);
    auto with = [&](const char *tag) {
      std::string s = name;
      s.insert(s.size() - 4, tag);
      return s;
    };
    dump(with("_I"), *is_map, 1.);
    dump(with("_Q"), *qs_map, 1.);
    dump(with("_U"), *us_map, 1.);
  }
  if (g.do_fd)
    dump(g.sim_fd_name, *fd_map, kM * kM);
}

// ------------------------------------------------------------- plugins --
Vec3 Breg::write_field(const Vec3 &, const Param *) const { return Vec3(); }
int Breg::device_model(const Param *, double *) const { return -1; }

Vec3 Breg::read_grid(const Vec3 &pos, const Param *par, const Grid_breg *grid) const {
  unsigned long long idx;
  double xd, yd, zd;
  if (!grid || !grid->bx || !host_cell(par->grid_breg, pos, idx, xd, yd, zd))
    return Vec3();
  return Vec3(host_lerp(grid->bx.get(), par->grid_breg, idx, xd, yd, zd),
              host_lerp(grid->by.get(), par->grid_breg, idx, xd, yd, zd),
              host_lerp(grid->bz.get(), par->grid_breg, idx, xd, yd, zd));
}

Vec3 Breg::read_field(const Vec3 &pos, const Param *par, const Grid_breg *grid) const {
  if (par->grid_breg.read_permission)
    return read_grid(pos, par, grid);
  if (par->grid_breg.build_permission)
    return write_field(pos, par);
  return Vec3();
}

void Breg::write_grid(const Param *par, Grid_breg *grid) const {
  sample_vertices(par->grid_breg, [&](const Vec3 &p, unsigned long long i) {
    const Vec3 b = write_field(p, par);
    grid->bx[i] = b[0];
    grid->by[i] = b[1];
    grid->bz[i] = b[2];
  });
}

// breg_lsa.cc:9-26 (host evaluator: baking and anisotropy checks)
Vec3 Breg_lsa::write_field(const Vec3 &pos, const Param *par) const {
  const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
  if (r > 20. * kKpc || r < 3. * kKpc)
    return Vec3();
  const double cp = pos[0] / r, sp = pos[1] / r;
  const double psi = par->breg_lsa.psi0 + par->breg_lsa.psi1 * std::log(r / (8. * kKpc));
  const double chi = par->breg_lsa.chi0 * std::tanh(pos[2] / kKpc);
  const double b0 = par->breg_lsa.b0;
  return Vec3(b0 * (std::sin(psi) * std::cos(chi) * cp - std::cos(psi) * std::cos(chi) * sp),
              b0 * (std::sin(psi) * std::cos(chi) * sp + std::cos(psi) * std::cos(chi) * cp),
              b0 * std::sin(chi));
}
int Breg_lsa::device_model(const Param *par, double *p) const {
  if (!exactly<Breg_lsa>(*this))
    return -1;
  p[0] = par->breg_lsa.b0;
  p[1] = par->breg_lsa.psi0;
  p[2] = par->breg_lsa.psi1;
  p[3] = par->breg_lsa.chi0;
  return HAMX_BREG_LSA;
}
Vec3 Breg_cart::write_field(const Vec3 &, const Param *par) const {
  return Vec3(par->breg_cart.bx, par->breg_cart.by, par->breg_cart.bz);
}
int Breg_cart::device_model(const Param *par, double *p) const {
  if (!exactly<Breg_cart>(*this))
    return -1;
  p[0] = par->breg_cart.bx;
  p[1] = par->breg_cart.by;
  p[2] = par->breg_cart.bz;
  return HAMX_BREG_CART;
}
Vec3 Breg_unif::write_field(const Vec3 &, const Param *par) const {
  return Vec3(par->breg_unif.bp * std::cos(par->breg_unif.l0),
              par->breg_unif.bp * std::sin(par->breg_unif.l0), par->breg_unif.bv);
}
int Breg_unif::device_model(const Param *par, double *p) const {
  if (!exactly<Breg_unif>(*this))
    return -1;
  p[0] = par->breg_unif.bp;
  p[1] = par->breg_unif.bv;
  p[2] = par->breg_unif.l0;
  return HAMX_BREG_UNIF;
}
// breg_helix.cc:13-30 and breg_jf12.cc:8-189, through the evaluators the
// device march uses (hamx_models.h)
namespace {
void helix_params(const Param *par, double *p) {
  p[0] = par->breg_helix.bx;
  p[1] = par->breg_helix.by;
  p[2] = par->breg_helix.bz;
  p[3] = par->breg_helix.r_min;
  p[4] = par->breg_helix.r_max;
}
} // namespace
int Breg_helix::device_model(const Param *par, double *p) const {
  if (!exactly<Breg_helix>(*this))
    return -1;
  helix_params(par, p);
  return HAMX_BREG_HELIX;
}
Vec3 Breg_helix::write_field(const Vec3 &pos, const Param *par) const {
  double p[HAMX_BREG_NPARAM] = {}, b[3];
  helix_params(par, p);
  hamx::breg_helix_eval(p, pos[0], pos[1], pos[2], b);
  return Vec3(b[0], b[1], b[2]);
}
int Breg_jf12::device_model(const Param *par, double *p) const {
  if (!exactly<Breg_jf12>(*this))
    return -1;
  for (int i = 0; i < 20; ++i)
    p[i] = par->breg_jf12[i];
  return HAMX_BREG_JF12;
}
Vec3 Breg_jf12::write_field(const Vec3 &pos, const Param *par) const {
  double p[HAMX_BREG_NPARAM] = {}, b[3];
  for (int i = 0; i < 20; ++i)
    p[i] = par->breg_jf12[i];
  hamx::breg_jf12_eval(p, pos[0], pos[1], pos[2], b);
  return Vec3(b[0], b[1], b[2]);
}
// breg_jaffe.cc:11-243
int Breg_jaffe::device_model(const Param *par, double *p) const {
  if (!exactly<Breg_jaffe>(*this))
    return -1;
  for (int i = 0; i < 33; ++i)
    p[i] = par->breg_jaffe[i];
  return HAMX_BREG_JAFFE;
}
Vec3 Breg_jaffe::write_field(const Vec3 &pos, const Param *par) const {
  double p[HAMX_BREG_NPARAM] = {}, b[3];
  for (int i = 0; i < 33; ++i)
    p[i] = par->breg_jaffe[i];
  hamx::breg_jaffe_eval(p, pos[0], pos[1], pos[2], b);
  return Vec3(b[0], b[1], b[2]);
}

// brnd.cc:12-88: the random field is the trilinear gather of its grid
Vec3 Brnd::read_field(const Vec3 &pos, const Param *par, const Grid_brnd *grid) const {
  return read_grid(pos, par, grid);
}
Vec3 Brnd::read_grid(const Vec3 &pos, const Param *par, const Grid_brnd *grid) const {
  unsigned long long idx;
  double xd, yd, zd;
  if (!grid || !grid->bx || !host_cell(par->grid_brnd, pos, idx, xd, yd, zd))
    return Vec3();
  return Vec3(host_lerp(grid->bx.get(), par->grid_brnd, idx, xd, yd, zd),
              host_lerp(grid->by.get(), par->grid_brnd, idx, xd, yd, zd),
              host_lerp(grid->bz.get(), par->grid_brnd, idx, xd, yd, zd));
}

double TEreg::write_field(const Vec3 &, const Param *) const { return 0.; }
int TEreg::device_model(const Param *, double *) const { return -1; }
double TEreg::read_grid(const Vec3 &pos, const Param *par, const Grid_tereg *grid) const {
  unsigned long long idx;
  double xd, yd, zd;
  if (!grid || !grid->te || !host_cell(par->grid_tereg, pos, idx, xd, yd, zd))
    return 0.;
  return host_lerp(grid->te.get(), par->grid_tereg, idx, xd, yd, zd);
}
double TEreg::read_field(const Vec3 &pos, const Param *par, const Grid_tereg *grid) const {
  if (par->grid_tereg.read_permission)
    return read_grid(pos, par, grid);
  if (par->grid_tereg.build_permission)
    return write_field(pos, par);
  return 0.;
}
void TEreg::write_grid(const Param *par, Grid_tereg *grid) const {
  sample_vertices(par->grid_tereg,
                  [&](const Vec3 &p, unsigned long long i) { grid->te[i] = write_field(p, par); });
}
double TEreg_unif::write_field(const Vec3 &pos, const Param *par) const {
  const Vec3 d(pos[0] - par->observer[0], pos[1] - par->observer[1], pos[2] - par->observer[2]);
  return length(d) > par->tereg_unif.r0 ? 0. : par->tereg_unif.n0;
}
int TEreg_unif::device_model(const Param *par, double *p) const {
  if (!exactly<TEreg_unif>(*this))
    return -1;
  p[0] = par->tereg_unif.n0;
  p[1] = par->tereg_unif.r0;
  return HAMX_TEREG_UNIF;
}
double TEreg_ymw16::write_field(const Vec3 &pos, const Param *par) const {
  return hamx::tereg_ymw16_eval(par->tereg_ymw16, pos[0], pos[1], pos[2]);
}
int TEreg_ymw16::device_model(const Param *par, double *p) const {
  if (!exactly<TEreg_ymw16>(*this))
    return -1;
  for (int i = 0; i < 73; ++i)
    p[i] = par->tereg_ymw16[i];
  return HAMX_TEREG_YMW16;
}

// ternd.cc:12-71
double TErnd::read_field(const Vec3 &pos, const Param *par, const Grid_ternd *grid) const {
  return read_grid(pos, par, grid);
}
double TErnd::read_grid(const Vec3 &pos, const Param *par, const Grid_ternd *grid) const {
  unsigned long long idx;
  double xd, yd, zd;
  if (!grid || !grid->te || !host_cell(par->grid_ternd, pos, idx, xd, yd, zd))
    return 0.;
  return host_lerp(grid->te.get(), par->grid_ternd, idx, xd, yd, zd);
}

// ------------------------------------------------------------------ CRE --
int CREfield::device_model(const Param *, double *) const { return -1; }
int CRE_num::device_model(const Param *, double *) const { return HAMX_CRE_GRID; }

double CREfield::write_field(const Vec3 &, const double &, const Param *) const { return 0.; }
double CREfield::flux_norm(const Vec3 &, const Param *) const {
  throw std::runtime_error("wrong inheritance");
}
double CREfield::flux_idx(const Vec3 &, const Param *) const {
  throw std::runtime_error("wrong inheritance");
}
double CREfield::spatial_profile(const Vec3 &, const Param *) const {
  throw std::runtime_error("wrong inheritance");
}

double CREfield::read_field(const Vec3 &pos, const double &E, const Param *par,
                            const Grid_cre *grid) const {
  if (par->grid_cre.read_permission)
    return read_grid(pos, E, par, grid);
  if (par->grid_cre.build_permission)
    return write_field(pos, E, par);
  return 0.;
}

namespace {
// One grid axis of box_cre: the lower vertex and the fraction, or false on
// and outside the faces (cre.cc:58-83).
bool cre_axis(double pos, double lo, double hi, unsigned long long n, unsigned long long &l,
              double &d) {
  const double t = (n - 1) * (pos - lo) / (hi - lo);
  if (t <= 0 || t >= n - 1)
    return false;
  l = static_cast<unsigned long long>(std::floor(t));
  d = t - l;
  return true;
}
// toolkit::index4d(n1, n2, n3, n4, e, i, j, l) = e n2 n3 n4 + i n3 n4 + j n4 + l
unsigned long long index4d(unsigned long long n2, unsigned long long n3, unsigned long long n4,
                           unsigned long long e, unsigned long long i, unsigned long long j,
                           unsigned long long l) {
  return e * n2 * n3 * n4 + i * n3 * n4 + j * n4 + l;
}
} // namespace

double CREfield::read_grid(const Vec3 &pos, const double &E, const Param *par,
                           const Grid_cre *grid) const {
  const auto &g = par->grid_cre;
  double t = std::log(E / g.E_min) / std::log(g.E_max / g.E_min);
  if (t <= 0 || t >= g.nE - 1)
    return 0.;
  const auto El = static_cast<unsigned long long>(std::floor(t));
  const double Ed = t - El;
  unsigned long long xl, yl, zl;
  double xd, yd, zd;
  if (!cre_axis(pos[0], g.x_min, g.x_max, g.nx, xl, xd) ||
      !cre_axis(pos[1], g.y_min, g.y_max, g.ny, yl, yd) ||
      !cre_axis(pos[2], g.z_min, g.z_max, g.nz, zl, zd))
    return 0.;
  const double *f = grid->cre_flux.get();
  // [x][y][z][E] flat index, with the argument order of the reference's
  // first pair of reads kept as written there (cre.cc:84-89): nE leads the
  // extents, so those two reads land elsewhere in the array
  auto at = [&](unsigned long long x, unsigned long long y, unsigned long long z,
                unsigned long long e) { return f[index4d(g.ny, g.nz, g.nE, x, y, z, e)]; };
  auto at_first = [&](unsigned long long x, unsigned long long y, unsigned long long z,
                      unsigned long long e) {
    const unsigned long long i = index4d(g.nx, g.ny, g.nz, x, y, z, e);
    return i < g.cre_size ? f[i] : 0.; // beyond the array when nx > nE
  };
  auto plane = [&](unsigned long long e, bool first) {
    const double i1 = first ? at_first(xl, yl, zl, e) * (1. - zd) + at_first(xl, yl, zl + 1, e) * zd
                            : at(xl, yl, zl, e) * (1. - zd) + at(xl, yl, zl + 1, e) * zd;
    const double i2 = at(xl, yl + 1, zl, e) * (1 - zd) + at(xl, yl + 1, zl + 1, e) * zd;
    const double j1 = at(xl + 1, yl, zl, e) * (1 - zd) + at(xl + 1, yl, zl + 1, e) * zd;
    const double j2 = at(xl + 1, yl + 1, zl, e) * (1 - zd) + at(xl + 1, yl + 1, zl + 1, e) * zd;
    const double w1 = i1 * (1 - yd) + i2 * yd;
    const double w2 = j1 * (1 - yd) + j2 * yd;
    return w1 * (1 - xd) + w2 * xd;
  };
  const double q1 = plane(El, true), q2 = plane(El + 1, false);
  return q1 * (1 - Ed) + q2 * Ed;
}

double CREfield::read_grid_num(const Vec3 &pos, const ham_uint &Eidx, const Param *par,
                               const Grid_cre *grid) const {
  const auto &g = par->grid_cre;
  unsigned long long xl, yl, zl;
  double xd, yd, zd;
  if (!cre_axis(pos[0], g.x_min, g.x_max, g.nx, xl, xd) ||
      !cre_axis(pos[1], g.y_min, g.y_max, g.ny, yl, yd) ||
      !cre_axis(pos[2], g.z_min, g.z_max, g.nz, zl, zd))
    return 0.;
  const double *f = grid->cre_flux.get();
  auto at = [&](unsigned long long x, unsigned long long y, unsigned long long z) {
    return f[index4d(g.ny, g.nz, g.nE, x, y, z, Eidx)];
  };
  const double i1 = at(xl, yl, zl) * (1. - zd) + at(xl, yl, zl + 1) * zd;
  const double i2 = at(xl, yl + 1, zl) * (1 - zd) + at(xl, yl + 1, zl + 1) * zd;
  const double j1 = at(xl + 1, yl, zl) * (1 - zd) + at(xl + 1, yl, zl + 1) * zd;
  const double j2 = at(xl + 1, yl + 1, zl) * (1 - zd) + at(xl + 1, yl + 1, zl + 1) * zd;
  const double w1 = i1 * (1 - yd) + i2 * yd;
  const double w2 = j1 * (1 - yd) + j2 * yd;
  return w1 * (1 - xd) + w2 * xd;
}

void CREfield::write_grid(const Param *par, Grid_cre *grid) const {
  const auto &g = par->grid_cre;
  if (!grid->cre_flux || g.cre_size == 0)
    throw std::runtime_error("hamx: CRE grid output needs box_cre");
  hamx_los_desc d;
  std::memset(&d, 0, sizeof(d));
  for (int i = 0; i < 3; ++i)
    d.observer[i] = par->observer[i];
  d.cre_model = device_model(par, d.cre);
  if (d.cre_model == HAMX_CRE_ANA || d.cre_model == HAMX_CRE_UNIF) {
    // the library's models: every (vertex, energy) on the device
    const hamx_box b = to_box(g);
    hamx_ctx *ctx = context();
    check(hamx_cre_write_grid(ctx, &d, &b, g.E_min, g.E_fact, static_cast<uint32_t>(g.nE),
                              grid->cre_flux.get(), 0),
          ctx);
  } else {
    // any other model: write_field at every vertex and energy (cre.cc:202-227)
    double *out = grid->cre_flux.get();
    sample_vertices(g, [&](const Vec3 &p, unsigned long long v) {
      for (unsigned long long m = 0; m != g.nE; ++m)
        out[v * g.nE + m] = write_field(p, g.E_min * std::exp(m * g.E_fact), par);
    });
  }
  grid->dev_valid = false;
}

// CRE_ana (cre_ana.cc:11-63)
double CRE_ana::spatial_profile(const Vec3 &pos, const Param *par) const {
  const double R0 = std::sqrt(par->observer[0] * par->observer[0] +
                              par->observer[1] * par->observer[1]);
  const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
  const double ch = std::cosh(pos[2] / par->cre_ana.z0);
  return std::exp((R0 - r) / par->cre_ana.r0) * (1. / (ch * ch));
}
double CRE_ana::flux_idx(const Vec3 &pos, const Param *par) const {
  const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]) / hamx::cgs::kpc;
  const double z = std::fabs(pos[2]) / hamx::cgs::kpc;
  return -par->cre_ana.alpha + par->cre_ana.beta * r + par->cre_ana.theta * z;
}
double CRE_ana::flux_norm(const Vec3 &pos, const Param *par) const {
  namespace c = hamx::cgs;
  const double gamma0 = par->cre_ana.E0 / c::mec2;
  const double beta0 = std::sqrt(1. - 1. / gamma0);
  const double unit = 4. * c::pi * c::mec / (c::GeV * c::m * c::m * c::sec * beta0);
  const double norm = par->cre_ana.j0 * std::pow(gamma0, -flux_idx(par->observer, par));
  return norm * unit * spatial_profile(pos, par);
}
double CRE_ana::write_field(const Vec3 &pos, const double &E, const Param *par) const {
  namespace c = hamx::cgs;
  const double gamma = E / c::mec2, gamma0 = par->cre_ana.E0 / c::mec2;
  const double beta_ratio = std::sqrt((1. - 1. / gamma) / (1. - 1. / gamma0));
  const double norm = par->cre_ana.j0 * std::pow(gamma0, -flux_idx(par->observer, par));
  return norm * beta_ratio * std::pow(gamma, flux_idx(pos, par)) * spatial_profile(pos, par);
}
int CRE_ana::device_model(const Param *par, double *p) const {
  if (!exactly<CRE_ana>(*this))
    return -1;
  p[0] = par->cre_ana.alpha;
  p[1] = par->cre_ana.beta;
  p[2] = par->cre_ana.theta;
  p[3] = par->cre_ana.r0;
  p[4] = par->cre_ana.z0;
  p[5] = par->cre_ana.E0;
  p[6] = par->cre_ana.j0;
  return HAMX_CRE_ANA;
}

// CRE_unif (cre_unif.cc:10-59)
double CRE_unif::spatial_profile(const Vec3 &pos, const Param *par) const {
  return (pos - par->observer).length() > par->cre_unif.r0 ? 0. : 1.;
}
double CRE_unif::flux_idx(const Vec3 &, const Param *par) const { return -(par->cre_unif.alpha); }
double CRE_unif::flux_norm(const Vec3 &pos, const Param *par) const {
  namespace c = hamx::cgs;
  const double gamma0 = par->cre_unif.E0 / c::mec2;
  const double beta0 = std::sqrt(1. - 1. / gamma0);
  const double unit = 4. * c::pi * c::mec / (c::GeV * c::m * c::m * c::sec * beta0);
  const double norm = par->cre_unif.j0 * std::pow(gamma0, -flux_idx(par->observer, par));
  return norm * unit * spatial_profile(pos, par);
}
double CRE_unif::write_field(const Vec3 &pos, const double &E, const Param *par) const {
  namespace c = hamx::cgs;
  const double gamma = E / c::mec2, gamma0 = par->cre_unif.E0 / c::mec2;
  const double beta_ratio = std::sqrt((1. - 1. / gamma) / (1. - 1. / gamma0));
  const double norm = par->cre_unif.j0 * std::pow(gamma0, -flux_idx(par->observer, par));
  return norm * beta_ratio * std::pow(gamma, flux_idx(pos, par)) * spatial_profile(pos, par);
}
int CRE_unif::device_model(const Param *par, double *p) const {
  if (!exactly<CRE_unif>(*this))
    return -1;
  p[0] = par->cre_unif.alpha;
  p[1] = par->cre_unif.r0;
  p[2] = par->cre_unif.E0;
  p[3] = par->cre_unif.j0;
  return HAMX_CRE_UNIF;
}

Grid_cre::Grid_cre(const Param *par) {
  if (par->grid_cre.read_permission || par->grid_cre.write_permission)
    build_grid(par);
}
Grid_cre::~Grid_cre() {
  if (dev_flux)
    hamx_dev_free(context(), dev_flux);
}
void Grid_cre::build_grid(const Param *par) {
  cre_flux = std::make_unique<double[]>(par->grid_cre.cre_size);
}
void Grid_cre::export_grid(const Param *par) {
  write_doubles(par->grid_cre.filename, {cre_flux.get()}, par->grid_cre.cre_size);
}
void Grid_cre::import_grid(const Param *par) {
  read_doubles(par->grid_cre.filename, {cre_flux.get()}, par->grid_cre.cre_size);
  dev_valid = false;
}

void Brnd::write_grid(const Param *, const Breg *, const Grid_breg *, Grid_brnd *) const {
  throw std::runtime_error("wrong inheritance");
}
void TErnd::write_grid(const Param *, const TEreg *, const Grid_tereg *, Grid_ternd *) const {
  throw std::runtime_error("wrong inheritance");
}

// The regular field a random generator reads (brnd_es.cc:19-22,
// brnd_mhd.cc:33): an imported grid, a device model, or a user model baked
// onto box_breg (kept alive in `baked`).
template <typename Desc>
void generator_breg(const Param *par, const Breg *breg, const Grid_breg *gbreg, Desc &d,
                    std::unique_ptr<Grid_breg> &baked) {
  if (par->grid_breg.read_permission && gbreg) {
    d.breg_model = HAMX_BREG_GRID;
  } else if (par->grid_breg.build_permission && breg) {
    d.breg_model = breg->device_model(par, d.breg);
    if (d.breg_model < 0) {
      if (par->grid_breg.full_size == 0)
        throw std::runtime_error("regular field model has no device evaluator and no "
                                 "box_breg grid to bake it on");
      baked = std::make_unique<Grid_breg>();
      baked->build_grid(par);
      breg->write_grid(par, baked.get());
      gbreg = baked.get();
      d.breg_model = HAMX_BREG_GRID;
    }
  } else {
    d.breg_model = HAMX_BREG_NONE;
  }
  if (d.breg_model == HAMX_BREG_GRID) {
    d.breg_x = gbreg->bx.get();
    d.breg_y = gbreg->by.get();
    d.breg_z = gbreg->bz.get();
    d.breg_box = to_box(par->grid_breg);
  }
}

// Brnd_mhd::write_grid (brnd_mhd.cc:16-167) on the device
void Brnd_mhd::write_grid(const Param *par, const Breg *breg, const Grid_breg *gbreg,
                          Grid_brnd *grid) const {
  hamx_mhd_desc d;
  std::memset(&d, 0, sizeof(d));
  d.box = to_box(par->grid_brnd);
  for (int i = 0; i < 3; ++i)
    d.observer[i] = par->observer[i];
  d.pa0 = par->brnd_mhd.pa0;
  d.pf0 = par->brnd_mhd.pf0;
  d.ps0 = par->brnd_mhd.ps0;
  d.k0 = par->brnd_mhd.k0;
  d.k1 = par->brnd_mhd.k1;
  d.aa0 = par->brnd_mhd.aa0;
  d.af0 = par->brnd_mhd.af0;
  d.as0 = par->brnd_mhd.as0;
  d.a1 = par->brnd_mhd.a1;
  d.beta = par->brnd_mhd.beta;
  d.ma = par->brnd_mhd.ma;
  std::unique_ptr<Grid_breg> baked;
  generator_breg(par, breg, gbreg, d, baked);
  hamx_ctx *ctx = context();
  check(hamx_brnd_mhd(ctx, &d, random_seed(par->brnd_seed), grid->bx.get(), grid->by.get(),
                      grid->bz.get(), 0),
        ctx);
  grid->invalidate_device();
}

// Brnd_es::write_grid (brnd_es.cc:76-308) on the device
void Brnd_es::write_grid(const Param *par, const Breg *breg, const Grid_breg *gbreg,
                         Grid_brnd *grid) const {
  hamx_es_desc d;
  std::memset(&d, 0, sizeof(d));
  d.box = to_box(par->grid_brnd);
  for (int i = 0; i < 3; ++i)
    d.observer[i] = par->observer[i];
  d.rms = par->brnd_es.rms;
  d.k0 = par->brnd_es.k0;
  d.k1 = par->brnd_es.k1;
  d.a0 = par->brnd_es.a0;
  d.a1 = par->brnd_es.a1;
  d.rho = par->brnd_es.rho;
  d.r0 = par->brnd_es.r0;
  d.z0 = par->brnd_es.z0;
  // anisotropy direction: the regular field (brnd_es.cc:19-22)
  std::unique_ptr<Grid_breg> baked;
  generator_breg(par, breg, gbreg, d, baked);
  for (int i = 0; i < 13; ++i)
    d.jf12[i] = par->brnd_jf12[i];
  hamx_ctx *ctx = context();
  check(generate(ctx, &d, random_seed(par->brnd_seed), grid), ctx);
  grid->invalidate_device();
}

int Brnd_es::generate(hamx_ctx *ctx, const hamx_es_desc *d, uint64_t seed,
                      Grid_brnd *grid) const {
  return hamx_brnd_es(ctx, d, seed, grid->bx.get(), grid->by.get(), grid->bz.get(), 0);
}

// Brnd_jf12::write_grid (brnd_jf12.cc:127-385) on the device
int Brnd_jf12::generate(hamx_ctx *ctx, const hamx_es_desc *d, uint64_t seed,
                        Grid_brnd *grid) const {
  return hamx_brnd_jf12(ctx, d, seed, grid->bx.get(), grid->by.get(), grid->bz.get(), 0);
}

// TErnd_dft::write_grid (ternd_dft.cc:35-128); seeded by brnd_seed as in
// the reference (ternd_dft.cc:44,48)
void TErnd_dft::write_grid(const Param *par, const TEreg *, const Grid_tereg *,
                           Grid_ternd *grid) const {
  hamx_dft_desc d;
  std::memset(&d, 0, sizeof(d));
  d.box = to_box(par->grid_ternd);
  for (int i = 0; i < 3; ++i)
    d.observer[i] = par->observer[i];
  d.rms = par->ternd_dft.rms;
  d.k0 = par->ternd_dft.k0;
  d.a0 = par->ternd_dft.a0;
  d.r0 = par->ternd_dft.r0;
  d.z0 = par->ternd_dft.z0;
  hamx_ctx *ctx = context();
  check(hamx_ternd_dft(ctx, &d, random_seed(par->brnd_seed), grid->te.get(), 0), ctx);
  grid->invalidate_device();
}

// ----------------------------------------------------------- integrator --
void Integrator::write_grid(const Breg *breg, const Brnd *, const TEreg *tereg,
                            const TErnd *, const CREfield *cre, const Grid_breg *gbreg,
                            const Grid_brnd *gbrnd, const Grid_tereg *gtereg,
                            const Grid_ternd *gternd, const Grid_cre *gcre, Grid_obs *gobs,
                            const Param *par) const {
  const auto &g = par->grid_obs;
  if (g.total_shell < 1 || g.total_shell > HAMX_MAX_SHELL)
    throw std::runtime_error("hamx: unsupported shell count");
  hamx_los_desc d;
  std::memset(&d, 0, sizeof(d));
  for (int i = 0; i < 3; ++i)
    d.observer[i] = par->observer[i];
  d.n_shell = static_cast<uint32_t>(g.total_shell);
  for (unsigned long long s = 0; s < g.total_shell; ++s)
    d.nside_shell[s] = static_cast<uint32_t>(g.nside_shell[s]);
  for (unsigned long long s = 0; s <= g.total_shell; ++s)
    d.radii_shell[s] = g.radii_shell[s];
  d.oc_r_res = g.oc_r_res;
  d.gc_r_min = g.gc_r_min;
  d.gc_r_max = g.gc_r_max;
  d.gc_z_min = g.gc_z_min;
  d.gc_z_max = g.gc_z_max;
  d.do_dm = g.do_dm;
  d.do_fd = g.do_fd;
  d.do_sync = g.do_sync.back();
  d.nside_dm = static_cast<uint32_t>(g.nside_dm);
  d.nside_fd = static_cast<uint32_t>(g.nside_fd);
  if (d.do_sync) {
    d.nside_sync = static_cast<uint32_t>(g.nside_sync.back());
    d.sync_freq = g.sim_sync_freq.back();
  }
  hamx_fields f;
  std::memset(&f, 0, sizeof(f));
  if (g.do_mask) { // Hampisk duplicated per shell on the device
    if (!gobs->mask_map)
      throw std::runtime_error("hamx: mask requested but Grid_obs holds no mask map");
    f.mask = gobs->mask_map->data.data();
    f.nside_mask = static_cast<uint32_t>(gobs->mask_map->nside);
  }
  // regular fields: device model, imported grid, or a user model baked
  std::unique_ptr<Grid_breg> baked_b;
  std::unique_ptr<Grid_tereg> baked_t;
  if (!par->grid_breg.build_permission) {
    d.breg_model = HAMX_BREG_NONE;
  } else if (par->grid_breg.read_permission) {
    d.breg_model = HAMX_BREG_GRID;
  } else {
    d.breg_model = breg ? breg->device_model(par, d.breg) : HAMX_BREG_NONE;
    if (d.breg_model < 0) {
      if (par->grid_breg.full_size == 0)
        throw std::runtime_error("regular B model has no device evaluator and no box_breg "
                                 "grid to bake it on");
      // what the model's read_field gives at the box_breg vertices
      baked_b = std::make_unique<Grid_breg>();
      baked_b->build_grid(par);
      Grid_breg *bb = baked_b.get();
      sample_vertices(par->grid_breg, [&](const Vec3 &p, unsigned long long i) {
        const Vec3 b = breg->read_field(p, par, gbreg);
        bb->bx[i] = b[0];
        bb->by[i] = b[1];
        bb->bz[i] = b[2];
      });
      gbreg = bb;
      d.breg_model = HAMX_BREG_GRID;
    }
  }
  if (d.breg_model == HAMX_BREG_GRID) {
    f.breg_x = gbreg->bx.get();
    f.breg_y = gbreg->by.get();
    f.breg_z = gbreg->bz.get();
    f.breg_box = to_box(par->grid_breg);
  }
  if (!par->grid_tereg.build_permission) {
    d.tereg_model = HAMX_TEREG_NONE;
  } else if (par->grid_tereg.read_permission) {
    d.tereg_model = HAMX_TEREG_GRID;
  } else {
    d.tereg_model = tereg ? tereg->device_model(par, d.tereg) : HAMX_TEREG_NONE;
    if (d.tereg_model < 0) {
      if (par->grid_tereg.full_size == 0)
        throw std::runtime_error("regular TE model has no device evaluator and no box_tereg "
                                 "grid to bake it on");
      baked_t = std::make_unique<Grid_tereg>();
      baked_t->build_grid(par);
      Grid_tereg *tb = baked_t.get();
      sample_vertices(par->grid_tereg, [&](const Vec3 &p, unsigned long long i) {
        tb->te[i] = tereg->read_field(p, par, gtereg);
      });
      gtereg = tb;
      d.tereg_model = HAMX_TEREG_GRID;
    }
  }
  if (d.tereg_model == HAMX_TEREG_GRID) {
    f.tereg = gtereg->te.get();
    f.tereg_box = to_box(par->grid_tereg);
  }
  d.cre_model = cre ? cre->device_model(par, d.cre) : HAMX_CRE_NONE;
  std::unique_ptr<Grid_cre> baked_c;
  if (par->grid_cre.read_permission) {
    d.cre_model = HAMX_CRE_GRID; // integrator.cc:287,356
  } else if (d.cre_model < 0) {
    // a model without a device evaluator: its flux on box_cre
    // (CREfield::write_grid, cre.cc:202-227), integrated as CRE_num
    if (par->grid_cre.cre_size == 0 || par->grid_cre.nE < 2)
      throw std::runtime_error("CRE model has no device evaluator and no box_cre grid to "
                               "write it on");
    baked_c = std::make_unique<Grid_cre>();
    baked_c->build_grid(par);
    cre->write_grid(par, baked_c.get());
    gcre = baked_c.get();
    d.cre_model = HAMX_CRE_GRID;
  }
  d.schedule = HAMX_SCHED_AUTO;

  // random grids: uploaded once, kept on the device across passes
  hamx_ctx *ctx = context();
  const bool brnd_on = par->grid_brnd.read_permission || par->grid_brnd.build_permission;
  const bool ternd_on = par->grid_ternd.read_permission || par->grid_ternd.build_permission;
  uint32_t flags = 0;
  const bool any_host_regular = d.breg_model == HAMX_BREG_GRID || d.tereg_model == HAMX_TEREG_GRID;
  if (brnd_on && gbrnd) {
    const size_t bytes = par->grid_brnd.full_size * sizeof(double);
    if (!gbrnd->dev_valid) {
      if (!gbrnd->dev_bx) {
        void *p;
        check(hamx_dev_alloc(ctx, bytes, &p), ctx);
        gbrnd->dev_bx = static_cast<double *>(p);
        check(hamx_dev_alloc(ctx, bytes, &p), ctx);
        gbrnd->dev_by = static_cast<double *>(p);
        check(hamx_dev_alloc(ctx, bytes, &p), ctx);
        gbrnd->dev_bz = static_cast<double *>(p);
      }
      check(hamx_memcpy_h2d(ctx, gbrnd->dev_bx, gbrnd->bx.get(), bytes), ctx);
      check(hamx_memcpy_h2d(ctx, gbrnd->dev_by, gbrnd->by.get(), bytes), ctx);
      check(hamx_memcpy_h2d(ctx, gbrnd->dev_bz, gbrnd->bz.get(), bytes), ctx);
      gbrnd->dev_valid = true;
    }
    f.brnd_x = gbrnd->dev_bx;
    f.brnd_y = gbrnd->dev_by;
    f.brnd_z = gbrnd->dev_bz;
    f.brnd_box = to_box(par->grid_brnd);
  }
  if (ternd_on && gternd) {
    const size_t bytes = par->grid_ternd.full_size * sizeof(double);
    if (!gternd->dev_valid) {
      if (!gternd->dev_te) {
        void *p;
        check(hamx_dev_alloc(ctx, bytes, &p), ctx);
        gternd->dev_te = static_cast<double *>(p);
      }
      check(hamx_memcpy_h2d(ctx, gternd->dev_te, gternd->te.get(), bytes), ctx);
      gternd->dev_valid = true;
    }
    f.ternd = gternd->dev_te;
    f.ternd_box = to_box(par->grid_ternd);
  }
  if (d.cre_model == HAMX_CRE_GRID) {
    if (!gcre || !gcre->cre_flux)
      throw std::runtime_error("hamx: CRE grid requested but Grid_cre holds no flux");
    const size_t bytes = par->grid_cre.cre_size * sizeof(double);
    if (!gcre->dev_valid) {
      if (!gcre->dev_flux) {
        void *p;
        check(hamx_dev_alloc(ctx, bytes, &p), ctx);
        gcre->dev_flux = static_cast<double *>(p);
      }
      check(hamx_memcpy_h2d(ctx, gcre->dev_flux, gcre->cre_flux.get(), bytes), ctx);
      gcre->dev_valid = true;
    }
    f.cre = gcre->dev_flux;
    f.cre_box = to_box(par->grid_cre);
    f.cre_E_min = par->grid_cre.E_min;
    f.cre_E_fact = par->grid_cre.E_fact;
    f.cre_nE = static_cast<uint32_t>(par->grid_cre.nE);
  }
  // every input of one call shares a residency: regular-field grids stay
  // on the host, so the random grids then go by their host copies
  if (any_host_regular) {
    if (f.brnd_x) {
      f.brnd_x = gbrnd->bx.get();
      f.brnd_y = gbrnd->by.get();
      f.brnd_z = gbrnd->bz.get();
    }
    if (f.ternd)
      f.ternd = gternd->te.get();
    if (f.cre)
      f.cre = gcre->cre_flux.get();
  } else if (f.brnd_x || f.ternd || f.cre) {
    flags |= HAMX_IN_DEVICE;
  }
  // the mask follows the residency of the other inputs
  struct DevMask {
    hamx_ctx *ctx = nullptr;
    void *p = nullptr;
    ~DevMask() {
      if (p)
        hamx_dev_free(ctx, p);
    }
  } dev_mask;
  if (f.mask && (flags & HAMX_IN_DEVICE)) {
    const size_t bytes = gobs->mask_map->npix() * sizeof(double);
    dev_mask.ctx = ctx;
    check(hamx_dev_alloc(ctx, bytes, &dev_mask.p), ctx);
    check(hamx_memcpy_h2d(ctx, dev_mask.p, f.mask, bytes), ctx);
    f.mask = static_cast<const double *>(dev_mask.p);
  }
  hamx_maps m;
  std::memset(&m, 0, sizeof(m));
  if (d.do_dm)
    m.dm = gobs->dm_map->data.data();
  if (d.do_sync) {
    m.is = gobs->is_map->data.data();
    m.qs = gobs->qs_map->data.data();
    m.us = gobs->us_map->data.data();
  }
  if (d.do_fd || d.do_sync)
    m.fd = gobs->fd_map->data.data();
  check(hamx_los_forward(ctx, &d, &f, &m, flags), ctx);
}

// ------------------------------------------------------------- pipeline --
Pipeline::Pipeline(const std::string &xml_file) { par = std::make_unique<Param>(xml_file); }

void Pipeline::assemble_grid() {
  grid_tereg = std::make_unique<Grid_tereg>(par.get());
  grid_breg = std::make_unique<Grid_breg>(par.get());
  grid_brnd = std::make_unique<Grid_brnd>(par.get());
  grid_ternd = std::make_unique<Grid_ternd>(par.get());
  grid_cre = std::make_unique<Grid_cre>(par.get());
  grid_obs = std::make_unique<Grid_obs>(par.get());
}

// pipeline.cc:33-85
void Pipeline::assemble_tereg() {
  if (!par->grid_tereg.build_permission) {
    tereg = std::make_unique<TEreg>();
    return;
  }
  if (par->grid_tereg.read_permission) {
    grid_tereg->import_grid(par.get());
    tereg = std::make_unique<TEreg>();
  } else if (par->tereg_type == "unif") {
    tereg = std::make_unique<TEreg_unif>();
  } else if (par->tereg_type == "ymw16") {
    tereg = std::make_unique<TEreg_ymw16>();
  } else {
    throw std::runtime_error("unsupported tereg model");
  }
  if (par->grid_tereg.write_permission) {
    tereg->write_grid(par.get(), grid_tereg.get());
    grid_tereg->export_grid(par.get());
  }
}

void Pipeline::assemble_breg() {
  if (!par->grid_breg.build_permission) {
    breg = std::make_unique<Breg>();
    return;
  }
  if (par->grid_breg.read_permission) {
    grid_breg->import_grid(par.get());
    breg = std::make_unique<Breg>();
  } else if (par->breg_type == "lsa") {
    breg = std::make_unique<Breg_lsa>();
  } else if (par->breg_type == "cart") {
    breg = std::make_unique<Breg_cart>();
  } else if (par->breg_type == "unif") {
    breg = std::make_unique<Breg_unif>();
  } else if (par->breg_type == "helix") {
    breg = std::make_unique<Breg_helix>();
  } else if (par->breg_type == "jf12") {
    breg = std::make_unique<Breg_jf12>();
  } else if (par->breg_type == "jaffe") {
    breg = std::make_unique<Breg_jaffe>();
  } else {
    throw std::runtime_error("unsupported breg model");
  }
  if (par->grid_breg.write_permission) {
    breg->write_grid(par.get(), grid_breg.get());
    grid_breg->export_grid(par.get());
  }
}

// pipeline.cc:87-110
void Pipeline::assemble_ternd() {
  if (par->grid_ternd.read_permission) {
    grid_ternd->import_grid(par.get());
    ternd = std::make_unique<TErnd>();
  } else if (par->grid_ternd.build_permission) {
    if (par->ternd_type == "global" && par->ternd_method == "dft")
      ternd = std::make_unique<TErnd_dft>();
    else
      throw std::runtime_error("unsupported ternd model");
    ternd->write_grid(par.get(), tereg.get(), grid_tereg.get(), grid_ternd.get());
  } else {
    ternd = std::make_unique<TErnd>();
  }
  if (par->grid_ternd.write_permission)
    grid_ternd->export_grid(par.get());
}

// pipeline.cc:112-143
void Pipeline::assemble_brnd() {
  if (par->grid_brnd.read_permission) {
    grid_brnd->import_grid(par.get());
    brnd = std::make_unique<Brnd>();
  } else if (par->grid_brnd.build_permission) {
    if (par->brnd_type == "global" && par->brnd_method == "es")
      brnd = std::make_unique<Brnd_es>();
    else if (par->brnd_type == "global" && par->brnd_method == "jf12")
      brnd = std::make_unique<Brnd_jf12>();
    else if (par->brnd_type == "local" && par->brnd_method == "mhd")
      brnd = std::make_unique<Brnd_mhd>();
    else
      throw std::runtime_error("unsupported brnd model");
    brnd->write_grid(par.get(), breg.get(), grid_breg.get(), grid_brnd.get());
  } else {
    brnd = std::make_unique<Brnd>();
  }
  if (par->grid_brnd.write_permission)
    grid_brnd->export_grid(par.get());
}

// pipeline.cc:145-163
void Pipeline::assemble_cre() {
  if (par->grid_cre.read_permission) {
    grid_cre->import_grid(par.get());
    cre = std::make_unique<CRE_num>();
  } else if (par->grid_cre.build_permission) {
    if (par->cre_type == "analytic")
      cre = std::make_unique<CRE_ana>();
    else if (par->cre_type == "unif")
      cre = std::make_unique<CRE_unif>();
    else
      throw std::runtime_error("unsupported cre model");
  }
  if (par->grid_cre.write_permission) {
    if (!cre)
      throw std::runtime_error("hamx: CRE grid output needs a CRE model");
    cre->write_grid(par.get(), grid_cre.get());
    grid_cre->export_grid(par.get());
  }
}

// pipeline.cc:165-190
void Pipeline::assemble_obs() {
  intobj = std::make_unique<Integrator>();
  if (!par->grid_obs.write_permission)
    return;
  const auto repeat = par->grid_obs.do_sync.size();
  for (size_t i = 0; i < repeat; ++i) {
    if (i > 0) {
      par->grid_obs.do_dm = false;
      par->grid_obs.do_fd = false;
      grid_obs->build_grid(par.get());
    }
    intobj->write_grid(breg.get(), brnd.get(), tereg.get(), ternd.get(), cre.get(),
                       grid_breg.get(), grid_brnd.get(), grid_tereg.get(), grid_ternd.get(),
                       grid_cre.get(), grid_obs.get(), par.get());
    grid_obs->export_grid(par.get());
    if (par->grid_obs.do_sync.back()) {
      par->grid_obs.nside_sync.pop_back();
      par->grid_obs.do_sync.pop_back();
      par->grid_obs.sim_sync_freq.pop_back();
      par->grid_obs.sim_sync_name.pop_back();
    }
  }
}

} // namespace hamx_host
