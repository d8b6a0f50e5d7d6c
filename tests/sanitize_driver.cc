// Host code under AddressSanitizer and UndefinedBehaviorSanitizer (built by
// tests/test_sanitizers.py with g++, no GPU code): the XML reader and Param
// on a valid parameter file and on malformed ones, and the CPU oracle's
// restatements on small inputs of each kind its tests use (a multi-shell LOS
// pass with random grids, a mask and an imported CRE grid; the generators on
// power-of-two, mixed-radix and prime grids; HEALPix and synchrotron helpers).
//
//   sanitize_driver <valid.xml> <malformed.xml>...
#include <cmath>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <vector>

#include "hamx_host.h"
#include "hamx_oracle.h"

namespace {

const double kKpc = 3.0856775806e+21;
const double kGeV = 1.60217733e-3;

hamx_box box(uint64_t nx, uint64_t ny, uint64_t nz) {
  hamx_box b;
  b.nx = nx;
  b.ny = ny;
  b.nz = nz;
  b.x_min = -20 * kKpc;
  b.x_max = 20 * kKpc;
  b.y_min = -20 * kKpc;
  b.y_max = 20 * kKpc;
  b.z_min = -4 * kKpc;
  b.z_max = 4 * kKpc;
  return b;
}

int los_pass(bool cre_grid) {
  hamx_los_desc d;
  std::memset(&d, 0, sizeof(d));
  d.observer[0] = -8.3 * kKpc;
  d.observer[2] = 0.006 * kKpc;
  d.n_shell = 2;
  d.nside_shell[0] = 2;
  d.nside_shell[1] = 4;
  d.radii_shell[0] = 0;
  d.radii_shell[1] = 5 * kKpc;
  d.radii_shell[2] = 12 * kKpc;
  d.oc_r_res = 0.2 * kKpc;
  d.gc_r_max = 20 * kKpc;
  d.gc_z_min = -10 * kKpc;
  d.gc_z_max = 10 * kKpc;
  d.do_dm = d.do_fd = d.do_sync = 1;
  d.nside_dm = d.nside_fd = d.nside_sync = 4;
  d.sync_freq = 23e9;
  d.breg_model = HAMX_BREG_LSA;
  d.breg[0] = 6e-6;
  d.breg[1] = 0.47;
  d.breg[2] = 0.016;
  d.breg[3] = 0.44;
  d.tereg_model = HAMX_TEREG_UNIF;
  d.tereg[0] = 0.01;
  d.tereg[1] = 3 * kKpc;
  d.cre_model = HAMX_CRE_ANA;
  const double cre[7] = {3.0, 0.0, 0.0, 5 * kKpc, 1 * kKpc, 20.6 * kGeV, 0.0217};
  std::memcpy(d.cre, cre, sizeof(cre));
  const hamx_box g = box(8, 8, 4);
  const uint64_t n = 8 * 8 * 4;
  std::vector<double> bx(n), by(n), bz(n), te(n), mask(12 * 2 * 2, 1.0);
  for (uint64_t i = 0; i < n; ++i) {
    bx[i] = 1e-6 * std::sin(0.7 * i);
    by[i] = 1e-6 * std::cos(0.3 * i);
    bz[i] = 5e-7 * std::sin(1.1 * i);
    te[i] = 0.002 * std::cos(0.9 * i);
  }
  mask[3] = 0.0;
  hamx_fields f;
  std::memset(&f, 0, sizeof(f));
  f.brnd_x = bx.data();
  f.brnd_y = by.data();
  f.brnd_z = bz.data();
  f.brnd_box = g;
  f.ternd = te.data();
  f.ternd_box = g;
  f.mask = mask.data();
  f.nside_mask = 2;
  std::vector<double> flux;
  if (cre_grid) {
    const hamx_box cb = box(6, 6, 4);
    const uint32_t nE = 8;
    flux.assign(6 * 6 * 4 * nE, 0.);
    if (orc_cre_write_grid(&d, &cb, 0.5 * kGeV, std::log(100.) / (nE - 1), nE, flux.data()))
      return 1;
    d.cre_model = HAMX_CRE_GRID;
    f.cre = flux.data();
    f.cre_box = cb;
    f.cre_E_min = 0.5 * kGeV;
    f.cre_E_fact = std::log(100.) / (nE - 1);
    f.cre_nE = nE;
  }
  std::vector<double> dm(192), fd(192), is(192), qs(192), us(192);
  hamx_maps m{dm.data(), fd.data(), is.data(), qs.data(), us.data()};
  if (orc_write_grid(&d, &f, &m, 2))
    return 1;
  for (double v : qs)
    if (!std::isfinite(v))
      return 1;
  return 0;
}

int generators() {
  const uint64_t shapes[][3] = {{8, 8, 4}, {12, 10, 6}, {7, 9, 5}, {13, 11, 3}};
  for (const auto &s : shapes) {
    const uint64_t n = s[0] * s[1] * s[2];
    std::vector<double> bx(n), by(n), bz(n), te(n), c(2 * n, 1.0);
    hamx_es_desc e;
    std::memset(&e, 0, sizeof(e));
    e.box = box(s[0], s[1], s[2]);
    e.observer[0] = -8.3 * kKpc;
    e.rms = 2e-6;
    e.k0 = e.k1 = 0.1;
    e.a0 = 1.7;
    e.rho = 0.5;
    e.r0 = 8 * kKpc;
    e.z0 = 1 * kKpc;
    e.breg_model = HAMX_BREG_LSA;
    e.breg[0] = 6e-6;
    e.breg[1] = 0.47;
    if (orc_brnd_es(&e, 3, bx.data(), by.data(), bz.data()))
      return 1;
    hamx_dft_desc t;
    std::memset(&t, 0, sizeof(t));
    t.box = e.box;
    t.observer[0] = -8.3 * kKpc;
    t.rms = 0.001;
    t.k0 = 0.05;
    t.a0 = -1.7;
    t.r0 = 8 * kKpc;
    t.z0 = 1 * kKpc;
    if (orc_ternd_dft(&t, 4, te.data()))
      return 1;
    orc_fft3d(c.data(), s[0], s[1], s[2], 1);
    orc_fft3d(c.data(), s[0], s[1], s[2], -1);
  }
  return 0;
}

int helpers() {
  std::vector<double> tp(2 * 12 * 16 * 16);
  orc_pix2ang_map(16, tp.data());
  std::vector<double> piv(12 * 8 * 8, 1.0), out(12 * 32 * 32);
  orc_mask_duplicate(piv.data(), 8, 32, out.data());
  std::vector<double> x{1e-6, 0.1, 1.0, 10.0, 99.0}, F(5), G(5);
  orc_sync_kernels(x.data(), x.size(), F.data(), G.data());
  std::vector<double> a(12 * 4 * 4, 1.0), b(12 * 8 * 8, 0.0);
  orc_accumulate(b.data(), 8, a.data(), 4);
  orc_accumulate(a.data(), 4, b.data(), 8);
  return 0;
}

} // namespace

int main(int argc, char **argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: sanitize_driver <valid.xml> <malformed.xml>...\n");
    return 2;
  }
  hamx_host::Param par(argv[1]);
  std::printf("valid: %llu shells, %zu sync entries\n",
              static_cast<unsigned long long>(par.grid_obs.total_shell),
              par.grid_obs.do_sync.size());
  for (int i = 2; i < argc; ++i) {
    try {
      hamx_host::Param bad(argv[i]);
      std::printf("malformed %s: accepted\n", argv[i]);
      return 3;
    } catch (const std::runtime_error &e) {
      std::printf("malformed %s: %s\n", argv[i], e.what());
    }
  }
  if (los_pass(false) || los_pass(true)) {
    std::printf("los pass failed\n");
    return 4;
  }
  if (generators() || helpers()) {
    std::printf("generators failed\n");
    return 5;
  }
  std::printf("sanitized run ok\n");
  return 0;
}
