// User field models written against the reference's plugin headers
// (bfield.h, tefield.h, crefield.h, pipeline.h) and installed the
// reference's way, by a Pipeline subclass.  Built by tests/test_host_plugins.py
// against libhamx_host with hammurabix_amd/host/include first on the path.
//
//   user_plugins <params.xml>               run the pipeline with the models
//   user_plugins --host-check <params.xml>  host-side evaluations only (no GPU)
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <stdexcept>
#include <string>

#include <bfield.h>
#include <crefield.h>
#include <hamtype.h>
#include <hamvec.h>
#include <param.h>
#include <pipeline.h>
#include <tefield.h>

namespace {
const ham_float kpc = 3.0856775806e+21;
const ham_float muG = 1e-6;
const ham_float GeV = 1.60217733e-3;
} // namespace

// a field that turns with position: not linear, so the baked grid is an
// approximation of it, as for any user model
class Breg_user final : public Breg {
public:
  Hamvec<3, ham_float> write_field(const Hamvec<3, ham_float> &pos,
                                   const Param *) const override {
    const ham_float a = 0.3 * pos[0] / kpc + 0.1 * pos[1] / kpc;
    return Hamvec<3, ham_float>(3. * muG * std::cos(a), 3. * muG * std::sin(a),
                                0.5 * muG * std::tanh(pos[2] / kpc));
  }
};

class TEreg_user final : public TEreg {
public:
  ham_float write_field(const Hamvec<3, ham_float> &pos, const Param *) const override {
    const ham_float r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
    return r > 15. * kpc ? 0. : 0.02 * std::exp(-std::fabs(pos[2]) / kpc);
  }
};

// PHI(E) [GeV m^2 s sr]^-1: a power law of index 3 above 20 GeV's norm with
// an exponential disc
class CRE_user final : public CREfield {
public:
  ham_float write_field(const Hamvec<3, ham_float> &pos, const ham_float &E,
                        const Param *) const override {
    const ham_float r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);
    return 0.02 * std::pow(E / (20. * GeV), -3.) * std::exp(-(r - 8.3 * kpc) / (5. * kpc)) *
           std::exp(-std::fabs(pos[2]) / kpc);
  }
};

class UserPipeline final : public Pipeline {
public:
  using Pipeline::Pipeline;
  void assemble_breg() override {
    breg = std::make_unique<Breg_user>();
    if (par->grid_breg.write_permission) {
      breg->write_grid(par.get(), grid_breg.get());
      grid_breg->export_grid(par.get());
    }
  }
  void assemble_tereg() override {
    tereg = std::make_unique<TEreg_user>();
    if (par->grid_tereg.write_permission) {
      tereg->write_grid(par.get(), grid_tereg.get());
      grid_tereg->export_grid(par.get());
    }
  }
  void assemble_cre() override {
    cre = std::make_unique<CRE_user>();
    if (par->grid_cre.write_permission) {
      cre->write_grid(par.get(), grid_cre.get());
      grid_cre->export_grid(par.get());
    }
  }
};

// Host evaluations printed as "key v0 v1 ..." lines.
void host_check(const std::string &xml) {
  Param par(xml);
  const Hamvec<3, ham_float> p(-5. * kpc, 2. * kpc, 0.4 * kpc);
  // vectors
  const Hamvec<3, ham_float> u(1., 2., 3.), v(-2., 0.5, 4.);
  const auto c = u.crossprod(v);
  std::printf("hamvec %.17g %.17g %.17g %.17g %.17g %.17g\n", c[0], c[1], c[2], u.dotprod(v),
              (u - v).length(), u.versor()[2]);
  // a user model through the base class dispatch (breg.cc:11-21)
  Breg_user b;
  const auto bf = b.read_field(p, &par, nullptr);
  const auto bw = b.write_field(p, &par);
  std::printf("breg_user %.17g %.17g %.17g %.17g %.17g %.17g\n", bf[0], bf[1], bf[2], bw[0],
              bw[1], bw[2]);
  TEreg_user t;
  std::printf("tereg_user %.17g %.17g\n", t.read_field(p, &par, nullptr), t.write_field(p, &par));
  // the library's CRE models on the host, at a position and energies
  CRE_ana ana;
  CRE_unif unif;
  for (const ham_float E : {5. * GeV, 50. * GeV}) {
    std::printf("cre_ana %.17g %.17g %.17g %.17g %.17g\n", E, ana.write_field(p, E, &par),
                ana.flux_idx(p, &par), ana.spatial_profile(p, &par), ana.flux_norm(p, &par));
  }
  std::printf("cre_unif %.17g %.17g %.17g\n", unif.write_field(par.observer, 7. * GeV, &par),
              unif.flux_idx(p, &par), unif.spatial_profile(p, &par));
  // the base class throws where the reference's does
  CREfield base;
  try {
    base.flux_idx(p, &par);
    std::printf("base_throws 0\n");
  } catch (const std::runtime_error &) {
    std::printf("base_throws 1\n");
  }
  // read_grid_num on a field linear in position and energy index
  Grid_cre g;
  g.build_grid(&par);
  const auto &gc = par.grid_cre;
  for (unsigned long long i = 0; i < gc.nx; ++i)
    for (unsigned long long j = 0; j < gc.ny; ++j)
      for (unsigned long long k = 0; k < gc.nz; ++k)
        for (unsigned long long m = 0; m < gc.nE; ++m) {
          const ham_float x = gc.x_min + i * (gc.x_max - gc.x_min) / (gc.nx - 1);
          const ham_float y = gc.y_min + j * (gc.y_max - gc.y_min) / (gc.ny - 1);
          const ham_float z = gc.z_min + k * (gc.z_max - gc.z_min) / (gc.nz - 1);
          g.cre_flux[((i * gc.ny + j) * gc.nz + k) * gc.nE + m] =
              1. + 2. * x / kpc - 0.5 * y / kpc + 3. * z / kpc + 0.25 * m;
        }
  CREfield reader;
  for (const ham_uint m : {ham_uint(0), ham_uint(5)})
    std::printf("read_grid_num %.17g %.17g\n", static_cast<double>(m),
                reader.read_grid_num(p, m, &par, &g));
}

int main(int argc, char **argv) {
  try {
    if (argc == 3 && std::string(argv[1]) == "--host-check") {
      host_check(argv[2]);
      return EXIT_SUCCESS;
    }
    if (argc != 2) {
      std::fprintf(stderr, "usage: user_plugins [--host-check] <params.xml>\n");
      return EXIT_FAILURE;
    }
    UserPipeline run(argv[1]);
    run.assemble_grid();
    run.assemble_tereg();
    run.assemble_breg();
    run.assemble_ternd();
    run.assemble_brnd();
    run.assemble_cre();
    run.assemble_obs();
  } catch (const std::exception &e) {
    std::fprintf(stderr, "%s\n", e.what());
    return EXIT_FAILURE;
  }
  return EXIT_SUCCESS;
}
