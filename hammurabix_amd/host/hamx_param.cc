// Param: XML parameter set -> CGS values (param.cc:12-778 restated over
// hamx::xml).  Unsupported type strings throw the reference's messages.
#include <cmath>
#include <stdexcept>
#include <string>

#include "hamx_host.h"
#include "hamx_xml.h"

namespace hamx_host {

namespace {

namespace units {
const double pi = 3.141592653589793238462643383279502884197;
const double rad = pi / 180.;
const double kpc = 3.0856775806e+21;
const double pc = 3.0856775806e+18;
const double eV = 1.60217733e-12;
const double GeV = 1.e9 * eV;
const double GHz = 1.e9;
const double muGauss = 1.e-6;
} // namespace units

using hamx::xml::Element;
using hamx::xml::trace;

double fval(const Element *e, const char *key) { return trace(e, {key})->attr_double("value"); }
unsigned long long uval(const Element *e, const char *key) {
  return trace(e, {key})->attr_uint("value");
}

void read_box(const Element *root, const char *name, Param::param_grid &g) {
  const Element *b = trace(root, {"grid", name});
  g.nx = uval(b, "nx");
  g.ny = uval(b, "ny");
  g.nz = uval(b, "nz");
  g.full_size = g.nx * g.ny * g.nz;
  g.x_max = units::kpc * fval(b, "x_max");
  g.x_min = units::kpc * fval(b, "x_min");
  g.y_max = units::kpc * fval(b, "y_max");
  g.y_min = units::kpc * fval(b, "y_min");
  g.z_max = units::kpc * fval(b, "z_max");
  g.z_min = units::kpc * fval(b, "z_min");
}

bool has_box(const Element *root, const char *name) {
  const Element *g = root->child("grid");
  return g && g->child(name);
}

void read_fieldio(const Element *root, const char *name, Param::param_grid &g) {
  const Element *fio = trace(root, {"fieldio"});
  if (const Element *e = fio->child(name)) {
    g.read_permission = e->attr_bool("read");
    g.write_permission = e->attr_bool("write");
    g.filename = e->attr_string("filename");
  }
}

} // namespace

Param::Param(const std::string &xml_file) {
  const hamx::xml::Document doc = hamx::xml::load(xml_file);
  const Element *root = doc.root.get();
  if (!root || root->name != "root")
    throw std::runtime_error("parameter file has no <root> element");
  const Element *o = trace(root, {"grid", "observer"});
  observer = Vec3(units::kpc * fval(o, "x"), units::kpc * fval(o, "y"),
                  units::kpc * fval(o, "z"));

  // observables and shells (param.cc:30-144)
  const Element *ob = trace(root, {"observable"});
  if (const Element *e = ob->child("dm")) {
    grid_obs.write_permission = true;
    grid_obs.do_dm = e->attr_bool("cue");
    grid_obs.sim_dm_name = e->attr_string("filename");
    grid_obs.nside_dm = e->attr_uint("nside");
  }
  if (const Element *e = ob->child("faraday")) {
    grid_obs.write_permission = true;
    grid_obs.do_fd = e->attr_bool("cue");
    grid_obs.sim_fd_name = e->attr_string("filename");
    grid_obs.nside_fd = e->attr_uint("nside");
  }
  const auto syncs = ob->children_named("sync");
  if (!syncs.empty()) {
    grid_obs.write_permission = true;
    for (const Element *e : syncs) {
      grid_obs.do_sync.push_back(e->attr_bool("cue"));
      grid_obs.sim_sync_freq.push_back(e->attr_double("freq") * units::GHz);
      grid_obs.sim_sync_name.push_back(e->attr_string("filename"));
      grid_obs.nside_sync.push_back(e->attr_uint("nside"));
    }
  } else {
    grid_obs.do_sync.push_back(false);
  }
  if (grid_obs.write_permission) {
    const Element *sh = trace(root, {"grid", "shell"});
    grid_obs.oc_r_min = fval(sh, "oc_r_min") * units::kpc;
    grid_obs.oc_r_max = fval(sh, "oc_r_max") * units::kpc;
    grid_obs.gc_r_min = fval(sh, "gc_r_min") * units::kpc;
    grid_obs.gc_r_max = fval(sh, "gc_r_max") * units::kpc;
    grid_obs.gc_z_min = fval(sh, "gc_z_min") * units::kpc;
    grid_obs.gc_z_max = fval(sh, "gc_z_max") * units::kpc;
    grid_obs.oc_r_res = fval(sh, "oc_r_res") * units::kpc;
    const std::string layer = trace(sh, {"layer"})->attr_string("type");
    if (layer == "auto") {
      const Element *a = trace(sh, {"layer", "auto"});
      grid_obs.total_shell = uval(a, "shell_num");
      const unsigned long long nside_sim = uval(a, "nside_sim");
      for (unsigned long long i = 0; i != grid_obs.total_shell; ++i)
        grid_obs.nside_shell.push_back(
            static_cast<unsigned long long>(std::pow(2, i) * nside_sim));
      grid_obs.radii_shell.push_back(grid_obs.oc_r_min);
      for (unsigned long long i = 0; i < grid_obs.total_shell; ++i)
        grid_obs.radii_shell.push_back(
            (grid_obs.oc_r_max - grid_obs.oc_r_min) *
                std::pow(0.5, static_cast<double>(grid_obs.total_shell - i - 1)) +
            grid_obs.oc_r_min);
    } else if (layer == "manual") {
      const Element *m = trace(sh, {"layer", "manual"});
      for (const Element *e : m->children_named("nside_sim"))
        grid_obs.nside_shell.push_back(e->attr_uint("value"));
      grid_obs.total_shell = grid_obs.nside_shell.size();
      for (const Element *e : m->children_named("cut"))
        grid_obs.cut_shell.push_back(e->attr_double("value"));
      grid_obs.cut_shell.push_back(1.);
      if (grid_obs.cut_shell.size() != grid_obs.total_shell)
        throw std::runtime_error("manual layer: cut count must be nside_sim count - 1");
      grid_obs.radii_shell.push_back(grid_obs.oc_r_min);
      for (double c : grid_obs.cut_shell)
        grid_obs.radii_shell.push_back((grid_obs.oc_r_max - grid_obs.oc_r_min) * c +
                                       grid_obs.oc_r_min);
    } else {
      throw std::runtime_error("unsupported layer option");
    }
    const Element *mk = trace(root, {"mask"});
    if (mk->attr_bool("cue")) {
      grid_obs.do_mask = true;
      grid_obs.mask_name = mk->attr_string("filename");
      grid_obs.nside_mask = mk->attr_uint("nside");
    }
  }

  // regular magnetic field (param.cc:146-343)
  read_fieldio(root, "breg", grid_breg);
  const Element *mf = trace(root, {"magneticfield"});
  grid_breg.build_permission = trace(mf, {"regular"})->attr_bool("cue");
  if (grid_breg.build_permission && !grid_breg.read_permission) {
    breg_type = trace(mf, {"regular"})->attr_string("type");
    if (breg_type == "lsa") {
      const Element *s = trace(mf, {"regular", "lsa"});
      breg_lsa.b0 = fval(s, "b0") * units::muGauss;
      breg_lsa.psi0 = fval(s, "psi0") * units::rad;
      breg_lsa.psi1 = fval(s, "psi1") * units::rad;
      breg_lsa.chi0 = fval(s, "chi0") * units::rad;
    } else if (breg_type == "unif") {
      const Element *s = trace(mf, {"regular", "unif"});
      breg_unif.bp = fval(s, "bp") * units::muGauss;
      breg_unif.bv = fval(s, "bv") * units::muGauss;
      breg_unif.l0 = fval(s, "l0") * units::rad;
    } else if (breg_type == "cart") {
      const Element *s = trace(mf, {"regular", "cart"});
      breg_cart.bx = fval(s, "bx") * units::muGauss;
      breg_cart.by = fval(s, "by") * units::muGauss;
      breg_cart.bz = fval(s, "bz") * units::muGauss;
    } else if (breg_type == "helix") { // param.cc:258-270
      const Element *s = trace(mf, {"regular", "helix"});
      breg_helix.bx = fval(s, "bx") * units::muGauss;
      breg_helix.by = fval(s, "by") * units::muGauss;
      breg_helix.bz = fval(s, "bz") * units::muGauss;
      breg_helix.r_min = fval(s, "r_min") * units::kpc;
      breg_helix.r_max = fval(s, "r_max") * units::kpc;
    } else if (breg_type == "jf12") { // param.cc:271-322
      const Element *s = trace(mf, {"regular", "jf12"});
      const char *mu[] = {"b_arm1", "b_arm2", "b_arm3", "b_arm4", "b_arm5", "b_arm6",
                          "b_arm7", "b_ring"};
      for (int i = 0; i < 8; ++i)
        breg_jf12[i] = fval(s, mu[i]) * units::muGauss;
      breg_jf12[8] = fval(s, "h_disk") * units::kpc;
      breg_jf12[9] = fval(s, "w_disk") * units::kpc;
      breg_jf12[10] = fval(s, "Bn") * units::muGauss;
      breg_jf12[11] = fval(s, "Bs") * units::muGauss;
      breg_jf12[12] = fval(s, "rn") * units::kpc;
      breg_jf12[13] = fval(s, "rs") * units::kpc;
      breg_jf12[14] = fval(s, "wh") * units::kpc;
      breg_jf12[15] = fval(s, "z0") * units::kpc;
      breg_jf12[16] = fval(s, "B0_X") * units::muGauss;
      breg_jf12[17] = fval(s, "Xtheta") * units::rad;
      breg_jf12[18] = fval(s, "rpc_X") * units::kpc;
      breg_jf12[19] = fval(s, "r0_X") * units::kpc;
    } else if (breg_type == "jaffe") { // param.cc:173-257
      const Element *s = trace(mf, {"regular", "jaffe"});
      auto cue = [&](const char *k) { return trace(s, {k})->attr_bool("cue") ? 1. : 0.; };
      double *q = breg_jaffe;
      q[0] = cue("quadruple");
      q[1] = cue("bss");
      q[2] = fval(s, "disk_amp") * units::muGauss;
      q[3] = fval(s, "disk_z0") * units::kpc;
      q[4] = fval(s, "halo_amp") * units::muGauss;
      q[5] = fval(s, "halo_z0") * units::kpc;
      q[6] = fval(s, "r_inner") * units::kpc;
      q[7] = fval(s, "r_scale") * units::kpc;
      q[8] = fval(s, "r_peak") * units::kpc;
      q[9] = cue("ring");
      q[10] = cue("bar");
      q[11] = fval(s, "ring_amp") * units::muGauss;
      q[12] = fval(s, "ring_r") * units::kpc;
      q[13] = fval(s, "bar_amp") * units::muGauss;
      q[14] = fval(s, "bar_a") * units::kpc;
      q[15] = fval(s, "bar_b") * units::kpc;
      q[16] = fval(s, "bar_phi0") * units::rad;
      const unsigned long long arms = uval(s, "arm_num");
      if (arms > 4)
        throw std::runtime_error("jaffe: at most 4 spiral arms (arm_phi1..4)");
      q[17] = static_cast<double>(arms);
      q[18] = fval(s, "arm_r0") * units::kpc;
      q[19] = fval(s, "arm_z0") * units::kpc;
      const char *phi[] = {"arm_phi1", "arm_phi2", "arm_phi3", "arm_phi4"};
      const char *amp[] = {"arm_amp1", "arm_amp2", "arm_amp3", "arm_amp4"};
      for (int i = 0; i < 4; ++i) {
        q[20 + i] = fval(s, phi[i]) * units::rad;
        q[24 + i] = fval(s, amp[i]) * units::muGauss;
      }
      q[28] = fval(s, "arm_pitch") * units::rad;
      q[29] = fval(s, "comp_c");
      q[30] = fval(s, "comp_d") * units::kpc;
      q[31] = fval(s, "comp_r") * units::kpc;
      q[32] = fval(s, "comp_p");
    } else {
      throw std::runtime_error("unsupported breg model");
    }
  }
  // also without field I/O when the file has the box: a model without a
  // device evaluator is baked on it
  if (grid_breg.read_permission || grid_breg.write_permission || has_box(root, "box_breg"))
    read_box(root, "box_breg", grid_breg);

  // random magnetic field (param.cc:345-460)
  read_fieldio(root, "brnd", grid_brnd);
  grid_brnd.build_permission = trace(mf, {"random"})->attr_bool("cue");
  if (grid_brnd.build_permission && !grid_brnd.read_permission) {
    brnd_seed = trace(mf, {"random"})->attr_uint("seed");
    brnd_type = trace(mf, {"random"})->attr_string("type");
    if (brnd_type == "global") {
      brnd_method = trace(mf, {"random", "global"})->attr_string("type");
      if (brnd_method == "es") {
        const Element *s = trace(mf, {"random", "global", "es"});
        brnd_es.rms = fval(s, "rms") * units::muGauss;
        brnd_es.k0 = fval(s, "k0");
        brnd_es.a0 = fval(s, "a0");
        brnd_es.k1 = fval(s, "k1");
        brnd_es.a1 = fval(s, "a1");
        brnd_es.rho = fval(s, "rho");
        brnd_es.r0 = fval(s, "r0") * units::kpc;
        brnd_es.z0 = fval(s, "z0") * units::kpc;
      } else if (brnd_method == "jf12") {
        const Element *s = trace(mf, {"random", "global", "jf12"});
        brnd_es.rms = fval(s, "rms") * units::muGauss;
        brnd_es.k0 = fval(s, "k0");
        brnd_es.a0 = fval(s, "a0");
        brnd_es.k1 = fval(s, "k1");
        brnd_es.a1 = fval(s, "a1");
        brnd_es.rho = fval(s, "rho");
        const char *keys[13] = {"b0_1", "b0_2", "b0_3", "b0_4", "b0_5", "b0_6", "b0_7",
                                "b0_8", "b0_int", "z0_spiral", "b0_halo", "r0_halo", "z0_halo"};
        for (int i = 0; i < 13; ++i) {
          const bool length = i == 9 || i == 11 || i == 12;
          brnd_jf12[i] = fval(s, keys[i]) * (length ? units::kpc : units::muGauss);
        }
      } else if (brnd_method != "jaffe") {
        throw std::runtime_error("unsupported brnd model");
      }
    } else if (brnd_type == "local") {
      brnd_method = trace(mf, {"random", "local"})->attr_string("type");
      if (brnd_method != "mhd")
        throw std::runtime_error("unsupported brnd model");
      const Element *s = trace(mf, {"random", "local", "mhd"});
      const double mu2 = units::muGauss * units::muGauss;
      brnd_mhd.pa0 = fval(s, "pa0") * mu2;
      brnd_mhd.pf0 = fval(s, "pf0") * mu2;
      brnd_mhd.ps0 = fval(s, "ps0") * mu2;
      brnd_mhd.k0 = fval(s, "k0");
      brnd_mhd.aa0 = fval(s, "aa0");
      brnd_mhd.af0 = fval(s, "af0");
      brnd_mhd.as0 = fval(s, "as0");
      brnd_mhd.k1 = fval(s, "k1");
      brnd_mhd.a1 = fval(s, "a1");
      brnd_mhd.beta = fval(s, "beta");
      brnd_mhd.ma = fval(s, "ma");
    } else {
      throw std::runtime_error("unsupported brnd type");
    }
  }
  if (grid_brnd.read_permission || grid_brnd.write_permission || grid_brnd.build_permission)
    read_box(root, "box_brnd", grid_brnd);

  // regular thermal electrons (param.cc:462-667)
  read_fieldio(root, "tereg", grid_tereg);
  const Element *te = trace(root, {"thermalelectron"});
  grid_tereg.build_permission = trace(te, {"regular"})->attr_bool("cue");
  if (grid_tereg.build_permission) {
    tereg_type = trace(te, {"regular"})->attr_string("type");
    if (tereg_type == "unif") {
      const Element *s = trace(te, {"regular", "unif"});
      tereg_unif.n0 = fval(s, "n0");
      tereg_unif.r0 = fval(s, "r0") * units::kpc;
    } else if (tereg_type == "ymw16") {
      auto blk = [&](const char *name) { return trace(te, {"regular", "ymw16", name}); };
      const Element *w = blk("warp"), *tk = blk("thickdisk"), *tn = blk("thindisk");
      const Element *sa = blk("spiralarm"), *gc = blk("galcenter"), *gn = blk("gumnebula");
      const Element *lb = blk("localbubble"), *li = blk("loopi");
      const double pc = units::pc, kpc = units::kpc, rad = units::rad;
      double *p = tereg_ymw16;
      int n = 0;
      p[n++] = fval(w, "r_warp") * kpc;
      p[n++] = fval(w, "r0") * kpc;
      p[n++] = fval(w, "gamma_w");
      for (const char *k : {"ad", "bd"})
        p[n++] = fval(tk, k) * pc;
      p[n++] = fval(tk, "n1");
      p[n++] = fval(tk, "h1") * pc;
      for (const char *k : {"a2", "b2"})
        p[n++] = fval(tn, k) * pc;
      p[n++] = fval(tn, "n2");
      p[n++] = fval(tn, "k2");
      p[n++] = fval(sa, "b2s") * pc;
      auto arm = [&](const char *stem, double unit) {
        for (int i = 0; i < 5; ++i)
          p[n++] = fval(sa, (std::string(stem) + std::to_string(i)).c_str()) * unit;
      };
      arm("ele_arm_", 1.);
      arm("wid_arm_", pc);
      arm("rref_arm_", kpc);
      arm("phiref_arm_", rad);
      for (int i = 0; i < 5; ++i)
        p[n++] = std::tan(fval(sa, ("pitch_arm_" + std::to_string(i)).c_str()) * rad);
      for (int i = 0; i < 5; ++i)
        p[n++] = std::cos(fval(sa, ("pitch_arm_" + std::to_string(i)).c_str()) * rad);
      p[n++] = fval(sa, "aa") * pc;
      for (const char *k : {"ka", "ncn", "thetacn", "wcn", "nsg", "thetasg", "wsg"})
        p[n++] = fval(sa, k);
      p[n++] = fval(gc, "ngc");
      p[n++] = fval(gc, "agc") * pc;
      p[n++] = fval(gc, "hgc") * pc;
      p[n++] = fval(gn, "ngn");
      p[n++] = fval(gn, "wgn") * pc;
      p[n++] = fval(gn, "agn") * pc;
      p[n++] = fval(gn, "kgn");
      p[n++] = fval(lb, "j_lb");
      for (const char *c : {"1", "2"}) {
        const std::string s(c);
        p[n++] = fval(lb, ("nlb" + s).c_str());
        p[n++] = fval(lb, ("thetalb" + s).c_str());
        p[n++] = fval(lb, ("detlb" + s).c_str());
        p[n++] = fval(lb, ("wlb" + s).c_str()) * pc;
        p[n++] = fval(lb, ("hlb" + s).c_str()) * pc;
      }
      p[n++] = fval(li, "nli");
      p[n++] = fval(li, "rli") * pc;
      p[n++] = fval(li, "wli") * pc;
      p[n++] = fval(li, "detthetali");
      p[n++] = fval(li, "thetali");
      if (n != 73)
        throw std::logic_error("ymw16 parameter count");
    } else {
      throw std::runtime_error("unsupported tereg model");
    }
  }
  if (grid_tereg.read_permission || grid_tereg.write_permission || has_box(root, "box_tereg"))
    read_box(root, "box_tereg", grid_tereg);

  // random thermal electrons (param.cc:669-716); "dgm" names the DFT method
  read_fieldio(root, "ternd", grid_ternd);
  grid_ternd.build_permission = trace(te, {"random"})->attr_bool("cue");
  if (grid_ternd.build_permission) {
    ternd_seed = trace(te, {"random"})->attr_uint("seed");
    ternd_type = trace(te, {"random"})->attr_string("type");
    if (ternd_type == "global") {
      ternd_method = trace(te, {"random", "global"})->attr_string("type");
      if (ternd_method == "dft" || ternd_method == "dgm") {
        ternd_method = "dft";
        const Element *s = trace(te, {"random", "global", "dft"});
        ternd_dft.rms = fval(s, "rms");
        ternd_dft.k0 = fval(s, "k0");
        ternd_dft.a0 = fval(s, "a0");
        ternd_dft.r0 = fval(s, "r0") * units::kpc;
        ternd_dft.z0 = fval(s, "z0") * units::kpc;
      } else {
        throw std::runtime_error("unsupported ternd model");
      }
    } else {
      throw std::runtime_error("unsupported ternd type");
    }
  }
  if (grid_ternd.read_permission || grid_ternd.write_permission || grid_ternd.build_permission)
    read_box(root, "box_ternd", grid_ternd);

  // cosmic-ray electrons (param.cc:718-778)
  read_fieldio(root, "cre", grid_cre);
  const Element *cr = trace(root, {"cre"});
  grid_cre.build_permission = cr->attr_bool("cue");
  if (grid_cre.build_permission) {
    cre_type = cr->attr_string("type");
    if (cre_type == "analytic") {
      const Element *s = trace(cr, {"analytic"});
      cre_ana.alpha = fval(s, "alpha");
      cre_ana.beta = fval(s, "beta");
      cre_ana.theta = fval(s, "theta");
      cre_ana.r0 = fval(s, "r0") * units::kpc;
      cre_ana.z0 = fval(s, "z0") * units::kpc;
      cre_ana.E0 = fval(s, "E0") * units::GeV;
      cre_ana.j0 = fval(s, "j0");
    } else if (cre_type == "unif") {
      const Element *s = trace(cr, {"unif"});
      cre_unif.alpha = fval(s, "alpha");
      cre_unif.r0 = fval(s, "r0") * units::kpc;
      cre_unif.E0 = fval(s, "E0") * units::GeV;
      cre_unif.j0 = fval(s, "j0");
    } else {
      throw std::runtime_error("unsupported cre model");
    }
  }
  if (grid_cre.read_permission || grid_cre.write_permission || has_box(root, "box_cre")) {
    const Element *b = trace(root, {"grid", "box_cre"});
    grid_cre.E_min = units::GeV * fval(b, "E_min");
    grid_cre.E_max = units::GeV * fval(b, "E_max");
    grid_cre.nE = uval(b, "nE");
    grid_cre.E_fact = std::log(grid_cre.E_max / grid_cre.E_min) / (grid_cre.nE - 1);
    read_box(root, "box_cre", grid_cre);
    grid_cre.cre_size = grid_cre.nE * grid_cre.full_size;
  }
}

} // namespace hamx_host
