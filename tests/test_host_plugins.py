"""User field models written against the reference's plugin headers
(bfield.h:44-97, tefield.h:40-81, crefield.h:24-84, pipeline.h:14-45):
tests/user_plugins.cc derives Breg, TEreg and CREfield models and a Pipeline
subclass that installs them, and compiles unchanged against this build's
host library with hammurabix_amd/host/include first on the include path.

* CPU: the program builds; its host-side evaluations hold -- Hamvec algebra,
  the base-class read_field dispatch (breg.cc:11-21, tereg.cc:11-20), the
  library's CRE_ana / CRE_unif write_field against the oracle's
  CREfield::write_grid at the same point, the base class's "wrong
  inheritance" throw (cre.cc:28-44), and CREfield::read_grid_num exact on a
  field linear in position and energy index.
* GPU: the pipeline with the user models, which the integrator bakes onto
  box_breg / box_tereg and writes onto box_cre (as imported grids, CRE_num),
  equals the oracle's write_grid passes on the grids the program exported
  through the reference's write_grid / export_grid calls, to 1e-10.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

from hammurabix_amd import Param, _capi as capi, configs
from hammurabix_amd.units import CGS
from oracle_lib import fields_struct, load_oracle, ptr, rel_l2, write_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "hammurabix_amd")
SRC = os.path.join(ROOT, "tests", "user_plugins.cc")
KPC = CGS.kpc


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libhamx_host.so")):
        pytest.skip("host library not built (python -m hammurabix_amd._build)")
    exe = tmp_path_factory.mktemp("plugins") / "user_plugins"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror",
                    "-I", os.path.join(LIBDIR, "host", "include"),
                    "-I", os.path.join(LIBDIR, "host"), "-I", os.path.join(ROOT, "include"),
                    SRC, "-o", str(exe), "-L", LIBDIR, "-lhamx_host", "-lhamx",
                    f"-Wl,-rpath,{LIBDIR}"], check=True)
    return str(exe)


def host_check(program, xml_text, tmp_path, name):
    x = tmp_path / name
    x.write_text(xml_text)
    r = subprocess.run([program, "--host-check", str(x)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        key, *vals = line.split()
        out.setdefault(key, []).append([float(v) for v in vals])
    return out


def oracle_flux(desc, pos, E):
    """The oracle's CREfield::write_grid value at one position and energy."""
    orc = load_oracle()
    x, y, z = pos
    box = capi.HamxBox(2, 2, 2, x, x + KPC, y, y + KPC, z, z + KPC)
    out = np.zeros(8)
    assert orc.orc_cre_write_grid(C.byref(desc), C.byref(box), E, 0.0, 1, ptr(out)) == 0
    return out[0]


def user_breg(p):
    a = 0.3 * p[0] / KPC + 0.1 * p[1] / KPC
    return np.array([3e-6 * math.cos(a), 3e-6 * math.sin(a), 0.5e-6 * math.tanh(p[2] / KPC)])


def test_user_models_compile_and_evaluate_on_the_host(program, tmp_path):
    p = np.array([-5 * KPC, 2 * KPC, 0.4 * KPC])
    out = host_check(program, configs.plain_forward(nside_sim=4, nside_out=8,
                                                    cre_grid=(10, 10, 6, 8), cre_write=True,
                                                    cre_file=str(tmp_path / "c.bin")),
                     tmp_path, "ana.xml")
    cx, cy, cz, dot, dist, vz = out["hamvec"][0]
    assert (cx, cy, cz, dot) == (6.5, -10.0, 4.5, 11.0)
    assert dist == 3.5 and vz == pytest.approx(3 / math.sqrt(14), rel=1e-15)
    b = out["breg_user"][0]
    assert b[:3] == b[3:]  # read_field dispatches to write_field when built
    np.testing.assert_allclose(b[:3], user_breg(p), rtol=1e-14)
    t = out["tereg_user"][0]
    assert t[0] == t[1] == pytest.approx(0.02 * math.exp(-0.4), rel=1e-14)
    # the library's CRE_ana on the host equals the oracle's
    par = Param(text=configs.plain_forward())
    d = par.los_desc()
    for E, phi, idx, prof, norm in out["cre_ana"]:
        assert phi == pytest.approx(oracle_flux(d, p, E), rel=1e-13)
        assert idx == -3.0
        r = math.hypot(p[0], p[1])
        assert prof == pytest.approx(math.exp((8.3 * KPC - r) / (5 * KPC)) /
                                     math.cosh(0.4) ** 2, rel=1e-14)
        assert norm > 0
    assert out["base_throws"][0] == [1.0]
    for m, v in out["read_grid_num"]:
        assert v == pytest.approx(1 - 10 - 1 + 1.2 + 0.25 * m, rel=1e-13)
    # the uniform CRE model on a parameter file that selects it
    out = host_check(program, configs.plain_forward(cre_type="unif", cre_write=True,
                                                    cre_grid=(10, 10, 6, 8),
                                                    cre_file=str(tmp_path / "c2.bin")),
                     tmp_path, "unif.xml")
    du = Param(text=configs.plain_forward(cre_type="unif")).los_desc()
    phi, idx, prof = out["cre_unif"][0]
    obs = np.array(du.observer[:])
    assert phi == pytest.approx(oracle_flux(du, obs, 7 * CGS.GeV), rel=1e-13)
    # p lies 3.9 kpc from the observer, inside the 5 kpc sphere
    assert idx == -3.0 and prof == 1.0


@pytest.mark.gpu
def test_user_models_run_through_the_pipeline(program, tmp_path):
    """The user models on the device (baked regular fields, the CRE written
    on box_cre and integrated as CRE_num) against the oracle's passes on the
    grids the program exported, for two frequencies in the reference
    pipeline's order."""
    print("case: user plugin models through the pipeline", flush=True)
    out = str(tmp_path) + "/"
    files = dict(breg_file=str(tmp_path / "breg.bin"), tereg_file=str(tmp_path / "tereg.bin"),
                 cre_file=str(tmp_path / "cre.bin"))
    ghz = [23.0, 1.4]
    text = configs.plain_forward(nside_sim=4, nside_out=8, oc_r_res=0.05, freqs=ghz,
                                 out=out, cre_write=True, cre_grid=(40, 40, 10, 24),
                                 regular_grid=(60, 60, 20), **files)
    x = tmp_path / "p.xml"
    x.write_text(text)
    r = subprocess.run([program, str(x)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    par = Param(text=text)
    gb, gt, gc = par.grid_breg, par.grid_tereg, par.grid_cre
    raw = np.fromfile(files["breg_file"], dtype="<f8").reshape(-1, 3)
    breg = tuple(np.ascontiguousarray(raw[:, i]) for i in range(3))
    tereg = np.fromfile(files["tereg_file"], dtype="<f8")
    cre = np.fromfile(files["cre_file"], dtype="<f8")
    assert breg[0].size == gb.nx * gb.ny * gb.nz and cre.size == gc.full_size * gc.nE
    # the exported B grid is the user model at the box vertices
    xs = np.linspace(gb.x_min, gb.x_max, gb.nx)
    ys = np.linspace(gb.y_min, gb.y_max, gb.ny)
    zs = np.linspace(gb.z_min, gb.z_max, gb.nz)
    for i, j, k in ((3, 7, 2), (59, 0, 19), (30, 31, 10)):
        v = (i * gb.ny + j) * gb.nz + k
        np.testing.assert_allclose([a[v] for a in breg], user_breg((xs[i], ys[j], zs[k])),
                                   rtol=1e-13, atol=1e-22)
    fields = fields_struct(breg=breg, breg_box=gb.box(), tereg=tereg, tereg_box=gt.box(),
                           cre=cre, cre_box=gc.box(), cre_E=(gc.E_min, gc.E_fact, gc.nE))
    freqs = par.grid_obs.sim_sync_freq
    for p, fi in enumerate(reversed(range(len(freqs)))):
        d = par.los_desc()
        d.breg_model = capi.HAMX_BREG_GRID
        d.tereg_model = capi.HAMX_TEREG_GRID
        d.cre_model = capi.HAMX_CRE_GRID
        d.sync_freq = freqs[fi]
        if p > 0:
            d.do_dm = d.do_fd = 0
        ref = write_grid(d, fields=fields)
        for c, k in (("I", "is"), ("Q", "qs"), ("U", "us")):
            got = np.fromfile(f"{out}sync_{ghz[fi]}_{c}.bin", dtype="<f8")
            err = rel_l2(got, ref[k])
            print(f"user models {k}@{freqs[fi] / 1e9:g} GHz: rel L2 vs oracle {err:.3e}",
                  flush=True)
            assert err <= 1e-10
        if p == 0:
            dm = np.fromfile(f"{out}dm.bin", dtype="<f8")
            fd = np.fromfile(f"{out}fd.bin", dtype="<f8")
            assert rel_l2(dm, ref["dm"] * CGS.ccm / CGS.pc) <= 1e-10
            assert rel_l2(fd, ref["fd"] * CGS.m ** 2) <= 1e-10
