"""CRE_ana in the oracle pinned away from the observer (cre_ana.cc:11-47).

The model: N(gamma) = N0(pos) gamma^idx(pos) with
  profile(pos) = exp((R0 - r) / r0) sech^2(z / z0), r the cylindrical
                 galactocentric radius, R0 the observer's,
  idx(pos)     = -alpha + beta r[kpc] + theta |z|[kpc].
The checks use points whose geometry fixes the answer, so that they are not
the formula typed again:

* radial: at fixed z the emissivity halves at r = R0 + r0 ln 2 and triples
  at r = R0 - r0 ln 3 (constant index, so the emissivity is proportional to
  the profile);
* vertical: at fixed r != R0 it falls to 1/4 at |z| = z0 acosh 2 and to
  1/9 at |z| = z0 acosh 3 (sech^2 is 1/cosh^2);
* cylindrical radius: invariance under rotation about z, and the vertical
  ratios above hold at r = 4, 8.3 and 12 kpc (a spherical radius would add
  a radial fall-off to them);
* the spectral index at a point, two ways: the polarised fraction of the
  power-law synchrotron emissivity, Pi = (p + 1) / (p + 7/3) with p = -idx
  (Rybicki & Lightman 6.38), and the logarithmic slope of the flux
  PHI(E) that CREfield::write_grid writes, between 100 GeV and 1 TeV (the
  sqrt(1 - 1/gamma) velocity factor changes that slope by < 1e-6 there);
* normalisation: PHI(E0) = j0 at the observer up to its own sech^2, and
  j0 / 2 at r = R0 + r0 ln 2 in the plane.

Mistakes planted in the oracle source (a sign, a spherical radius in either
function, sech for sech^2, z0 swapped for r0, a dropped |.|, beta and
theta swapped, a spherical R0) are compiled into copies of the oracle on the
CPU; each copy must fail at least one check, and the unchanged oracle none.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

from hammurabix_amd import _capi as capi
from hammurabix_amd.units import CGS
from oracle_lib import ORACLE_DIR, PROTOS, load_oracle, ptr

KPC = CGS.kpc
OBS = (-8.3 * KPC, 0.0, 0.006 * KPC)
R0 = 8.3  # kpc, the observer's cylindrical radius
ALPHA, BETA, THETA, RS, ZS = 3.0, 0.1, 0.05, 5.0, 1.0  # r0 = 5 kpc, z0 = 1 kpc
E0 = 10 * CGS.GeV
J0 = 0.0217
BPER = 3e-6  # G


def desc(beta=0.0, theta=0.0):
    d = capi.HamxLosDesc()
    d.observer[:] = OBS
    d.cre_model = capi.HAMX_CRE_ANA
    d.cre[0:7] = [ALPHA, beta, theta, RS * KPC, ZS * KPC, E0, J0]
    d.sync_freq = 23e9
    return d


def at(r_kpc, z_kpc, phi=0.0):
    """A point at cylindrical radius r, azimuth phi, height z (kpc)."""
    return np.array([r_kpc * KPC * math.cos(math.pi + phi),
                     r_kpc * KPC * math.sin(math.pi + phi), z_kpc * KPC])


def emissivity(lib, d, pos):
    jt, jp = C.c_double(), C.c_double()
    lib.orc_sync_emissivity(C.byref(d), ptr(np.ascontiguousarray(pos)), BPER,
                            C.byref(jt), C.byref(jp))
    return jt.value, jp.value


def flux(lib, d, pos, E, nE=1, E_fact=0.0):
    """PHI(pos, E .. E e^((nE-1) E_fact)) from CREfield::write_grid on a 2^3 box
    whose lowest vertex is pos (the other vertices lie 1 kpc beyond)."""
    x, y, z = pos
    box = capi.HamxBox(2, 2, 2, x, x + KPC, y, y + KPC, z, z + KPC)
    out = np.zeros(8 * nE)
    assert lib.orc_cre_write_grid(C.byref(d), C.byref(box), E, E_fact, nE, ptr(out)) == 0
    return out.reshape(2, 2, 2, nE)[0, 0, 0]


def index_from_pol(jt, jp):
    # Pi = (1 - idx) / (7/3 - idx)  =>  idx = (7/3 Pi - 1) / (Pi - 1)
    pi = jp / jt
    return (7.0 / 3.0 * pi - 1.0) / (pi - 1.0)


# every check returns a list of failures (empty = pass)

def check_radial(lib):
    bad = []
    d = desc()
    for z in (0.0, 0.4, -1.3):
        ref = emissivity(lib, d, at(R0, z))[0]
        for r, want in ((R0 + RS * math.log(2.0), 0.5), (R0 - RS * math.log(3.0), 3.0),
                        (R0 + RS * math.log(10.0), 0.1)):
            got = emissivity(lib, d, at(r, z))[0] / ref
            if abs(got / want - 1) > 1e-12:
                bad.append(f"radial r={r:.4f} z={z}: {got} != {want}")
    return bad


def check_vertical(lib):
    bad = []
    d = desc()
    for r in (12.0, 4.0, R0):
        ref = emissivity(lib, d, at(r, 0.0))[0]
        for z, want in ((ZS * math.acosh(2.0), 0.25), (-ZS * math.acosh(3.0), 1.0 / 9.0),
                        (ZS * math.acosh(1.5), 1.0 / 2.25)):
            got = emissivity(lib, d, at(r, z))[0] / ref
            if abs(got / want - 1) > 1e-12:
                bad.append(f"vertical r={r} z={z:.4f}: {got} != {want}")
    return bad


def check_rotation(lib):
    bad = []
    for d in (desc(), desc(BETA, THETA)):
        for r, z in ((12.0, 0.7), (4.5, -2.0)):
            ref = emissivity(lib, d, at(r, z))
            for phi in (0.3, 1.9, -2.7):
                got = emissivity(lib, d, at(r, z, phi))
                for g, w in zip(got, ref):
                    if abs(g / w - 1) > 1e-12:
                        bad.append(f"rotation r={r} z={z} phi={phi}: {g} != {w}")
    return bad


POINTS = ((4.0, 0.5), (10.0, -2.0), (15.0, 2.6), (R0, 0.0), (17.0, -0.1))


def check_index_pol(lib):
    bad = []
    d = desc(BETA, THETA)
    for r, z in POINTS:
        want = -ALPHA + BETA * r + THETA * abs(z)
        got = index_from_pol(*emissivity(lib, d, at(r, z)))
        if abs(got - want) > 1e-11:
            bad.append(f"index (polarised fraction) r={r} z={z}: {got} != {want}")
    return bad


def check_index_slope(lib):
    bad = []
    d = desc(BETA, THETA)
    for r, z in POINTS:
        want = -ALPHA + BETA * r + THETA * abs(z)
        phi = flux(lib, d, at(r, z), 100 * CGS.GeV, nE=2, E_fact=math.log(10.0))
        got = math.log(phi[1] / phi[0]) / math.log(10.0)
        if abs(got - want) > 2e-6:
            bad.append(f"index (flux slope) r={r} z={z}: {got} != {want}")
    return bad


def check_norm(lib):
    bad = []
    d = desc()
    got = flux(lib, d, np.array(OBS), E0)[0]
    want = J0 / math.cosh(0.006 / ZS) ** 2
    if abs(got / want - 1) > 1e-13:
        bad.append(f"PHI(E0) at the observer: {got} != {want}")
    got = flux(lib, d, at(R0 + RS * math.log(2.0), 0.0, 0.8), E0)[0]
    if abs(got / (J0 / 2) - 1) > 1e-12:
        bad.append(f"PHI(E0) at r = R0 + r0 ln 2: {got} != {J0 / 2}")
    return bad


CHECKS = (check_radial, check_vertical, check_rotation, check_index_pol, check_index_slope,
          check_norm)


@pytest.mark.parametrize("check", CHECKS, ids=lambda c: c.__name__)
def test_cre_ana_oracle(check):
    bad = check(load_oracle())
    assert not bad, "\n".join(bad)


# ------------------------------------------------------- planted mistakes

MUTANTS = {
    "profile sign": ("std::exp((R0 - r) / p[3])", "std::exp((r - R0) / p[3])"),
    "profile spherical r": ("const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]);",
                            "const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1] + "
                            "pos[2] * pos[2]);"),
    "sech for sech^2": ("(1. / (std::cosh(pos[2] / p[4]) * std::cosh(pos[2] / p[4])))",
                        "(1. / std::cosh(pos[2] / p[4]))"),
    "z0 swapped for r0": ("(1. / (std::cosh(pos[2] / p[4]) * std::cosh(pos[2] / p[4])))",
                          "(1. / (std::cosh(pos[2] / p[3]) * std::cosh(pos[2] / p[3])))"),
    "index spherical r": ("const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1]) / kKpc;",
                          "const double r = std::sqrt(pos[0] * pos[0] + pos[1] * pos[1] + "
                          "pos[2] * pos[2]) / kKpc;"),
    "index without |z|": ("const double z = std::fabs(pos[2]) / kKpc;",
                          "const double z = pos[2] / kKpc;"),
    "beta and theta swapped": ("return -p[0] + p[1] * r + p[2] * z;",
                               "return -p[0] + p[2] * r + p[1] * z;"),
    "spherical R0": ("const double R0 = std::sqrt(obs[0] * obs[0] + obs[1] * obs[1]);",
                     "const double R0 = std::sqrt(obs[0] * obs[0] + obs[1] * obs[1] + "
                     "obs[2] * obs[2]);"),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    """Copies of the oracle, each with one mistake planted in CRE_ana."""
    src = open(os.path.join(ORACLE_DIR, "hamx_oracle.cc")).read()
    # the mistakes go into the CRE_ana restatement only: struct Cre { ... };
    a = src.index("struct Cre {")
    b = src.index("\n};\n", a)
    tmp = tmp_path_factory.mktemp("cre_mutants")
    procs = {}
    for i, (name, (old, new)) in enumerate(MUTANTS.items()):
        block = src[a:b]
        assert block.count(old) == 1, f"mutant {name!r}: the planted text is not unique"
        path = tmp / f"m{i}.cc"
        path.write_text(src[:a] + block.replace(old, new) + src[b:])
        so = tmp / f"m{i}.so"
        procs[name] = (so, subprocess.Popen(
            ["g++", "-O1", "-fopenmp", "-ffp-contract=off", "-fPIC", "-std=c++17",
             "-I", os.path.join(ORACLE_DIR, "..", "include"), "-I", ORACLE_DIR, "-shared",
             "-o", str(so), str(path), os.path.join(ORACLE_DIR, "hamx_oracle_rnd.cc")]))
    libs = {}
    for name, (so, p) in procs.items():
        assert p.wait() == 0, f"mutant {name!r} did not compile"
        lib = C.CDLL(str(so))
        for fn, (res, args) in PROTOS.items():
            f = getattr(lib, fn)
            f.restype, f.argtypes = res, args
        libs[name] = lib
    return libs


@pytest.mark.parametrize("name", list(MUTANTS))
def test_cre_ana_planted_mistakes_fail(mutant_libs, name):
    lib = mutant_libs[name]
    failed = [c.__name__ for c in CHECKS if c(lib)]
    print(f"{name}: fails {failed}")
    assert failed, f"planted mistake {name!r} passes every CRE_ana check"
