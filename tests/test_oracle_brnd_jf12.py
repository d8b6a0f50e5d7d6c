"""The oracle's Brnd_jf12 (brnd_jf12.cc:54-385) pinned without the reference
build.

* The energy profile (brnd_jf12.cc:54-114) at points whose answer the model's
  geometry fixes by hand: zero inside the 1 kpc sphere and beyond r = 20 kpc,
  the inner disk b0_int, the eight spiral regions read off where the
  logarithmic spiral of pitch 11.5 degrees through the point crosses the
  negative x-axis (on the axis itself, a quarter and half a turn along the
  spiral, and the two wraps the reference takes when the first crossing lies
  beyond 15.5 kpc), the 5 kpc / r fall-off, the Gaussian disk height and the
  exponential halo.
* The generator: with a halo-only profile tuned to equal four times the
  Brnd_es profile on a box that avoids both cut-offs, Brnd_jf12 must give
  twice Brnd_es's field times the fixed factors its literal constants imply: the 0.33333333 of the
  spectrum cancels in the variance normalisation (brnd_jf12.cc:209-213), the
  anisotropy norm rhonorm(0.33333333, 0.66666667) / rhonorm(1/3, 2/3)
  scales every mode, and the Gram-Schmidt factor 1.73205081 / sqrt 3 every
  mode but k = 0, which the projection passes through unscaled
  (brnd_jf12.cc:122-134), so the grid mean and the fluctuation about it
  carry different factors."""
import ctypes as C
import math

import numpy as np
import pytest

from hammurabix_amd import Param, _capi as capi, configs
from hammurabix_amd.param import BRND_JF12_PARAMS
from hammurabix_amd.units import CGS
from oracle_lib import ptr, rel_l2

KPC = CGS.kpc
MU = 1e-6
TANP = math.tan(math.radians(11.5))
RC = (5.1, 6.3, 7.1, 8.3, 9.8, 11.4, 12.7, 15.5)


def template(**kw):
    """The template's <jf12> random block (templates/params_template.xml) in
    hamx_es_desc.jf12 order and CGS units."""
    v = dict(b0_1=10.81 * MU, b0_2=6.96 * MU, b0_3=9.59 * MU, b0_4=6.96 * MU,
             b0_5=1.96 * MU, b0_6=16.34 * MU, b0_7=37.29 * MU, b0_8=10.35 * MU,
             b0_int=7.63 * MU, z0_spiral=0.61 * KPC, b0_halo=4.68 * MU,
             r0_halo=10.97 * KPC, z0_halo=2.84 * KPC)
    v.update(kw)
    return np.array([v[k] for k, _ in BRND_JF12_PARAMS])


def prof(orc, q, pos_kpc):
    pos = np.asarray(pos_kpc, dtype=np.float64) * KPC
    return orc.orc_jf12_profile(ptr(q), ptr(pos))


def region(r_cross):
    """Index of the spiral region whose outer edge first exceeds the crossing
    radius on the negative x-axis (rc_B of brnd_jf12.cc:84-85)."""
    return min(i for i, rc in enumerate(RC) if r_cross < rc)


def test_profile_cutoffs_inner_disk_and_halo(orc):
    q = template()
    assert prof(orc, q, [0.5, 0.3, 0.2]) == 0.0
    assert prof(orc, q, [0.0, 0.0, 0.0]) == 0.0
    assert prof(orc, q, [20.5, 0.0, 0.0]) == 0.0
    assert prof(orc, q, [15.0, 14.0, 0.0]) == 0.0
    # inner disk: b0_int with the Gaussian height, plus the halo
    x, y, z = 3.0, 2.0, 0.4
    r = math.hypot(x, y)
    disk = 7.63 * math.exp(-0.5 * (z / 0.61) ** 2)
    halo = 4.68 * math.exp(-r / 10.97) * math.exp(-z / 2.84)
    assert prof(orc, q, [x, y, z]) == pytest.approx(disk ** 2 + halo ** 2, rel=1e-13, abs=0)
    # halo alone, both signs of z
    qh = template(**{f"b0_{i}": 0.0 for i in range(1, 9)}, b0_int=0.0)
    for pos in ([3.0, 4.0, -1.5], [-9.0, 6.0, 2.5], [0.2, 0.1, 1.2]):
        r = math.hypot(pos[0], pos[1])
        want = (4.68 * math.exp(-r / 10.97) * math.exp(-abs(pos[2]) / 2.84)) ** 2
        assert prof(orc, qh, pos) == pytest.approx(want, rel=1e-13, abs=0)


ARMS = (10.81, 6.96, 9.59, 6.96, 1.96, 16.34, 37.29, 10.35)


def _disk_only():
    return template(b0_halo=0.0)


@pytest.mark.parametrize("rx,want", [(5.05, 0), (6.0, 1), (7.0, 2), (8.0, 3), (9.0, 4),
                                     (11.0, 5), (12.0, 6), (15.0, 7)])
def test_profile_regions_on_the_negative_x_axis(orc, rx, want):
    assert region(rx) == want
    got = prof(orc, _disk_only(), [-rx, 0.0, 0.0])
    assert got == pytest.approx((ARMS[want] * 5.0 / rx) ** 2, rel=1e-13, abs=0)
    z = 0.35
    got = prof(orc, _disk_only(), [-rx, 0.0, z])
    assert got == pytest.approx((ARMS[want] * 5.0 / rx * math.exp(-0.5 * (z / 0.61) ** 2)) ** 2,
                                rel=1e-13, abs=0)


@pytest.mark.parametrize("rx,dphi", [(12.0, math.pi / 2), (12.0, math.pi), (15.0, 1.5 * math.pi),
                                     (9.5, math.pi / 3)])
def test_profile_constant_region_along_the_spiral(orc, rx, dphi):
    """A point reached from (-rx, 0, 0) by turning dphi clockwise along the
    spiral r = rx exp(tan(11.5 deg) (phi - pi)) keeps that crossing's region
    and scales as 5 kpc / r."""
    phi = math.pi - dphi
    r = rx * math.exp(-TANP * dphi)
    assert r > 5.0
    got = prof(orc, _disk_only(), [r * math.cos(phi), r * math.sin(phi), 0.0])
    assert got == pytest.approx((ARMS[region(rx)] * 5.0 / r) ** 2, rel=1e-10, abs=0)


def test_profile_wraps_to_inner_crossings(orc):
    """When the crossing half a turn outward lies beyond 15.5 kpc, the region
    comes from the crossing reached by following the spiral inward
    (decreasing phi) to phi = -pi, and failing that, one more turn in."""
    q = _disk_only()
    # (0, -10): outward crossing at 10 e^{3 pi tan/2} > 15.5; inward at phi = -pi
    r_in = 10.0 * math.exp(-TANP * math.pi / 2)
    assert 10.0 * math.exp(TANP * 1.5 * math.pi) > 15.5
    assert prof(orc, q, [0.0, -10.0, 0.0]) == pytest.approx(
        (ARMS[region(r_in)] * 5.0 / 10.0) ** 2, rel=1e-13, abs=0)
    assert region(r_in) == 3
    # (-16, 0): on the axis but beyond 15.5 kpc; one turn inward
    r_in = 16.0 * math.exp(-TANP * 2 * math.pi)
    assert prof(orc, q, [-16.0, 0.0, 0.0]) == pytest.approx(
        (ARMS[region(r_in)] * 5.0 / 16.0) ** 2, rel=1e-13, abs=0)
    assert region(r_in) == 0
    # just below the negative x-axis (phi ~ -pi) at r = 17: the crossings at
    # phi = pi and -pi are both beyond 15.5 kpc, the one a turn further in is not
    phi = -math.pi + 1e-3
    r = 17.0
    r_in = r * math.exp(-TANP * (phi + 3 * math.pi))
    assert r * math.exp(-TANP * (phi + math.pi)) > 15.5
    got = prof(orc, q, [r * math.cos(phi), r * math.sin(phi), 0.0])
    assert got == pytest.approx((ARMS[region(r_in)] * 5.0 / r) ** 2, rel=1e-12, abs=0)


def _es_desc(shape, rho):
    d = capi.HamxEsDesc()
    d.box = capi.HamxBox(*shape, 2 * KPC, 14 * KPC, -6 * KPC, 6 * KPC, -2 * KPC, 2 * KPC)
    d.observer[:] = [-8.3 * KPC, 0.0, 0.006 * KPC]
    d.rms, d.k0, d.k1, d.a0, d.a1, d.rho = 2e-6, 0.5, 0.5, 1.7, 0.0, rho
    d.breg_model = capi.HAMX_BREG_CART
    d.breg[0:3] = [1e-6, 2e-6, 0.5e-6]
    return d


@pytest.mark.parametrize("rho", [1.0, 0.5])
def test_generator_is_es_with_its_constants(orc, rho):
    shape = (16, 16, 8)
    r0h, z0h = 10.97 * KPC, 2.84 * KPC
    d = _es_desc(shape, rho)
    # ES profile exp(-(r - 8.3)/r0) exp(-(|z| - 0.006)/z0) with r0 = r0h/2,
    # z0 = z0h/2 is a quarter of the halo-only JF12 profile with this b0_halo
    d.r0, d.z0 = r0h / 2, z0h / 2
    b0h = 2 * MU * math.exp(8.3 * KPC / r0h + 0.006 * KPC / z0h)
    d.jf12[:] = template(**{f"b0_{i}": 0.0 for i in range(1, 9)}, b0_int=0.0, b0_halo=b0h)
    n = int(np.prod(shape))
    es = [np.zeros(n) for _ in range(3)]
    jf = [np.zeros(n) for _ in range(3)]
    assert orc.orc_brnd_es(C.byref(d), 5, *(ptr(a) for a in es)) == 0
    assert orc.orc_brnd_jf12(C.byref(d), 5, *(ptr(a) for a in jf)) == 0
    rhn_j = 1 / math.sqrt(0.33333333 * rho ** 2 + 0.66666667 / rho ** 2)
    rhn_e = 1 / math.sqrt(rho ** 2 / 3 + 2 / (3 * rho ** 2))
    f_mean = 2 * rhn_j / rhn_e
    factor = 1.73205081 / math.sqrt(3) * f_mean
    assert abs(factor / 2 - 1) > 5e-10
    for a, b in zip(jf, es):
        fluct = b - b.mean()
        rms = np.linalg.norm(fluct) / math.sqrt(n)
        assert rms > 0 and abs(b.mean()) > 1e-5 * rms
        assert rel_l2(a - a.mean(), fluct * factor) < 1e-13
        # the mean differs from the fluctuation's factor by 1.4e-9
        assert a.mean() == pytest.approx(b.mean() * f_mean, rel=1e-10, abs=0)
        assert rel_l2(a, 2 * b) > 0.5 * abs(factor / 2 - 1)


def test_param_reads_the_jf12_random_block():
    text = configs.random_forward().replace('<global type="es">', '<global type="jf12">')
    block = text[text.index("<es>"):text.index("</es>") + 5]
    jf = ("<jf12><rms value=\"1.5\"/><k0 value=\"10.0\"/><k1 value=\"0.1\"/>"
          "<a0 value=\"1.7\"/><a1 value=\"0.0\"/><rho value=\"0.8\"/>"
          + "".join(f'<b0_{i} value="{v}"/>' for i, v in enumerate(ARMS, 1))
          + '<b0_int value="7.63"/><z0_spiral value="0.61"/><b0_halo value="4.68"/>'
          '<r0_halo value="10.97"/><z0_halo value="2.84"/></jf12>')
    p = Param(text=text.replace(block, jf))
    assert p.brnd_method == "jf12"
    d = p.es_desc()
    assert d.rms == pytest.approx(1.5 * MU, rel=1e-15, abs=0) and d.k0 == 10.0 and d.rho == 0.8
    np.testing.assert_allclose(np.array(d.jf12), template(), rtol=1e-15)
