"""The oracle's YMW16 thermal electron density (tereg_ymw16.cc:10-333) pinned
without the reference build, component by component: every other density is
set to zero and the component is evaluated where its geometry gives a known
answer.

* Frame: YMW16's (X, Y, Z) is our (y, -x, z); nothing beyond 25 kpc.
* Thick disk: n1 sech^2(Z/h1), the sech^2 radial fall-off beyond bd, the cut
  above Z = 10 h1 only (negative Z is never cut), the warp beyond r_warp
  (Z - gamma_w (R - r_warp) cos(theta)).
* Thin disk: n2 at R = b2, Z = 0, and sech^2(1) at Z = h0(R).
* Spiral arms: n_arm sech^2((R - b2s)/aa) on each arm's ridge
  R = rmin exp((phi - phimin) tan(pitch)); the Local arm (index 4) is never
  reached by the reference's loop; arm 2 carries the Carina-Sagittarius
  factors with the reference's theta * rad expression.
* Galactic centre: ngc at (50, 0, -7) pc, e^-1 one agc away, sech^2(1) one
  hgc away.
* Gum Nebula: ngn on its ellipsoidal shell.
* Local Bubble: nlb1 where l = thetalb1, sech^2 of the distance from its
  110 pc boundary.
* Loop I: nli on its shell along its axis, e^-1 one width or one angular
  width away.
* Combination: thick + max(thin, arms); inside the Local Bubble the thick
  disk scaled by j_lb."""
import math

import numpy as np
import pytest
from scipy import optimize

from hammurabix_amd import Param, configs
from hammurabix_amd.units import CGS
from oracle_lib import ptr

PC, KPC, RAD = CGS.pc, CGS.kpc, CGS.rad
N1, N2, NARM, NGC, NGN, NLB1, NLB2, NLI = 5, 9, 12, 50, 53, 58, 63, 68


def template():
    return list(Param(text=configs.plain_forward(tereg="ymw16")).tereg_ymw16)


def only(*keep, warp=False):
    """The template with every density but those listed set to zero, and
    (unless warp) the warp moved beyond the 25 kpc cut."""
    p = template()
    if not warp:
        p[0] = 100 * KPC
    dens = [N1, N2, NGC, NGN, NLB1, NLB2, NLI] + [NARM + i for i in range(5)]
    for i in dens:
        if i not in keep:
            p[i] = 0.0
    return np.array(p)


def ne(orc, p, X, Y, Z):
    """The density at YMW16 frame point (X, Y, Z) [cm]: our (-Y, X, Z)."""
    return orc.orc_tereg_ymw16(ptr(p), ptr(np.array([-Y, X, Z], dtype=np.float64)))


def sech2(v):
    return 1.0 / math.cosh(v) ** 2


def test_frame_and_outer_cut(orc):
    p = only(N1)
    n1, h1 = p[5], p[6]
    # the frame: YMW16 (X, Y, Z) = our (y, -x, z); a rotation about z keeps
    # the thick disk, so the same density at (2, 3) kpc in both frames' terms
    assert ne(orc, p, 2 * KPC, 3 * KPC, 0.2 * KPC) == pytest.approx(
        n1 * sech2(0.2 * KPC / h1), rel=1e-13, abs=0)
    assert ne(orc, p, 20 * KPC, 0.0, -16 * KPC) == 0.0  # 25.6 kpc
    assert ne(orc, p, 3 * KPC, 0.0, -24 * KPC) > 0.0    # 24.19 kpc, negative Z never cut


def test_thick_disk(orc):
    p = only(N1)
    ad, bd, n1, h1 = p[3], p[4], p[5], p[6]
    assert ne(orc, p, 1 * KPC, 2 * KPC, 0.0) == pytest.approx(n1, rel=1e-15, abs=0)
    assert ne(orc, p, 1 * KPC, 2 * KPC, h1) == pytest.approx(n1 * sech2(1.0), rel=1e-13, abs=0)
    assert ne(orc, p, 1 * KPC, 2 * KPC, 10.01 * h1) == 0.0
    assert ne(orc, p, 1 * KPC, 2 * KPC, -10.01 * h1) == pytest.approx(
        n1 * sech2(10.01), rel=1e-12, abs=0)
    # beyond bd on the Y axis (cos(theta) = 0: no warp)
    R = bd + 0.4 * ad
    assert ne(orc, p, 0.0, R, 0.0) == pytest.approx(n1 * sech2(0.4), rel=1e-12, abs=0)
    # the warp on the X axis (theta = 0): Z' = Z - gamma_w (R - r_warp)
    p = only(N1, warp=True)
    r_warp, gamma_w = p[0], p[2]
    R, Z = 12 * KPC, 0.5 * KPC
    Zw = Z - gamma_w * (R - r_warp)
    assert ne(orc, p, R, 0.0, Z) == pytest.approx(n1 * sech2(Zw / h1), rel=1e-11, abs=0)
    assert abs(Zw) < 0.01 * KPC  # the warp, not the plain height, sets the value


def test_thin_disk(orc):
    p = only(N2)
    a2, b2, n2, k2 = p[7], p[8], p[9], p[10]
    # R = b2 on the Y axis, Z = 0
    assert ne(orc, p, 0.0, b2, 0.0) == pytest.approx(n2, rel=1e-13, abs=0)
    R = 6 * KPC
    h0 = k2 * (32 * PC + 1.6e-3 * R + 4e-7 / PC * R * R)
    assert ne(orc, p, 0.0, R, h0) == pytest.approx(n2 * sech2((R - b2) / a2) * sech2(1.0),
                                                    rel=1e-12, abs=0)
    assert ne(orc, p, 0.0, R, 10.01 * h0) == 0.0


def arm_point(p, i, dphi):
    """A point on arm i's ridge, dphi past its reference angle."""
    rmin, phimin, tpitch = p[22 + i], p[27 + i], p[32 + i]
    R = rmin * math.exp(dphi * tpitch)
    th = phimin + dphi
    return R, R * math.cos(th), R * math.sin(th), th % (2 * math.pi)


@pytest.mark.parametrize("arm", [0, 1, 3])
def test_spiral_arm_ridges(orc, arm):
    p = only(NARM + arm)
    b2s, aa, bd = p[11], p[42], p[4]
    for dphi in (0.3, 1.2, 2.5):
        R, X, Y, _ = arm_point(p, arm, dphi)
        assert R < bd
        want = p[NARM + arm] * sech2((R - b2s) / aa)
        assert ne(orc, p, X, Y, 0.0) == pytest.approx(want, rel=1e-9, abs=0), (arm, dphi)


def test_spiral_arm_two_and_the_local_arm(orc):
    p = only(NARM + 2)
    b2s, aa = p[11], p[42]
    # the second point lies beyond 6 kpc at 145 degrees, where a comparison
    # of theta in degrees with thetacn = 109 would take the (1 + ncn) branch
    for dphi in (1.0, 5.0):
        R, X, Y, th = arm_point(p, 2, dphi)
        # theta * rad (radians times pi/180) against the degree constants
        t = th * RAD
        g = (1 - p[47] * math.exp(-((t - p[48]) / p[49]) ** 2)) * \
            (1 + p[44] * math.exp(-((t - p[45]) / p[46]) ** 2))
        want = p[NARM + 2] * sech2((R - b2s) / aa) * g
        assert ne(orc, p, X, Y, 0.0) == pytest.approx(want, rel=1e-9, abs=0)
        assert abs(g - 1) < 1e-5
    assert R > 6 * KPC and th / RAD > 109
    # the Local arm: its density never enters
    p4 = only(NARM + 4)
    R, X, Y, _ = arm_point(p4, 4, 0.5)
    assert ne(orc, p4, X, Y, 0.0) == 0.0


def test_galactic_centre(orc):
    p = only(NGC)
    ngc, agc, hgc = p[50], p[51], p[52]
    assert ne(orc, p, 50 * PC, 0.0, -7 * PC) == pytest.approx(ngc, rel=1e-15, abs=0)
    assert ne(orc, p, 50 * PC + agc, 0.0, -7 * PC) == pytest.approx(ngc / math.e, rel=1e-13,
                                                                    abs=0)
    assert ne(orc, p, 50 * PC, 0.0, -7 * PC + hgc) == pytest.approx(ngc * sech2(1.0),
                                                                    rel=1e-13, abs=0)


def test_gum_nebula_shell(orc):
    p = only(NGN)
    ngn, a, k, r0 = p[53], p[55], p[56], p[1]
    lc, bc, dc = 264 * RAD, -4 * RAD, 450 * PC
    xc, yc, zc = dc * math.cos(bc) * math.sin(lc), r0 - dc * math.cos(bc) * math.cos(lc), \
        dc * math.sin(bc)
    for th, ph in ((0.5, 0.8), (1.1, 2.0), (0.2, 4.0)):
        # where the ray from the centre at elevation th meets the ellipsoid
        # (rho/a)^2 + (z/(a k))^2 = 1
        s = 1.0 / math.sqrt((math.cos(th) / a) ** 2 + (math.sin(th) / (a * k)) ** 2)
        X, Y, Z = xc + s * math.cos(th) * math.cos(ph), yc + s * math.cos(th) * math.sin(ph), \
            zc + s * math.sin(th)
        assert ne(orc, p, X, Y, Z) == pytest.approx(ngn, rel=1e-9, abs=0), (th, ph)
    assert ne(orc, p, xc + 3 * a, yc, zc + 0.3 * a) == 0.0


def test_local_bubble(orc):
    p = only(NLB1)
    nlb1, thetalb1, wlb1 = p[58], p[59], p[61]
    r0 = p[1]
    l = thetalb1 * RAD
    d_edge = 5 * PC

    def r_lb(s):
        X, Y = s * math.sin(l), r0 - s * math.cos(l)
        return math.hypot((Y - 8.34 * KPC) * 0.94, X) - (110 * PC + d_edge)
    s = optimize.brentq(r_lb, 1 * PC, 2 * KPC, xtol=1e-10 * PC)
    X, Y = s * math.sin(l), r0 - s * math.cos(l)
    assert ne(orc, p, X, Y, 0.0) == pytest.approx(nlb1 * sech2(d_edge / wlb1), rel=1e-9, abs=0)


def test_loop_one(orc):
    p = only(NLI)
    nli, rli, wli, det, th_li = p[68], p[69], p[70], p[71], p[72] * RAD
    c = np.array([-10.156 * PC, 8106.207 * PC, 10.467 * PC])
    axis = np.array([math.cos(th_li), 0.0, math.sin(th_li)])
    X, Y, Z = c + rli * axis
    assert ne(orc, p, X, Y, Z) == pytest.approx(nli, rel=1e-9, abs=0)
    X, Y, Z = c + (rli + wli) * axis
    assert ne(orc, p, X, Y, Z) == pytest.approx(nli / math.e, rel=1e-9, abs=0)
    # det degrees off the axis, in the plane of the axis and y
    ang = th_li + det * RAD
    X, Y, Z = c + rli * np.array([math.cos(ang), 0.0, math.sin(ang)])
    assert ne(orc, p, X, Y, Z) == pytest.approx(nli / math.e, rel=1e-9, abs=0)


def test_combination_rules(orc):
    # inside the Local Bubble the thick disk is scaled by j_lb
    p = only(N1)
    j_lb, n1 = p[57], p[5]
    assert ne(orc, p, 0.0, 8.34 * KPC, 0.0) == pytest.approx(j_lb * n1, rel=1e-15, abs=0)
    # elsewhere thick + max(thin, arms)
    p = only(N1, N2, NARM + 0)
    R, X, Y, _ = arm_point(p, 0, 1.2)
    thick = ne(orc, only(N1), X, Y, 0.0)
    thin = ne(orc, only(N2), X, Y, 0.0)
    arm = ne(orc, only(NARM + 0), X, Y, 0.0)
    assert thin > 0 and arm > 0 and thin != arm
    assert ne(orc, p, X, Y, 0.0) == pytest.approx(thick + max(thin, arm), rel=1e-14, abs=0)


def test_template_sky_is_finite_and_positive(orc):
    """Every sample of a sight line through the template model is finite and
    non-negative, and the Galactic plane is denser than high latitudes."""
    p = np.array(template())
    sun = np.array([-8.3 * KPC, 0.0, 0.006 * KPC])
    rng = np.random.default_rng(0)
    for _ in range(400):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        for s in np.linspace(0.01, 20, 7) * KPC:
            v = orc.orc_tereg_ymw16(ptr(p), ptr(sun + s * d))
            assert math.isfinite(v) and v >= 0
    plane = orc.orc_tereg_ymw16(ptr(p), ptr(np.array([-5 * KPC, 0.0, 0.0])))
    high = orc.orc_tereg_ymw16(ptr(p), ptr(np.array([-5 * KPC, 0.0, 3 * KPC])))
    assert plane > 10 * high > 0
