"""The LOS adjoint over the forward's other field models and over masks.

The forward integrates any regular B model (LSA, cart, unif, helix, JF12,
Jaffe, an imported grid) and any TE model (uniform sphere, YMW16, an
imported grid) with the random grids added, and a mask restricts it to the
pixels the mask keeps (integrator.cc:63-70).  hamx_los_adjoint covers the
same set (the CRE stays analytic or uniform with a constant index):

* the LSA parameters (b0, psi0, psi1, chi0) enter the forward only through
  the LSA field, so their gradient is exactly zero for the other B models,
  and n0 enters only through the uniform TE sphere;
* alpha (the CRE index), n0 where it applies, and the random-B and random-TE
  grid gradients along random directions equal central differences of the
  CPU oracle's maps (h = 1e-5 relative; rel 1e-5);
* a masked-out ray contributes nothing: the masked adjoint equals the
  oracle's central differences of the masked forward.
"""
import os

import numpy as np
import pytest

from hammurabix_amd import HamxError, Param, _capi as capi, configs
from hammurabix_amd.units import CGS
from oracle_lib import write_grid
from test_gpu_adjoint import adjoint, random_grids, scaled_weights

pytestmark = pytest.mark.gpu

RTOL = 1e-5
H_REL = 1e-5
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def case(breg, tereg, mask=False, seed=21):
    from test_oracle_models import jf12_params
    # shells at nside 4/8/16 when masked: the nside-8 mask is downgraded,
    # used as is and upgraded
    par = Param(text=configs.random_forward(
        nside_sim=4, nside_out=16 if mask else 8, oc_r_res=0.1, grid=(16, 16, 8),
        **({"tereg": "ymw16"} if tereg == "ymw16" else {})))
    d = par.los_desc()
    if breg == "jf12":
        d.breg_model = capi.HAMX_BREG_JF12
        d.breg[0:20] = list(jf12_params())
    elif breg == "helix":
        d.breg_model = capi.HAMX_BREG_HELIX
        d.breg[0:5] = [1e-6, 2e-6, 0.5e-6, 3 * CGS.kpc, 12 * CGS.kpc]
    elif breg == "cart":
        d.breg_model = capi.HAMX_BREG_CART
        d.breg[0:3] = [1e-6, -2e-6, 0.5e-6]
    b, te, box = random_grids((16, 16, 8), seed, te_scale=0.0003 if tereg == "ymw16" else 0.003)
    grids = dict(brnd=b, brnd_box=box, ternd=te, ternd_box=box)
    if mask:
        grids["mask"] = np.ascontiguousarray(
            np.fromfile(os.path.join(GOLDEN, "mask_map_nside8.bin"), dtype="<f8"))
        grids["nside_mask"] = 8
    return d, grids


def oracle_loss(d, w, grids):
    m = write_grid(d, **grids)
    return sum(float(np.dot(w[k], m[k])) for k in w)


CASES = [("jf12", "unif", False), ("helix", "ymw16", False), ("cart", "unif", False),
         ("lsa", "ymw16", False), ("lsa", "unif", True), ("jf12", "ymw16", True)]


@pytest.mark.parametrize("breg,tereg,mask", CASES)
def test_adjoint_of_other_models_vs_oracle(ctx, breg, tereg, mask):
    print(f"case: adjoint breg {breg} tereg {tereg} mask {mask}", flush=True)
    d, grids = case(breg, tereg, mask)
    w = scaled_weights(ctx, d, grids, seed=22)
    dp, g = adjoint(ctx, d, w, grids, with_grid=True)
    if breg != "lsa":
        assert np.all(dp[0:4] == 0.0)
    if tereg != "unif":
        assert dp[5] == 0.0
    # alpha, and n0 with the uniform sphere, against the oracle
    checks = [("cre", 0, 4)] + ([("tereg", 0, 5)] if tereg == "unif" else []) + \
        ([("breg", j, j) for j in range(4)] if breg == "lsa" else [])
    for arr, j, i in checks:
        p = getattr(d, arr)
        p0 = p[j]
        h = H_REL * abs(p0)
        p[j] = p0 + h
        lp = oracle_loss(d, w, grids)
        p[j] = p0 - h
        lm = oracle_loss(d, w, grids)
        p[j] = p0
        fd = (lp - lm) / (2 * h)
        print(f"{breg}/{tereg}/{mask} d/d{capi.PARAM_NAMES[i]}: adjoint {dp[i]:.10e} "
              f"oracle fd {fd:.10e}", flush=True)
        assert abs(dp[i] - fd) <= RTOL * abs(fd), capi.PARAM_NAMES[i]
    # grid gradients along a random direction (the TE one on a positive grid
    # so that no sample sits at the te > 0 clamp)
    rng = np.random.default_rng(23)
    b = grids["brnd"]
    v = [rng.normal(size=a.shape) for a in b]
    h = 1e-5 * 2e-6
    plus = dict(grids, brnd=tuple(np.ascontiguousarray(a + h * va) for a, va in zip(b, v)))
    minus = dict(grids, brnd=tuple(np.ascontiguousarray(a - h * va) for a, va in zip(b, v)))
    fd = (oracle_loss(d, w, plus) - oracle_loss(d, w, minus)) / (2 * h)
    got = sum(float(np.dot(ga, va)) for ga, va in zip(g["dbrnd"], v))
    print(f"brnd directional: adjoint {got:.10e} oracle fd {fd:.10e}", flush=True)
    assert abs(got - fd) <= RTOL * abs(fd)
    te_pos = np.ascontiguousarray(np.abs(grids["ternd"]) + 1e-4)
    gp = dict(grids, ternd=te_pos)
    wp = scaled_weights(ctx, d, gp, seed=24)
    _, g_pos = adjoint(ctx, d, wp, gp, with_grid=True)
    vt = rng.normal(size=te_pos.shape)
    h = 1e-5 * 1e-3
    fd = (oracle_loss(d, wp, dict(gp, ternd=np.ascontiguousarray(te_pos + h * vt))) -
          oracle_loss(d, wp, dict(gp, ternd=np.ascontiguousarray(te_pos - h * vt)))) / (2 * h)
    got = float(np.dot(g_pos["dternd"], vt))
    print(f"ternd directional: adjoint {got:.10e} oracle fd {fd:.10e}", flush=True)
    assert abs(got - fd) <= RTOL * abs(fd)


def test_mask_changes_the_gradient(ctx):
    """The masked adjoint is not the unmasked one: the masked-out rays take
    their share of the parameter and grid gradients with them."""
    print("case: adjoint masked vs unmasked", flush=True)
    d, grids = case("lsa", "unif", mask=True, seed=25)
    w = scaled_weights(ctx, d, grids, seed=26)
    dp_m, g_m = adjoint(ctx, d, w, grids, with_grid=True)
    full = {k: v for k, v in grids.items() if k not in ("mask", "nside_mask")}
    dp_f, g_f = adjoint(ctx, d, w, full, with_grid=True)
    assert np.linalg.norm(g_m["dbrnd"][0]) < np.linalg.norm(g_f["dbrnd"][0])
    assert not np.allclose(dp_m, dp_f, rtol=1e-3)


def test_adjoint_still_rejects_grid_cre_and_variable_index(ctx):
    d = Param(text=configs.plain_forward(nside_sim=4, nside_out=4)).los_desc()
    w = {k: np.ones(12 * 16) for k in ("is", "qs", "us")}
    d.cre[1] = 0.1 / CGS.kpc  # beta != 0: the CRE index varies along the ray
    with pytest.raises(HamxError) as e:
        ctx.los_adjoint(d, w, np.zeros(capi.HAMX_NPARAM))
    assert e.value.code == capi.HAMX_ERR_UNSUPPORTED
