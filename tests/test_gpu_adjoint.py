"""GPU checks of the LOS adjoint (hamx_los_adjoint) against central finite
differences of the GPU forward: the six model parameters, single grid
vertices and random directions over the whole random-B and random-TE
grids, with FD carried across shells.  L = sum over maps of w * map with
random weights w, so dL/d(map) = w.

Finite-difference error: with step h relative to the parameter, the
truncation term is O(h^2) and the rounding term O(eps |L| / h); h = 1e-5
keeps both below 1e-7 of the gradient on these workloads, and the check is
rel 1e-6 (plus an absolute floor of 1e-9 of the gradient's scale for
components that nearly vanish).
"""
import numpy as np
import pytest

from hammurabix_amd import HamxError, Param, _capi as capi, configs
from hammurabix_amd.units import CGS
from oracle_lib import npix, rel_l2

pytestmark = pytest.mark.gpu

RTOL = 1e-6
H_REL = 1e-5


def out_maps(d):
    m = {}
    if d.do_dm:
        m["dm"] = np.zeros(npix(d.nside_dm))
    for k in ("is", "qs", "us"):
        m[k] = np.zeros(npix(d.nside_sync))
    if d.do_fd:
        m["fd"] = np.zeros(npix(d.nside_fd))
    return m


def weights(d, seed):
    rng = np.random.default_rng(seed)
    w = {}
    for k, v in out_maps(d).items():
        w[k] = np.ascontiguousarray(rng.normal(size=v.shape))
    return w


def loss(ctx, d, w, grids):
    m = out_maps(d)
    ctx.los_forward(d, m, **grids)
    return sum(float(np.dot(w[k], m[k])) for k in m)


def scaled_weights(ctx, d, grids, seed):
    """Random weights divided by each map's rms, so that every map
    contributes to L at the same order."""
    m = out_maps(d)
    ctx.los_forward(d, m, **grids)
    w = weights(d, seed)
    for k in w:
        s = float(np.sqrt(np.mean(m[k] ** 2)))
        w[k] = w[k] / s if s > 0 else w[k]
    return w


def random_grids(shape, seed, te_scale=0.003):
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    b = tuple(np.ascontiguousarray(rng.normal(0, 2e-6, n)) for _ in range(3))
    te = np.ascontiguousarray(rng.normal(0.0, te_scale, n))
    box = capi.HamxBox(*shape, -20 * CGS.kpc, 20 * CGS.kpc, -20 * CGS.kpc, 20 * CGS.kpc,
                       -4 * CGS.kpc, 4 * CGS.kpc)
    return b, te, box


PARAMS = [("breg", 0), ("breg", 1), ("breg", 2), ("breg", 3), ("cre", 0), ("tereg", 0)]


def get_param(d, i):
    arr, j = PARAMS[i]
    return getattr(d, arr)[j]


def set_param(d, i, v):
    arr, j = PARAMS[i]
    getattr(d, arr)[j] = v


def adjoint(ctx, d, w, grids, with_grid):
    dp = np.zeros(capi.HAMX_NPARAM)
    kw = {}
    if with_grid:
        n = grids["brnd"][0].shape[0]
        kw["dbrnd"] = tuple(np.zeros(n) for _ in range(3))
        kw["dternd"] = np.zeros(n)
    ctx.los_adjoint(d, w, dp, **grids, **kw)
    return dp, kw


def check_params(ctx, d, w, grids, label):
    dp, _ = adjoint(ctx, d, w, grids, with_grid=False)
    scale = np.max(np.abs(dp * np.array([abs(get_param(d, i)) or 1.0 for i in range(6)])))
    for i, name in enumerate(capi.PARAM_NAMES):
        p0 = get_param(d, i)
        h = H_REL * (abs(p0) if p0 != 0 else 1.0)
        set_param(d, i, p0 + h)
        lp = loss(ctx, d, w, grids)
        set_param(d, i, p0 - h)
        lm = loss(ctx, d, w, grids)
        set_param(d, i, p0)
        fd = (lp - lm) / (2 * h)
        err = abs(dp[i] - fd)
        floor = 1e-9 * scale / (abs(p0) or 1.0)
        print(f"{label} d/d{name}: adjoint {dp[i]:.10e} fd {fd:.10e} "
              f"rel {err / max(abs(fd), 1e-300):.2e}", flush=True)
        assert err <= RTOL * abs(fd) + floor, f"{label} {name}: {dp[i]} vs {fd}"


def multishell_case(seed=3):
    par = Param(text=configs.random_forward(nside_sim=8, nside_out=32, oc_r_res=0.05,
                                            grid=(32, 32, 16)))
    d = par.los_desc()
    b, te, box = random_grids((32, 32, 16), seed)
    return d, dict(brnd=b, brnd_box=box, ternd=te, ternd_box=box)


def test_param_gradients_plain(ctx):
    """One shell, analytic fields only, DM + FD + sync."""
    print("case: adjoint params plain nside 16", flush=True)
    d = Param(text=configs.plain_forward(nside_sim=16, nside_out=16, oc_r_res=0.02)).los_desc()
    w = scaled_weights(ctx, d, {}, seed=1)
    check_params(ctx, d, w, {}, "plain")


def test_param_gradients_multishell_grids(ctx):
    """Three shells (8/16/32), random grids, FD carried through the shells."""
    print("case: adjoint params multishell grids", flush=True)
    d, grids = multishell_case()
    w = scaled_weights(ctx, d, grids, seed=2)
    check_params(ctx, d, w, grids, "multishell")


def test_param_gradients_mixed_nsides_cre_unif(ctx):
    """Output maps at other nsides than the shells (transposed interpolation
    in every accumulate) and the uniform CRE sphere; alpha != 3 takes the
    pow() emissivity branch."""
    print("case: adjoint params mixed nsides, CRE unif", flush=True)
    d, grids = multishell_case(seed=4)
    d.nside_dm, d.nside_fd, d.nside_sync = 16, 64, 8
    d.cre_model = capi.HAMX_CRE_UNIF
    d.cre[0:4] = [2.6, 6 * CGS.kpc, 20.6 * CGS.GeV, 0.0217]
    w = scaled_weights(ctx, d, grids, seed=5)
    check_params(ctx, d, w, grids, "mixed")


def test_grid_gradients_directional(ctx):
    """<dL/dgrid, v> against the central difference along a random v, for
    the random-B grid and the random-TE grid, plus single vertices."""
    print("case: adjoint grid directional", flush=True)
    d, grids = multishell_case(seed=6)
    w = scaled_weights(ctx, d, grids, seed=7)
    _, g = adjoint(ctx, d, w, grids, with_grid=True)
    rng = np.random.default_rng(8)
    b = grids["brnd"]
    te = grids["ternd"]
    # random B direction
    v = [rng.normal(size=a.shape) for a in b]
    h = 1e-5 * 2e-6
    plus = dict(grids, brnd=tuple(np.ascontiguousarray(a + h * va) for a, va in zip(b, v)))
    minus = dict(grids, brnd=tuple(np.ascontiguousarray(a - h * va) for a, va in zip(b, v)))
    fd = (loss(ctx, d, w, plus) - loss(ctx, d, w, minus)) / (2 * h)
    adj = sum(float(np.dot(ga, va)) for ga, va in zip(g["dbrnd"], v))
    print(f"brnd directional: adjoint {adj:.10e} fd {fd:.10e}", flush=True)
    assert abs(adj - fd) <= 1e-6 * abs(fd)
    # random TE direction.  The te > 0 clamp makes L piecewise smooth in the
    # TE grid: a central difference across a sample whose density crosses 0
    # is off by up to half that sample's term, so the tight check runs on a
    # positive TE grid (no sample at the clamp) and the grid with negative
    # densities gets a loose one.
    vt = rng.normal(size=te.shape)
    h = 1e-5 * 0.003
    te_pos = np.ascontiguousarray(np.abs(te) + 1e-4)
    gp = dict(grids, ternd=te_pos)
    wp = scaled_weights(ctx, d, gp, seed=7)
    _, g_pos = adjoint(ctx, d, wp, gp, with_grid=True)
    fd = (loss(ctx, d, wp, dict(gp, ternd=np.ascontiguousarray(te_pos + h * vt))) -
          loss(ctx, d, wp, dict(gp, ternd=np.ascontiguousarray(te_pos - h * vt)))) / (2 * h)
    adj = float(np.dot(g_pos["dternd"], vt))
    print(f"ternd directional (positive grid): adjoint {adj:.10e} fd {fd:.10e}", flush=True)
    assert abs(adj - fd) <= 1e-6 * abs(fd)
    fd = (loss(ctx, d, w, dict(grids, ternd=np.ascontiguousarray(te + h * vt))) -
          loss(ctx, d, w, dict(grids, ternd=np.ascontiguousarray(te - h * vt)))) / (2 * h)
    adj = float(np.dot(g["dternd"], vt))
    print(f"ternd directional (clamped grid): adjoint {adj:.10e} fd {fd:.10e}", flush=True)
    assert abs(adj - fd) <= 1e-3 * abs(fd)
    # the vertices with the largest B gradient, one at a time
    gb = g["dbrnd"][0]
    for i in np.argsort(-np.abs(gb))[:3]:
        bp = [a.copy() for a in b]
        bm = [a.copy() for a in b]
        hh = 1e-4 * 2e-6
        bp[0][i] += hh
        bm[0][i] -= hh
        fd = (loss(ctx, d, w, dict(grids, brnd=tuple(bp))) -
              loss(ctx, d, w, dict(grids, brnd=tuple(bm)))) / (2 * hh)
        print(f"vertex {i}: adjoint {gb[i]:.10e} fd {fd:.10e}", flush=True)
        assert abs(gb[i] - fd) <= 1e-5 * abs(fd)


def test_linearity_in_the_adjoint_maps(ctx):
    """The gradient is linear in dL/d(maps): adj(w1 + 2 w2) = adj(w1) + 2 adj(w2)."""
    d, grids = multishell_case(seed=9)
    w1 = scaled_weights(ctx, d, grids, seed=10)
    w2 = scaled_weights(ctx, d, grids, seed=11)
    w3 = {k: np.ascontiguousarray(w1[k] + 2 * w2[k]) for k in w1}
    r1, g1 = adjoint(ctx, d, w1, grids, True)
    r2, g2 = adjoint(ctx, d, w2, grids, True)
    r3, g3 = adjoint(ctx, d, w3, grids, True)
    np.testing.assert_allclose(r3, r1 + 2 * r2, rtol=1e-9, atol=1e-12 * np.abs(r3).max())
    for a, b_, c in zip(g1["dbrnd"], g2["dbrnd"], g3["dbrnd"]):
        np.testing.assert_allclose(c, a + 2 * b_, rtol=1e-9, atol=1e-12 * np.abs(c).max())


def test_autograd_function_matches_c_abi_and_fd(ctx):
    """torch.autograd over the adjoint: the gradient of a Gaussian
    log-likelihood of Q/U and FD maps equals the C-ABI adjoint with
    dL/dmap = -(map - data)/sigma^2, and a central difference in b0."""
    import torch
    from hammurabix_amd.autograd import desc_params, gaussian_loglike, los_maps
    print("case: autograd likelihood", flush=True)
    d, grids = multishell_case(seed=12)
    dev = torch.device("cuda:0")
    b = tuple(torch.from_numpy(a).to(dev).requires_grad_(True) for a in grids["brnd"])
    te = torch.from_numpy(grids["ternd"]).to(dev).requires_grad_(True)
    p = desc_params(d).requires_grad_(True)
    kw = dict(brnd_box=grids["brnd_box"], ternd_box=grids["ternd_box"])
    maps = los_maps(ctx, d, p, brnd=b, ternd=te, **kw)
    rng = np.random.default_rng(13)
    data, sigma = {}, {}
    for k in ("qs", "us", "fd"):
        m = maps[k].detach()
        s = 0.1 * float(torch.sqrt(torch.mean(m * m)))
        data[k] = m + s * torch.from_numpy(rng.normal(size=m.shape[0])).to(dev)
        sigma[k] = s
    ll = gaussian_loglike(maps, data, sigma)
    ll.backward()
    adj = {k: np.ascontiguousarray((-(maps[k].detach() - data[k]) / sigma[k] ** 2).cpu().numpy())
           for k in data}
    dp, g = adjoint(ctx, d, adj, grids, with_grid=True)
    # FP64 atomics sum in a run-dependent order: equal to ~1e-15 of each sum
    np.testing.assert_allclose(p.grad.numpy(), dp, rtol=1e-10, atol=0)
    for got, ref in ((b[0].grad.cpu().numpy(), g["dbrnd"][0]),
                     (te.grad.cpu().numpy(), g["dternd"])):
        assert rel_l2(got, ref) <= 1e-10
    # central difference of the likelihood in b0
    h = H_REL * float(p[0])

    def ll_at(v):
        q = p.detach().clone()
        q[0] = v
        with torch.no_grad():
            mm = los_maps(ctx, d, q, brnd=tuple(x.detach() for x in b), ternd=te.detach(), **kw)
            return float(gaussian_loglike(mm, data, sigma))
    fd = (ll_at(float(p[0]) + h) - ll_at(float(p[0]) - h)) / (2 * h)
    print(f"autograd d/db0: {float(p.grad[0]):.10e} fd {fd:.10e}", flush=True)
    assert abs(float(p.grad[0]) - fd) <= 1e-5 * abs(fd)


def test_param_gradients_vs_oracle_central_differences(ctx):
    """The GPU adjoint against central differences of the FP64 CPU oracle's
    maps (the CPU reference gradient), 3 shells with random grids."""
    from oracle_lib import write_grid
    print("case: adjoint vs oracle FD", flush=True)
    par = Param(text=configs.random_forward(nside_sim=4, nside_out=16, oc_r_res=0.1,
                                            grid=(16, 16, 8)))
    d = par.los_desc()
    b, te, box = random_grids((16, 16, 8), seed=14)
    grids = dict(brnd=b, brnd_box=box, ternd=te, ternd_box=box)
    w = scaled_weights(ctx, d, grids, seed=15)
    dp, _ = adjoint(ctx, d, w, grids, with_grid=False)

    def oracle_loss():
        m = write_grid(d, **grids)
        return sum(float(np.dot(w[k], m[k])) for k in w)
    for i, name in enumerate(capi.PARAM_NAMES):
        p0 = get_param(d, i)
        h = H_REL * (abs(p0) if p0 != 0 else 1.0)
        set_param(d, i, p0 + h)
        lp = oracle_loss()
        set_param(d, i, p0 - h)
        lm = oracle_loss()
        set_param(d, i, p0)
        fd = (lp - lm) / (2 * h)
        print(f"oracle d/d{name}: adjoint {dp[i]:.10e} fd {fd:.10e}", flush=True)
        assert abs(dp[i] - fd) <= 1e-5 * abs(fd), name


def test_tape_reuse_matches_recompute(ctx):
    """hamx_los_adjoint on the kept tape of a forward equals the adjoint that
    re-runs the forward; a tape of another descriptor is refused."""
    print("case: adjoint tape reuse", flush=True)
    d, grids = multishell_case(seed=16)
    w = scaled_weights(ctx, d, grids, seed=17)
    r_re, g_re = adjoint(ctx, d, w, grids, with_grid=True)
    m = out_maps(d)
    ctx.los_forward(d, m, keep_tape=True, **grids)
    dp = np.zeros(capi.HAMX_NPARAM)
    n = grids["brnd"][0].shape[0]
    gb = tuple(np.zeros(n) for _ in range(3))
    gt = np.zeros(n)
    ctx.los_adjoint(d, w, dp, dbrnd=gb, dternd=gt, use_tape=True, **grids)
    # same tape values; only the order of the FP64 atomic sums differs
    np.testing.assert_allclose(dp, r_re, rtol=1e-10, atol=0)
    assert all(rel_l2(a, b) <= 1e-10 for a, b in zip(gb, g_re["dbrnd"]))
    assert rel_l2(gt, g_re["dternd"]) <= 1e-10
    d2, _ = multishell_case(seed=16)
    d2.breg[0] *= 1.1
    with pytest.raises(HamxError) as e:
        ctx.los_adjoint(d2, w, dp, use_tape=True, **grids)
    assert e.value.code == capi.HAMX_ERR_ARG
