"""GPU checks of the batched LOS path (hamx_batch_*) against the plain path
run pass by pass the way Pipeline::assemble_obs runs it (pipeline.cc:165-190):
the last-listed frequency first with DM / FD, then the others without.

FP64 batches must equal the plain path to rounding (1e-12 relative L2 per
map: per-pass emissivity factors are formed as ratios, which moves the
last bits); FP32 batches must stay within 1e-5 relative L2 per map, the
batched path's stated tolerance.  Every replay of the captured graph must
reproduce the first, and a replay after the input grids change must follow
the new grids.
"""
import ctypes as C

import numpy as np
import pytest

from hammurabix_amd import HamxError, Param, _capi as capi, configs
from hammurabix_amd.units import CGS
from oracle_lib import npix, rel_l2

pytestmark = pytest.mark.gpu

FREQS = [0.408e9, 1.4e9, 23e9, 30e9]


def torch_dev():
    import torch
    return torch, torch.device("cuda:0")


def grids(shape, seed):
    torch, dev = torch_dev()
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    b = tuple(torch.from_numpy(rng.normal(0, 2e-6, n)).to(dev) for _ in range(3))
    te = torch.from_numpy(rng.normal(0.0, 0.003, n)).to(dev)
    box = capi.HamxBox(*shape, -20 * CGS.kpc, 20 * CGS.kpc, -20 * CGS.kpc, 20 * CGS.kpc,
                       -4 * CGS.kpc, 4 * CGS.kpc)
    return b, te, box


def instance(nside_sim, nside_out, freqs, b0=None, psi0=None, g=None, alpha=None,
             cre_unif=False):
    torch, dev = torch_dev()
    d = Param(text=configs.random_forward(nside_sim=nside_sim, nside_out=nside_out,
                                          oc_r_res=0.05, grid=(32, 32, 16))).los_desc()
    if b0 is not None:
        d.breg[0] = b0
    if psi0 is not None:
        d.breg[1] = psi0
    if alpha is not None:
        d.cre[0] = alpha
    if cre_unif:
        d.cre_model = capi.HAMX_CRE_UNIF
        d.cre[0:4] = [alpha or 3.0, 6 * CGS.kpc, 20.6 * CGS.GeV, 0.0217]
    z = lambda n: torch.zeros(npix(n), dtype=torch.float64, device=dev)  # noqa: E731
    maps = dict(dm=z(d.nside_dm), fd=z(d.nside_fd), **{k: [z(d.nside_sync) for _ in freqs]
                                                        for k in ("is", "qs", "us")})
    spec = dict(desc=d, freqs=list(freqs), maps=maps)
    if g is not None:
        spec.update(brnd=g[0], ternd=g[1], box=g[2])
    return spec


def plain_passes(ctx, spec):
    """The reference pipeline's passes on the plain path -> expected maps."""
    d0 = spec["desc"]
    freqs = spec["freqs"]
    exp = {"is": [None] * len(freqs), "qs": [None] * len(freqs), "us": [None] * len(freqs)}
    kw = {}
    if "brnd" in spec:
        b = tuple(a.cpu().numpy() for a in spec["brnd"])
        kw = dict(brnd=b, brnd_box=spec["box"], ternd=spec["ternd"].cpu().numpy(),
                  ternd_box=spec["box"])
    for p, fi in enumerate(reversed(range(len(freqs)))):
        d = capi.HamxLosDesc()
        C.memmove(C.byref(d), C.byref(d0), C.sizeof(d))
        d.sync_freq = freqs[fi]
        d.do_sync = 1
        if p > 0:
            d.do_dm = d.do_fd = 0
        m = {k: np.zeros(npix(d.nside_sync)) for k in ("is", "qs", "us")}
        m["fd"] = np.zeros(npix(d.nside_fd if d.do_fd else d.nside_sync))
        if d.do_dm:
            m["dm"] = np.zeros(npix(d.nside_dm))
        ctx.los_forward(d, m, **kw)
        for k in ("is", "qs", "us"):
            exp[k][fi] = m[k]
        if p == 0:
            exp["dm"], exp["fd"] = m["dm"], m["fd"]
    return exp


def compare(spec, exp, tol, label):
    worst = 0.0
    m = spec["maps"]
    pairs = [("dm", m["dm"], exp["dm"]), ("fd", m["fd"], exp["fd"])]
    for k in ("is", "qs", "us"):
        pairs += [(f"{k}[{i}]", a, e) for i, (a, e) in enumerate(zip(m[k], exp[k]))]
    for name, got, ref in pairs:
        g = got.cpu().numpy()
        assert np.all(np.isfinite(g)), name
        err = rel_l2(g, ref)
        worst = max(worst, err)
        assert err <= tol, f"{label} {name}: rel L2 {err:.3e} > {tol}"
    print(f"{label}: worst rel L2 {worst:.3e}", flush=True)
    return worst


def mixed_batch():
    g1 = grids((32, 32, 16), 1)
    g2 = grids((32, 32, 16), 2)
    return [
        instance(8, 32, FREQS[2:3], g=g1),
        instance(4, 16, FREQS[:2], b0=5e-6, psi0=0.45, g=g2),
        instance(8, 32, FREQS, b0=7e-6, psi0=0.55, g=g1),
        instance(4, 16, FREQS[1:3], g=None),               # no random grids
        instance(4, 16, FREQS[2:], g=g2, alpha=2.6, cre_unif=True),  # pow() emissivity
    ]


@pytest.mark.parametrize("prec,tol", [(capi.HAMX_PREC_F64, 1e-12), (capi.HAMX_PREC_F32, 1e-5)])
def test_batch_matches_plain_passes(ctx, prec, tol):
    print(f"case: batch vs plain, precision {prec}", flush=True)
    specs = mixed_batch()
    with ctx.batch(specs, precision=prec, tune=False) as b:
        b.run()
        for i, s in enumerate(specs):
            compare(s, plain_passes(ctx, s), tol, f"prec{prec} inst{i}")
        info = b.info()
        assert info["levels"] == 3
        assert info["samples_ref"] > info["samples_geo"] > 0


def test_graph_replays_and_follows_new_grids(ctx):
    """Three replays give identical maps; scaling the random grids in place
    and replaying gives the plain path's maps for the new grids."""
    print("case: batch graph replays", flush=True)
    g = grids((32, 32, 16), 3)
    specs = [instance(8, 32, FREQS[1:3], g=g), instance(4, 16, FREQS[2:3], g=g, b0=4e-6)]
    with ctx.batch(specs, precision=capi.HAMX_PREC_F64, tune=False) as b:
        b.run()
        first = [{k: ([t.clone() for t in v] if isinstance(v, list) else v.clone())
                  for k, v in s["maps"].items()} for s in specs]
        for _ in range(2):
            b.run()
            for s, f in zip(specs, first):
                for k, v in s["maps"].items():
                    vs = v if isinstance(v, list) else [v]
                    fs = f[k] if isinstance(f[k], list) else [f[k]]
                    for a, c in zip(vs, fs):
                        assert bool((a == c).all()), k
        for a in g[0]:
            a.mul_(1.7)
        g[1].mul_(0.5)
        b.run()
        for i, s in enumerate(specs):
            compare(s, plain_passes(ctx, s), 1e-12, f"regrid inst{i}")


def test_tuned_shapes_give_identical_maps(ctx):
    """The lanes-per-sample launch shapes compute each ray the same way, so
    a tuned plan that picks one of them equals the untuned plan bit for bit;
    the one-ray-per-lane forward sums its rays in another order (FP64 lane
    sums instead of FP32 lane partials and a wave sum) and equals it to the
    FP32 batch's rounding."""
    print("case: batch tuned vs untuned", flush=True)
    import torch
    specs_a = mixed_batch()
    with ctx.batch(specs_a, precision=capi.HAMX_PREC_F32, tune=False) as b:
        b.run()
    specs_b = mixed_batch()
    with ctx.batch(specs_b, precision=capi.HAMX_PREC_F32, tune=True) as b:
        b.run()
        info = b.info()
        print("tuned:", info["shape"], info["variant_ms"], flush=True)
        assert len(info["variant_ms"]) >= 4
    lane_shape = "lane per ray" in info["shape"]
    for sa, sb in zip(specs_a, specs_b):
        for k in sa["maps"]:
            va, vb = sa["maps"][k], sb["maps"][k]
            for a, c in zip(va if isinstance(va, list) else [va],
                            vb if isinstance(vb, list) else [vb]):
                if lane_shape:
                    err = rel_l2(c.cpu().numpy(), a.cpu().numpy())
                    assert err <= 1e-6, (k, err)
                else:
                    assert torch.equal(a, c), k


@pytest.mark.parametrize("adjoint", [False, True])
def test_lane_per_ray_forward(ctx, adjoint):
    """The FP32 forward with one ray per lane (HAMX_BATCH_LANES) against the
    plain FP64 passes, at the FP32 batch's map tolerance; with the reverse
    graph captured too, the same maps."""
    print(f"case: batch lane per ray, adjoint {adjoint}", flush=True)
    specs = mixed_batch()
    if adjoint:
        specs = with_adjoint(specs, seed=23)
    with ctx.batch(specs, precision=capi.HAMX_PREC_F32, lanes=True, adjoint=adjoint) as b:
        assert "lane per ray" in b.info()["shape"]
        b.run()
        for i, s in enumerate(specs):
            compare(s, plain_passes(ctx, s), 1e-5, f"lanes inst{i}")
        if not adjoint:
            return
        b.run_adjoint()
    # the reverse graph reads the forward's tape: the same gradients as a
    # plan with the lanes-per-sample forward, to the FP32 batch's rounding
    ref = with_adjoint(mixed_batch(), seed=23)
    with ctx.batch(ref, precision=capi.HAMX_PREC_F32, tune=False, adjoint=True) as b:
        b.run()
        b.run_adjoint()
    for i, (s, r) in enumerate(zip(specs, ref)):
        for name in ("dparams", "dbrnd", "dternd"):
            if name not in r["grads"]:
                continue
            got, exp = s["grads"][name], r["grads"][name]
            for g, e in zip(got if isinstance(got, tuple) else [got],
                            exp if isinstance(exp, tuple) else [exp]):
                err = rel_l2(g.cpu().numpy(), e.cpu().numpy())
                print(f"lanes inst{i} {name}: rel L2 vs lanes-per-sample {err:.3e}", flush=True)
                assert err <= 1e-5, (i, name, err)


def test_lsa_angles_beyond_a_quarter_turn_in_batches(ctx):
    """psi0 or chi0 beyond 45 degrees: the FP32 LSA's sines and cosines take
    the quadrant-reducing path instead of the series, in both precisions."""
    print("case: batch LSA angles beyond pi/4", flush=True)
    g1 = grids((32, 32, 16), 1)
    for prec, tol in ((capi.HAMX_PREC_F32, 1e-5), (capi.HAMX_PREC_F64, 1e-12)):
        specs = [instance(4, 16, FREQS[2:3], psi0=1.2, g=g1),
                 instance(4, 16, FREQS[1:3], g=g1)]
        specs[1]["desc"].breg[3] = -1.3  # chi0
        with ctx.batch(specs, precision=prec, tune=False) as b:
            b.run()
            for i, s in enumerate(specs):
                compare(s, plain_passes(ctx, s), tol, f"quarter-turn prec{prec} inst{i}")


def test_lane_per_ray_needs_f32(ctx):
    with pytest.raises(RuntimeError, match="HAMX_BATCH_LANES"):
        ctx.batch(mixed_batch(), precision=capi.HAMX_PREC_F64, lanes=True)


def test_batch_rejects_bad_instances(ctx):
    s = instance(4, 16, FREQS[2:3])
    s["desc"].breg_model = capi.HAMX_BREG_CART
    with pytest.raises(HamxError) as e:
        ctx.batch([s], tune=False)
    assert e.value.code == capi.HAMX_ERR_UNSUPPORTED
    s = instance(4, 16, FREQS[2:3])
    s["desc"].nside_shell[0] = 6
    with pytest.raises(HamxError) as e:
        ctx.batch([s], tune=False)
    assert e.value.code == capi.HAMX_ERR_NSIDE
    s = instance(4, 16, FREQS[2:3])
    s["maps"]["is"][0] = s["maps"]["is"][0].cpu()
    with pytest.raises(ValueError):
        ctx.batch([s], tune=False)


def with_adjoint(specs, seed):
    """Random dL/d(maps) (scaled to each map's size) and gradient outputs."""
    torch, dev = torch_dev()
    gen = torch.Generator(device=dev).manual_seed(seed)
    for sp in specs:
        m = sp["maps"]
        rnd = lambda t: torch.randn(t.shape[0], dtype=torch.float64, device=dev,  # noqa: E731
                                    generator=gen)
        sp["adj"] = {"dm": rnd(m["dm"]), "fd": rnd(m["fd"]),
                     **{k: [rnd(t) for t in m[k]] for k in ("is", "qs", "us")}}
        gr = {"dparams": torch.zeros(capi.HAMX_NPARAM, dtype=torch.float64, device=dev)}
        if "brnd" in sp:
            n = sp["brnd"][0].shape[0]
            gr["dbrnd"] = tuple(torch.zeros(n, dtype=torch.float64, device=dev)
                                for _ in range(3))
            gr["dternd"] = torch.zeros(n, dtype=torch.float64, device=dev)
        sp["grads"] = gr
    return specs


def scale_adjoint(specs):
    """Divide each adjoint map by its map's rms, so every map weighs alike."""
    import torch
    for sp in specs:
        for k, v in sp["maps"].items():
            vs = v if isinstance(v, list) else [v]
            adj = sp["adj"][k] if isinstance(v, list) else [sp["adj"][k]]
            for a, mm in zip(adj, vs):
                r = float(torch.sqrt(torch.mean(mm * mm)))
                if r > 0:
                    a.div_(r)


def plain_adjoint_passes(ctx, spec):
    """Sum over the reference pipeline's passes of the plain-path adjoint."""
    d0, freqs, adj = spec["desc"], spec["freqs"], spec["adj"]
    kw = {}
    n = 0
    if "brnd" in spec:
        kw = dict(brnd=tuple(a.cpu().numpy() for a in spec["brnd"]), brnd_box=spec["box"],
                  ternd=spec["ternd"].cpu().numpy(), ternd_box=spec["box"])
        n = spec["brnd"][0].shape[0]
    dp_tot = np.zeros(capi.HAMX_NPARAM)
    gb_tot = [np.zeros(n) for _ in range(3)]
    gt_tot = np.zeros(n)
    for p, fi in enumerate(reversed(range(len(freqs)))):
        d = capi.HamxLosDesc()
        C.memmove(C.byref(d), C.byref(d0), C.sizeof(d))
        d.sync_freq, d.do_sync = freqs[fi], 1
        if p > 0:
            d.do_dm = d.do_fd = 0
        a = {k: np.ascontiguousarray(adj[k][fi].cpu().numpy()) for k in ("is", "qs", "us")}
        if p == 0:
            a["dm"] = np.ascontiguousarray(adj["dm"].cpu().numpy())
            a["fd"] = np.ascontiguousarray(adj["fd"].cpu().numpy())
        dp = np.zeros(capi.HAMX_NPARAM)
        gkw = {}
        if n:
            gkw = dict(dbrnd=tuple(np.zeros(n) for _ in range(3)), dternd=np.zeros(n))
        ctx.los_adjoint(d, a, dp, **kw, **gkw)
        dp_tot += dp
        if n:
            for t, g in zip(gb_tot, gkw["dbrnd"]):
                t += g
            gt_tot += gkw["dternd"]
    return dp_tot, gb_tot, gt_tot


@pytest.mark.parametrize("prec,tol", [(capi.HAMX_PREC_F64, 1e-9), (capi.HAMX_PREC_F32, 1e-4)])
def test_batch_adjoint_matches_plain_adjoint_passes(ctx, prec, tol):
    """Batched reverse mode against the plain hamx_los_adjoint summed over
    the pipeline's passes: parameters per component and the B / TE grid
    gradients in relative L2."""
    print(f"case: batch adjoint vs plain, precision {prec}", flush=True)
    specs2 = with_adjoint(mixed_batch(), seed=21)
    with ctx.batch(specs2, precision=prec, tune=False, adjoint=True) as b:
        b.run()
        scale_adjoint(specs2)  # in place: the plan's addresses stay valid
        b.run_adjoint()
        for i, sp in enumerate(specs2):
            dp, gb, gt = plain_adjoint_passes(ctx, sp)
            got = sp["grads"]["dparams"].cpu().numpy()
            rel = np.abs(got - dp) / np.maximum(np.abs(dp), 1e-300)
            print(f"prec{prec} inst{i} dparams rel {np.array2string(rel, precision=2)}",
                  flush=True)
            assert np.all(rel <= tol), (i, got, dp)
            if "dbrnd" in sp["grads"]:
                for name, g, r in (("bx", sp["grads"]["dbrnd"][0], gb[0]),
                                   ("by", sp["grads"]["dbrnd"][1], gb[1]),
                                   ("bz", sp["grads"]["dbrnd"][2], gb[2]),
                                   ("te", sp["grads"]["dternd"], gt)):
                    err = rel_l2(g.cpu().numpy(), r)
                    print(f"prec{prec} inst{i} d{name} rel L2 {err:.2e}", flush=True)
                    assert err <= tol, (i, name, err)
        # a second replay of both graphs gives the same gradients
        first = [sp["grads"]["dparams"].clone() for sp in specs2]
        b.run()
        b.run_adjoint()
        for sp, f in zip(specs2, first):
            np.testing.assert_allclose(sp["grads"]["dparams"].cpu().numpy(), f.cpu().numpy(),
                                       rtol=1e-12 if prec == capi.HAMX_PREC_F64 else 1e-5)


def test_batch_adjoint_needs_the_flag(ctx):
    s = instance(4, 16, FREQS[2:3])
    with ctx.batch([s], tune=False) as b:
        with pytest.raises(HamxError) as e:
            b.run_adjoint()
        assert e.value.code == capi.HAMX_ERR_ARG


def large_specs():
    """Two instances in the shape of the large batched adjoint (512x512x128
    grids, 4 shells, 4 frequencies, DM / FD / sync), at a small sky
    resolution: shells at nside 4..32 and 8..64 (3 and 12 NESTED patches of
    64 rays in the innermost shell of the reverse march)."""
    torch, dev = torch_dev()
    specs = []
    for i, ns in enumerate((4, 8)):
        par = Param(text=configs.large_instance(i, nside_sim=ns, nside_out=8 * ns,
                                                oc_r_res=0.02))
        d = par.los_desc()
        n = par.grid_brnd.full_size
        gen = torch.Generator(device=dev).manual_seed(100 + i)
        b = tuple(torch.randn(n, dtype=torch.float64, device=dev, generator=gen) * 2e-6
                  for _ in range(3))
        te = torch.randn(n, dtype=torch.float64, device=dev, generator=gen) * 0.003
        z = lambda k: torch.zeros(npix(k), dtype=torch.float64, device=dev)  # noqa: E731
        freqs = list(par.grid_obs.sim_sync_freq)
        maps = dict(dm=z(d.nside_dm), fd=z(d.nside_fd),
                    **{k: [z(d.nside_sync) for _ in freqs] for k in ("is", "qs", "us")})
        specs.append(dict(desc=d, freqs=freqs, maps=maps, brnd=b, ternd=te,
                          box=par.grid_brnd.box()))
    return specs


def plain_dev(ctx, spec):
    """Plain forward and adjoint passes with device-resident grids and maps:
    (expected maps, dparams, (dbx, dby, dbz), dte) summed over the passes."""
    torch, dev = torch_dev()
    d0, freqs, adj = spec["desc"], spec["freqs"], spec["adj"]
    kw = dict(brnd=spec["brnd"], brnd_box=spec["box"], ternd=spec["ternd"],
              ternd_box=spec["box"])
    n = spec["brnd"][0].shape[0]
    exp = {"is": [None] * len(freqs), "qs": [None] * len(freqs), "us": [None] * len(freqs)}
    dp_tot = torch.zeros(capi.HAMX_NPARAM, dtype=torch.float64, device=dev)
    gb_tot = [torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3)]
    gt_tot = torch.zeros(n, dtype=torch.float64, device=dev)
    for p, fi in enumerate(reversed(range(len(freqs)))):
        d = capi.HamxLosDesc()
        C.memmove(C.byref(d), C.byref(d0), C.sizeof(d))
        d.sync_freq, d.do_sync = freqs[fi], 1
        if p > 0:
            d.do_dm = d.do_fd = 0
        z = lambda k: torch.zeros(npix(k), dtype=torch.float64, device=dev)  # noqa: E731
        m = {k: z(d.nside_sync) for k in ("is", "qs", "us")}
        m["fd"] = z(d.nside_fd if d.do_fd else d.nside_sync)
        if d.do_dm:
            m["dm"] = z(d.nside_dm)
        ctx.los_forward(d, m, keep_tape=True, **kw)
        for k in ("is", "qs", "us"):
            exp[k][fi] = m[k].cpu().numpy()
        if p == 0:
            exp["dm"], exp["fd"] = m["dm"].cpu().numpy(), m["fd"].cpu().numpy()
        a = {k: adj[k][fi] for k in ("is", "qs", "us")}
        if p == 0:
            a["dm"], a["fd"] = adj["dm"], adj["fd"]
        dp = torch.zeros(capi.HAMX_NPARAM, dtype=torch.float64, device=dev)
        gb = tuple(torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3))
        gt = torch.zeros(n, dtype=torch.float64, device=dev)
        ctx.los_adjoint(d, a, dp, dbrnd=gb, dternd=gt, use_tape=True, **kw)
        dp_tot += dp
        for t, g in zip(gb_tot, gb):
            t += g
        gt_tot += gt
    return exp, dp_tot.cpu().numpy(), [g.cpu().numpy() for g in gb_tot], gt_tot.cpu().numpy()


def test_large_shape_batch_forward_and_adjoint(ctx):
    """FP32 batch on 512x512x128 grids with 4 shells and 4 frequencies: maps
    within 1e-5 and gradients within 1e-4 of the FP64 plain path."""
    print("case: large-shape FP32 batch, forward and adjoint", flush=True)
    specs = with_adjoint(large_specs(), seed=31)
    with ctx.batch(specs, precision=capi.HAMX_PREC_F32, tune=False, adjoint=True) as b:
        b.run()
        scale_adjoint(specs)
        b.run_adjoint()
        for i, sp in enumerate(specs):
            exp, dp, gb, gt = plain_dev(ctx, sp)
            compare(sp, exp, 1e-5, f"large inst{i}")
            got = sp["grads"]["dparams"].cpu().numpy()
            rel = np.abs(got - dp) / np.maximum(np.abs(dp), 1e-300)
            print(f"large inst{i} dparams rel {np.array2string(rel, precision=2)}", flush=True)
            assert np.all(rel <= 1e-4), (i, got, dp)
            for name, g, r in (("bx", sp["grads"]["dbrnd"][0], gb[0]),
                               ("by", sp["grads"]["dbrnd"][1], gb[1]),
                               ("bz", sp["grads"]["dbrnd"][2], gb[2]),
                               ("te", sp["grads"]["dternd"], gt)):
                err = rel_l2(g.cpu().numpy(), r)
                print(f"large inst{i} d{name} rel L2 {err:.2e}", flush=True)
                assert err <= 1e-4, (i, name, err)
