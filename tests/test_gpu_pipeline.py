"""The reference's run sequence on the device path: multi-frequency passes
(last <sync> first, DM/FD only in the first pass, so later passes see a zero
Faraday cache -- pipeline.cc:165-190), export units and file names, grid
export/import round trips, and the in-process hampyx keys."""
import os

import numpy as np
import pytest

from hammurabix_amd import Context, Param, _capi as capi, configs
from hammurabix_amd.hampyx import Hampyx
from hammurabix_amd.pipeline import Pipeline
from hammurabix_amd.units import CGS
from oracle_lib import rel_l2, write_grid

pytestmark = pytest.mark.gpu


def small_random(tmp, **kw):
    args = dict(nside_sim=4, nside_out=16, oc_r_res=0.1, grid=(32, 32, 16), te_k0=0.3,
                out=str(tmp) + "/")
    args.update(kw)
    return configs.random_forward(**args)


def test_multifrequency_passes_match_the_reference_sequence(ctx, tmp_path):
    text = small_random(tmp_path, freqs=[23.0, 1.4])
    pipe = Pipeline(text=text, ctx=ctx)
    maps = pipe.run()
    # oracle: the same two passes built by hand on the same grids
    par = Param(text=text)
    grids = dict(brnd=tuple(a.cpu().numpy() for a in pipe.brnd),
                 brnd_box=par.grid_brnd.box(), ternd=pipe.ternd.cpu().numpy(),
                 ternd_box=par.grid_ternd.box())
    d0 = par.los_desc()                 # pass 0: 1.4 GHz with DM and FD
    ref0 = write_grid(d0, **grids)
    par.grid_obs.do_dm = par.grid_obs.do_fd = False
    for lst in (par.grid_obs.nside_sync, par.grid_obs.do_sync, par.grid_obs.sim_sync_freq,
                par.grid_obs.sim_sync_name, par.grid_obs.sim_sync_freq_str):
        lst.pop()
    d1 = par.los_desc()                 # pass 1: 23 GHz, DM/FD off
    assert d1.do_dm == 0 and d1.do_fd == 0 and d1.sync_freq == 23 * CGS.GHz
    ref1 = write_grid(d1, **grids)
    assert rel_l2(maps[("dm", "nan", "16", "nan")], ref0["dm"] * CGS.ccm / CGS.pc) < 1e-10
    assert rel_l2(maps[("fd", "nan", "16", "nan")], ref0["fd"] * CGS.m * CGS.m) < 1e-10
    for comp, k in zip("IQU", ("is", "qs", "us")):
        assert rel_l2(maps[("sync", "1.4", "16", comp)], ref0[k]) < 1e-10
        assert rel_l2(maps[("sync", "23.0", "16", comp)], ref1[k]) < 1e-10
    # exported files: raw FP64, DM/FD rescaled, _I/_Q/_U names
    dm = np.fromfile(str(tmp_path / "dm.bin"), dtype="<f8")
    assert np.array_equal(dm, maps[("dm", "nan", "16", "nan")])
    q = np.fromfile(str(tmp_path / "sync_23.0_Q.bin"), dtype="<f8")
    assert np.array_equal(q, maps[("sync", "23.0", "16", "Q")])


def test_grid_export_import_round_trip(ctx, tmp_path):
    bfile, tfile = str(tmp_path / "brnd.bin"), str(tmp_path / "ternd.bin")
    text = small_random(tmp_path)
    text_w = text.replace('<brnd read="0" write="0" filename="brnd.bin"/>',
                          f'<brnd read="0" write="1" filename="{bfile}"/>')
    text_w = text_w.replace('<ternd read="0" write="0" filename="ternd.bin"/>',
                            f'<ternd read="0" write="1" filename="{tfile}"/>')
    first = Pipeline(text=text_w, ctx=ctx, export=False).run()
    assert os.path.getsize(bfile) == 3 * 32 * 32 * 16 * 8
    assert os.path.getsize(tfile) == 32 * 32 * 16 * 8
    # read them back (<fieldio read="1">): the same maps
    text_r = small_random(tmp_path, brnd_file=bfile, ternd_file=tfile)
    second = Pipeline(text=text_r, ctx=ctx, export=False).run()
    for k in first:
        assert np.array_equal(first[k], second[k]), k


def oracle_passes(text, brnd, ternd):
    """sim_map of the reference pipeline's passes on the oracle: the
    last-listed frequency first with DM and FD, the others sync only."""
    par = Param(text=text)
    g = par.grid_obs
    grids = dict(brnd=brnd, brnd_box=par.grid_brnd.box(), ternd=ternd,
                 ternd_box=par.grid_ternd.box())
    out = {}
    n = len(g.do_sync)
    for p in range(n):
        fi = n - 1 - p
        d = par.los_desc()
        d.sync_freq, d.nside_sync = g.sim_sync_freq[fi], g.nside_sync[fi]
        if p > 0:
            d.do_dm = d.do_fd = 0
        m = write_grid(d, **grids)
        for comp, k in zip("IQU", ("is", "qs", "us")):
            out[("sync", g.sim_sync_freq_str[fi], str(d.nside_sync), comp)] = m[k]
        if p == 0:
            out[("dm", "nan", str(d.nside_dm), "nan")] = m["dm"] * CGS.ccm / CGS.pc
            out[("fd", "nan", str(d.nside_fd), "nan")] = m["fd"] * CGS.m * CGS.m
    return out, d


def read_grids(bfile, tfile):
    raw = np.fromfile(bfile, dtype="<f8").reshape(-1, 3)
    return (tuple(np.ascontiguousarray(raw[:, i]) for i in range(3)),
            np.fromfile(tfile, dtype="<f8"))


def test_hampyx_in_process(tmp_path, monkeypatch):
    """The in-process hampyx: the reference's keys, and every map equal to the
    oracle's pass on the grids the run exported."""
    bfile, tfile = str(tmp_path / "brnd.bin"), str(tmp_path / "ternd.bin")
    text = small_random(tmp_path, freqs=[30.0, 0.408])
    text = text.replace('<brnd read="0" write="0" filename="brnd.bin"/>',
                        f'<brnd read="0" write="1" filename="{bfile}"/>')
    text = text.replace('<ternd read="0" write="0" filename="ternd.bin"/>',
                        f'<ternd read="0" write="1" filename="{tfile}"/>')
    x = tmp_path / "params.xml"
    x.write_text(text)
    monkeypatch.chdir(tmp_path)
    h = Hampyx()
    h.mod_par(keychain=["magneticfield", "regular", "lsa", "b0"], attrib={"value": "4.0"})
    h()
    keys = set(h.sim_map)
    for f in ("30.0", "0.408"):
        for c in "IQU":
            assert ("sync", f, "16", c) in keys
    assert ("fd", "nan", "16", "nan") in keys and ("dm", "nan", "16", "nan") in keys
    assert all(np.all(np.isfinite(v)) for v in h.sim_map.values())
    b, te = read_grids(bfile, tfile)
    import xml.etree.ElementTree as et
    ref, _ = oracle_passes(et.tostring(h.tree.getroot(), encoding="unicode"), b, te)
    assert set(ref) == keys
    for k, v in ref.items():
        err = rel_l2(h.sim_map[k], v)
        assert err <= 1e-10, (k, err)


def test_hampyx_differentiable_vs_oracle_central_differences(tmp_path, monkeypatch):
    """Hampyx.differentiable(): L = sum over every sim_map entry (two <sync>
    passes, DM and FD) of w * map with fixed random weights; backward through
    the passes gives dL/d(b0, psi0, psi1, chi0, alpha, n0), compared with
    central differences (h = 1e-6 relative) of the same L on the oracle's
    passes over the grids the run exported; and the random-B grid gradient
    along a random direction likewise."""
    import torch
    print("case: hampyx differentiable", flush=True)
    bfile, tfile = str(tmp_path / "brnd.bin"), str(tmp_path / "ternd.bin")
    text = small_random(tmp_path, freqs=[23.0, 1.4])
    text = text.replace('<brnd read="0" write="0" filename="brnd.bin"/>',
                        f'<brnd read="0" write="1" filename="{bfile}"/>')
    text = text.replace('<ternd read="0" write="0" filename="ternd.bin"/>',
                        f'<ternd read="0" write="1" filename="{tfile}"/>')
    x = tmp_path / "params.xml"
    x.write_text(text)
    monkeypatch.chdir(tmp_path)
    h = Hampyx(xml_path=str(x))
    maps, params, grids = h.differentiable(grids_grad=True)
    b, te = read_grids(bfile, tfile)
    ref, _ = oracle_passes(text, b, te)
    assert set(maps) == set(ref)
    rng = np.random.default_rng(3)
    w = {}
    for k, v in ref.items():
        s = float(np.sqrt(np.mean(v ** 2))) or 1.0
        w[k] = rng.normal(size=v.shape) / s
        assert rel_l2(maps[k].detach().cpu().numpy(), v) <= 1e-10, k
    L = sum(torch.dot(torch.from_numpy(w[k]).to(maps[k].device), maps[k]) for k in maps)
    L.backward()
    grad = params.grad.cpu().numpy()

    def oracle_loss(text_p, bb=b):
        r, _ = oracle_passes(text_p, bb, te)
        return sum(float(np.dot(w[k], r[k])) for k in r)
    # parameters, stepped through the XML values the descriptor is read from
    xml_keys = [("magneticfield/regular/lsa/b0", 6.0), ("magneticfield/regular/lsa/psi0", 27.0),
                ("magneticfield/regular/lsa/psi1", 0.9), ("magneticfield/regular/lsa/chi0", 25.0),
                ("cre/analytic/alpha", 3.0), ("thermalelectron/regular/unif/n0", 0.01)]
    import xml.etree.ElementTree as et
    for i, ((path, v0), name) in enumerate(zip(xml_keys, capi.PARAM_NAMES)):
        hstep = 1e-6 * v0
        vals = []
        for v in (v0 + hstep, v0 - hstep):
            root = et.fromstring(text)
            root.find("./" + path).set("value", repr(v))
            vals.append(oracle_loss(et.tostring(root, encoding="unicode")))
        # the XML value is in muG / degrees / cm^-3: the step in the
        # descriptor's units, through the descriptor itself
        dp = abs(d_last_param(text, path, v0 + hstep) - d_last_param(text, path, v0 - hstep))
        fd = (vals[0] - vals[1]) / dp
        print(f"hampyx dL/d{name}: autograd {grad[i]:.10e} oracle FD {fd:.10e}", flush=True)
        assert abs(grad[i] - fd) <= 1e-5 * abs(fd), name
    # the random-B grid along a random direction
    v = [rng.normal(size=a.shape) for a in b]
    hb = 1e-6 * 2e-6
    fd = (oracle_loss(text, tuple(np.ascontiguousarray(a + hb * va) for a, va in zip(b, v))) -
          oracle_loss(text, tuple(np.ascontiguousarray(a - hb * va) for a, va in zip(b, v)))) \
        / (2 * hb)
    got = sum(float(np.dot(g.grad.cpu().numpy(), va)) for g, va in zip(grids[:3], v))
    print(f"hampyx <dL/dB, v>: autograd {got:.10e} oracle FD {fd:.10e}", flush=True)
    assert abs(got - fd) <= 1e-5 * abs(fd)


def d_last_param(text, path, value):
    """The descriptor value (CGS / radians) of the XML parameter at `path`
    set to `value`."""
    import xml.etree.ElementTree as et
    from hammurabix_amd.autograd import desc_params
    root = et.fromstring(text)
    root.find("./" + path).set("value", repr(value))
    d = Param(text=et.tostring(root, encoding="unicode")).los_desc()
    idx = {"b0": 0, "psi0": 1, "psi1": 2, "chi0": 3, "alpha": 4, "n0": 5}[path.split("/")[-1]]
    return float(desc_params(d)[idx])
