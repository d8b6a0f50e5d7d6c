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


def test_hampyx_in_process(tmp_path, monkeypatch):
    x = tmp_path / "params.xml"
    x.write_text(small_random(tmp_path, freqs=[30.0, 0.408]))
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
