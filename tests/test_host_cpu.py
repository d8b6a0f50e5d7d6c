"""Host-side mirrors of the reference interface, no GPU needed: Param parsing
(param.cc), the hampyx XML-editing API (hampyx.py:364-470), and the map /
grid file formats (grid_obs.cc:53-83, grid_brnd.cc:57-98)."""
import os

import numpy as np
import pytest

from hammurabix_amd import Param, _capi as capi, configs
from hammurabix_amd.hampyx import Hampyx
from hammurabix_amd.pipeline import export_map, import_grid, sync_names
from hammurabix_amd.units import CGS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_param_units_and_models():
    p = Param(text=configs.random_forward())
    assert p.observer == [-8.3 * CGS.kpc, 0.0, 0.006 * CGS.kpc]
    g = p.grid_obs
    assert g.nside_shell == [64, 128, 256]
    assert g.radii_shell == [0.0, 7.5 * CGS.kpc, 15 * CGS.kpc, 30 * CGS.kpc]
    assert g.sim_sync_freq == [23.0 * CGS.GHz]
    assert p.breg_type == "lsa" and p.breg_lsa["b0"] == 6.0 * CGS.muGauss
    assert p.breg_lsa["psi0"] == 27.0 * CGS.rad
    assert p.brnd_es["rms"] == 2.0 * CGS.muGauss and p.brnd_seed == 1
    assert p.ternd_dft["rms"] == 0.001 and p.ternd_dft["a0"] == -1.7
    assert p.cre_ana["E0"] == 20.6 * CGS.GeV
    assert p.grid_brnd.nx == 256 and p.grid_brnd.z_max == 4 * CGS.kpc
    d = p.los_desc()
    assert d.breg_model == capi.HAMX_BREG_LSA and d.tereg_model == capi.HAMX_TEREG_UNIF
    assert d.cre_model == capi.HAMX_CRE_ANA and d.n_shell == 3


def test_dgm_is_an_alias_of_dft():
    text = configs.random_forward().replace('<global type="dft">', '<global type="dgm">')
    p = Param(text=text)
    assert p.ternd_method == "dft" and p.ternd_dft["k0"] == 2.0


@pytest.mark.parametrize("old,new,msg", [
    ('<regular cue="1" type="lsa">', '<regular cue="1" type="bogus">', "unsupported breg model"),
    ('<cre cue="1" type="analytic">', '<cre cue="1" type="bogus">', "unsupported cre model"),
    ('<layer type="auto">', '<layer type="bogus">', "unsupported layer option"),
    ('<global type="dft">', '<global type="bogus">', "unsupported ternd model"),
])
def test_unsupported_strings_raise_like_the_reference(old, new, msg):
    with pytest.raises(RuntimeError, match=msg):
        Param(text=configs.random_forward().replace(old, new))


def test_multifrequency_parse_order():
    p = Param(text=configs.plain_forward(freqs=[30.0, 1.4, 0.408]))
    assert p.grid_obs.sim_sync_freq_str == ["30.0", "1.4", "0.408"]
    # the pass for the last-listed entry is built first
    assert p.los_desc().sync_freq == 0.408 * CGS.GHz


def test_sync_names_and_map_export(tmp_path):
    assert sync_names("out/sync_23.bin") == ("out/sync_23_I.bin", "out/sync_23_Q.bin",
                                             "out/sync_23_U.bin")
    v = np.arange(48, dtype=np.float64)
    path = tmp_path / "m.bin"
    export_map(str(path), v)
    assert path.stat().st_size == 48 * 8
    assert np.array_equal(np.fromfile(path, dtype="<f8"), v)


def test_grid_import_checks_size(tmp_path):
    path = tmp_path / "g.bin"
    np.zeros(10).tofile(path)
    with pytest.raises(RuntimeError, match="expected"):
        import_grid(str(path), 4, 3)
    assert import_grid(str(path), 5, 2).shape == (10,)


def test_hampyx_xml_editing(tmp_path, capsys):
    x = tmp_path / "p.xml"
    x.write_text(configs.plain_forward())
    h = Hampyx(xml_path=str(x))
    h.mod_par(keychain=["magneticfield", "regular", "lsa", "b0"], attrib={"value": "3.0"})
    assert h.tree.getroot().find("./magneticfield/regular/lsa/b0").get("value") == "3.0"
    h.add_par(keychain=["observable"], subkey="sync",
              attrib={"cue": "1", "freq": "1.4", "filename": "s.bin", "nside": "32"})
    assert len(h.tree.getroot().findall("./observable/sync")) == 2
    with pytest.raises(ValueError, match="repeatitive"):
        h.add_par(keychain=["observable"], subkey="sync",
                  attrib={"cue": "1", "freq": "1.4", "filename": "s.bin", "nside": "32"})
    h.del_par(keychain=["observable", "sync"], opt="all")
    assert h.tree.getroot().findall("./observable/sync") == []
    with pytest.raises(ValueError, match="wrong path"):
        h.mod_par(keychain=["nothere"], attrib={"a": "1"})
    with pytest.raises(ValueError, match="wrong input"):
        h.mod_par(keychain="grid", attrib={"a": "1"})
    h.print_par(keychain=["grid", "observer"])
    assert "observer" in capsys.readouterr().out
