"""The C++ host mirror (Param / plugins / Grid_* / Integrator / Pipeline over
the C-ABI) driven through the reference's command line, `hamx <params.xml>`."""
import os
import subprocess

import numpy as np
import pytest

from hammurabix_amd import Param, configs
from hammurabix_amd.pipeline import Pipeline
from hammurabix_amd.units import CGS
from oracle_lib import rel_l2, write_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HAMX = os.path.join(ROOT, "hammurabix_amd", "host", "hamx")


def need_cli():
    if not os.path.exists(HAMX):
        pytest.skip("hamx CLI not built (python -m hammurabix_amd._build)")


def run(*args):
    return subprocess.run([HAMX, *args], capture_output=True, text=True, timeout=300)


def test_cli_usage_and_bad_file():
    need_cli()
    r = run()
    assert r.returncode == 1 and "usage" in r.stderr
    r = run("/nonexistent/params.xml")
    assert r.returncode == 1 and "cannot open parameter file" in r.stderr


def test_cli_rejects_unsupported_models(tmp_path):
    need_cli()
    x = tmp_path / "p.xml"
    x.write_text(configs.plain_forward().replace('type="analytic"', 'type="bogus"'))
    r = run(str(x))
    assert r.returncode == 1 and "unsupported cre model" in r.stderr


@pytest.mark.gpu
def test_cli_matches_python_pipeline_and_oracle(tmp_path, ctx):
    need_cli()
    out = str(tmp_path) + "/"
    text = configs.random_forward(nside_sim=4, nside_out=16, oc_r_res=0.1, grid=(32, 32, 16),
                                  te_k0=0.3,
                                  freqs=[23.0, 1.4], out=out)
    bfile = str(tmp_path / "brnd.bin")
    text = text.replace('<brnd read="0" write="0" filename="brnd.bin"/>',
                        f'<brnd read="0" write="1" filename="{bfile}"/>')
    tfile = str(tmp_path / "ternd.bin")
    text = text.replace('<ternd read="0" write="0" filename="ternd.bin"/>',
                        f'<ternd read="0" write="1" filename="{tfile}"/>')
    x = tmp_path / "p.xml"
    x.write_text(text)
    r = run(str(x))
    assert r.returncode == 0, r.stderr
    py = Pipeline(text=text, ctx=ctx, export=False).run()
    files = {("dm", "nan", "16", "nan"): "dm.bin", ("fd", "nan", "16", "nan"): "fd.bin"}
    for f in ("23.0", "1.4"):
        for c in "IQU":
            files[("sync", f, "16", c)] = f"sync_{f}_{c}.bin"
    for key, name in files.items():
        got = np.fromfile(str(tmp_path / name), dtype="<f8")
        assert np.array_equal(got, py[key]), key
    # the first pass against the oracle on the exported grids
    par = Param(text=text)
    raw = np.fromfile(bfile, dtype="<f8").reshape(-1, 3)
    b = tuple(np.ascontiguousarray(raw[:, i]) for i in range(3))
    te = np.fromfile(tfile, dtype="<f8")
    ref = write_grid(par.los_desc(), brnd=b, brnd_box=par.grid_brnd.box(), ternd=te,
                     ternd_box=par.grid_ternd.box())
    dm = np.fromfile(str(tmp_path / "dm.bin"), dtype="<f8")
    assert rel_l2(dm, ref["dm"] * CGS.ccm / CGS.pc) < 1e-10
    q = np.fromfile(str(tmp_path / "sync_1.4_Q.bin"), dtype="<f8")
    assert rel_l2(q, ref["qs"]) < 1e-10


@pytest.mark.gpu
def test_cli_and_pipeline_with_mask(tmp_path, ctx):
    """<mask cue="1">: the CLI and the Python pipeline read the mask file and
    give the oracle's masked maps; masked pixels are zero."""
    need_cli()
    import os
    out = str(tmp_path) + "/"
    mfile = os.path.join(os.path.dirname(__file__), "golden", "mask_map_nside8.bin")
    text = configs.plain_forward(nside_sim=8, nside_out=8, oc_r_res=0.1, out=out,
                                 mask_file=mfile, nside_mask=8)
    x = tmp_path / "m.xml"
    x.write_text(text)
    r = run(str(x))
    assert r.returncode == 0, r.stderr
    py = Pipeline(text=text, ctx=ctx, export=False).run()
    piv = np.ascontiguousarray(np.fromfile(mfile, dtype="<f8"))
    ref = write_grid(Param(text=text).los_desc(), mask=piv, nside_mask=8)
    q_cli = np.fromfile(str(tmp_path / "sync_23.0_Q.bin"), dtype="<f8")
    assert np.array_equal(q_cli, py[("sync", "23.0", "8", "Q")])
    assert rel_l2(q_cli, ref["qs"]) < 1e-10
    assert np.all(q_cli[piv != 1.0] == 0.0)


@pytest.mark.gpu
def test_cli_jf12_helix_jaffe_match_python_pipeline_and_oracle(tmp_path, ctx):
    """The C++ host reads the reference's <jf12> / <helix> / <jaffe> blocks and runs them
    through the device evaluators: its maps equal the Python pipeline's and
    the oracle's (no random fields)."""
    need_cli()
    jf12 = ('<jf12><b_arm1 value="0.1"/><b_arm2 value="3.0"/><b_arm3 value="-0.9"/>'
            '<b_arm4 value="-0.8"/><b_arm5 value="-2.0"/><b_arm6 value="-4.2"/>'
            '<b_arm7 value="0.0"/><b_ring value="0.1"/><h_disk value="0.40"/>'
            '<w_disk value="0.27"/><Bn value="1.4"/><Bs value="-1.1"/><rn value="9.22"/>'
            '<rs value="16.7"/><wh value="0.20"/><z0 value="5.3"/><B0_X value="4.6"/>'
            '<Xtheta value="49"/><rpc_X value="4.8"/><r0_X value="2.9"/></jf12>')
    helix = ('<helix><bx value="1.0"/><by value="2.0"/><bz value="0.5"/>'
             '<r_min value="3.0"/><r_max value="12.0"/></helix>')
    jaffe = ('<jaffe><quadruple cue="0"/><bss cue="0"/><disk_amp value="0.167"/>'
             '<disk_z0 value="0.1"/><halo_amp value="1.38"/><halo_z0 value="3.0"/>'
             '<r_inner value="0.5"/><r_scale value="20"/><r_peak value="0"/><ring cue="0"/>'
             '<bar cue="1"/><ring_amp value="0.023"/><ring_r value="5.0"/>'
             '<bar_amp value="0.023"/><bar_a value="5.0"/><bar_b value="3.0"/>'
             '<bar_phi0 value="45.0"/><arm_num value="4"/><arm_r0 value="7.1"/>'
             '<arm_z0 value="0.1"/><arm_phi1 value="70"/><arm_phi2 value="160"/>'
             '<arm_phi3 value="250"/><arm_phi4 value="340"/><arm_amp1 value="2"/>'
             '<arm_amp2 value="0.133"/><arm_amp3 value="-3.78"/><arm_amp4 value="0.32"/>'
             '<arm_pitch value="11.5"/><comp_c value="0.5"/><comp_d value="0.3"/>'
             '<comp_r value="12"/><comp_p value="3"/></jaffe>')
    for name, block in (("jf12", jf12), ("helix", helix), ("jaffe", jaffe)):
        out = str(tmp_path / name) + "_"
        text = configs.plain_forward(nside_sim=8, nside_out=8, oc_r_res=0.05, out=out)
        text = text.replace('type="lsa"', f'type="{name}"').replace('<lsa>', block + '<lsa>')
        x = tmp_path / f"{name}.xml"
        x.write_text(text)
        r = run(str(x))
        assert r.returncode == 0, r.stderr
        py = Pipeline(text=text, ctx=ctx, export=False).run()
        ref = write_grid(Param(text=text).los_desc())
        fd = np.fromfile(out + "fd.bin", dtype="<f8")
        assert np.array_equal(fd, py[("fd", "nan", "8", "nan")]), name
        assert rel_l2(fd, ref["fd"] * CGS.m * CGS.m) < 1e-10, name
        for c, k in zip("IQU", ("is", "qs", "us")):
            got = np.fromfile(f"{out}sync_23.0_{c}.bin", dtype="<f8")
            assert np.array_equal(got, py[("sync", "23.0", "8", c)]), (name, c)
            assert rel_l2(got, ref[k]) < 1e-10, (name, c)


@pytest.mark.gpu
def test_cli_ymw16_matches_python_pipeline_and_oracle(tmp_path, ctx):
    """<thermalelectron><regular type="ymw16">: the C++ host and the Python
    pipeline read the reference's block (param.cc:480-641) and give the
    oracle's maps."""
    need_cli()
    out = str(tmp_path) + "/"
    text = configs.plain_forward(nside_sim=8, nside_out=8, oc_r_res=0.05, out=out,
                                 tereg="ymw16")
    x = tmp_path / "y.xml"
    x.write_text(text)
    r = run(str(x))
    assert r.returncode == 0, r.stderr
    py = Pipeline(text=text, ctx=ctx, export=False).run()
    ref = write_grid(Param(text=text).los_desc())
    dm = np.fromfile(out + "dm.bin", dtype="<f8")
    assert np.array_equal(dm, py[("dm", "nan", "8", "nan")])
    assert rel_l2(dm, ref["dm"] * CGS.ccm / CGS.pc) < 1e-10
    q = np.fromfile(out + "sync_23.0_Q.bin", dtype="<f8")
    assert rel_l2(q, ref["qs"]) < 1e-10


@pytest.mark.gpu
def test_cli_random_mhd_matches_python_pipeline_and_oracle(tmp_path, ctx):
    """<random type="local"><mhd>: the C++ host and the Python pipeline read
    the reference's block (param.cc:419-440), write the same Brnd_mhd grid,
    and that grid is the oracle's."""
    import ctypes as C
    from oracle_lib import load_oracle, ptr
    need_cli()
    out = str(tmp_path) + "/"
    text = configs.random_forward(nside_sim=4, nside_out=16, oc_r_res=0.1, grid=(32, 32, 16),
                                  te_k0=0.3, freqs=[23.0], out=out, brnd_type="local")
    bfile = str(tmp_path / "brnd.bin")
    text = text.replace('<brnd read="0" write="0" filename="brnd.bin"/>',
                        f'<brnd read="0" write="1" filename="{bfile}"/>')
    x = tmp_path / "p.xml"
    x.write_text(text)
    r = run(str(x))
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(bfile, dtype="<f8").reshape(-1, 3)
    pipe = Pipeline(text=text, ctx=ctx, export=False)
    py = pipe.run()
    for i in range(3):
        assert np.array_equal(raw[:, i], pipe.brnd[i].cpu().numpy()), i
    q = np.fromfile(str(tmp_path / "sync_23.0_Q.bin"), dtype="<f8")
    assert np.array_equal(q, py[("sync", "23.0", "16", "Q")])
    par = Param(text=text)
    n = par.grid_brnd.full_size
    ref = [np.zeros(n) for _ in range(3)]
    assert load_oracle().orc_brnd_mhd(C.byref(par.mhd_desc()), par.brnd_seed,
                                      *(ptr(a) for a in ref)) == 0
    for i in range(3):
        assert np.linalg.norm(ref[i]) > 0
        assert rel_l2(raw[:, i], ref[i]) < 1e-11, i


@pytest.mark.gpu
def test_cli_random_jf12_matches_python_pipeline_and_oracle(tmp_path, ctx):
    """<random><global type="jf12">: the C++ host and the Python pipeline read
    the reference's block (param.cc:375-409), write the same Brnd_jf12 grid,
    and that grid is the oracle's."""
    import ctypes as C
    from oracle_lib import load_oracle, ptr
    need_cli()
    out = str(tmp_path) + "/"
    text = configs.random_forward(nside_sim=4, nside_out=16, oc_r_res=0.1, grid=(32, 32, 16),
                                  te_k0=0.3, freqs=[23.0], out=out)
    es = text[text.index('<global type="es">'):text.index("</es>") + 5]
    jf = ('<global type="jf12"><jf12><rms value="1.0"/><k0 value="0.5"/><k1 value="0.1"/>'
          '<a0 value="1.7"/><a1 value="0.0"/><rho value="0.7"/>'
          '<b0_1 value="10.81"/><b0_2 value="6.96"/><b0_3 value="9.59"/><b0_4 value="6.96"/>'
          '<b0_5 value="1.96"/><b0_6 value="16.34"/><b0_7 value="37.29"/><b0_8 value="10.35"/>'
          '<b0_int value="7.63"/><z0_spiral value="0.61"/><b0_halo value="4.68"/>'
          '<r0_halo value="10.97"/><z0_halo value="2.84"/></jf12>')
    text = text.replace(es, jf)
    bfile = str(tmp_path / "brnd.bin")
    text = text.replace('<brnd read="0" write="0" filename="brnd.bin"/>',
                        f'<brnd read="0" write="1" filename="{bfile}"/>')
    x = tmp_path / "p.xml"
    x.write_text(text)
    r = run(str(x))
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(bfile, dtype="<f8").reshape(-1, 3)
    pipe = Pipeline(text=text, ctx=ctx, export=False)
    py = pipe.run()
    for i in range(3):
        assert np.array_equal(raw[:, i], pipe.brnd[i].cpu().numpy()), i
    q = np.fromfile(str(tmp_path / "sync_23.0_Q.bin"), dtype="<f8")
    assert np.array_equal(q, py[("sync", "23.0", "16", "Q")])
    par = Param(text=text)
    n = par.grid_brnd.full_size
    ref = [np.zeros(n) for _ in range(3)]
    assert load_oracle().orc_brnd_jf12(C.byref(par.es_desc()), par.brnd_seed,
                                       *(ptr(a) for a in ref)) == 0
    for i in range(3):
        assert rel_l2(raw[:, i], ref[i]) < 1e-11, i
