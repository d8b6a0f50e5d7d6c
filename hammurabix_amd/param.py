"""Param: the XML parameter set of hammurabiX, parsed into CGS units.

Mirrors the reference's ``Param`` (include/param.h:13-248, source/param/
param.cc:12-778): the same XML schema (templates/params_template.xml), the
same attribute defaults (missing numeric attribute -> 0, missing bool ->
false, as tinyxml2 returns), the same unit conversions, and the same
``std::runtime_error`` messages for unsupported model strings.  It then
builds the POD descriptors of the C-ABI for the device path.
"""
from __future__ import annotations

import math
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import Optional

from . import _capi as capi
from .units import CGS

# --------------------------------------------------------------- XML access


class _El:
    """Thin tinyxml2-like view: missing attributes read as 0 / false / ''."""

    def __init__(self, el: Optional[ET.Element]):
        self.el = el

    def child(self, key: str) -> "_El":
        if self.el is None:
            raise RuntimeError(f"missing XML element <{key}>")
        c = self.el.find(key)
        if c is None:
            raise RuntimeError(f"missing XML element <{key}> under <{self.el.tag}>")
        return _El(c)

    def has(self, key: str) -> bool:
        return self.el is not None and self.el.find(key) is not None

    def children(self, key: str) -> list["_El"]:
        return [] if self.el is None else [_El(c) for c in self.el.findall(key)]

    def s(self, att: str, key: Optional[str] = None) -> str:
        e = self.child(key).el if key else self.el
        v = e.get(att)
        if v is None:
            raise RuntimeError(f"missing attribute {att} on <{e.tag}>")
        return v

    def f(self, att: str, key: Optional[str] = None) -> float:
        e = self.child(key).el if key else self.el
        v = e.get(att)
        return float(v) if v is not None else 0.0

    def u(self, att: str, key: Optional[str] = None) -> int:
        e = self.child(key).el if key else self.el
        v = e.get(att)
        return int(v) if v is not None else 0

    def b(self, att: str, key: Optional[str] = None) -> bool:
        e = self.child(key).el if key else self.el
        v = e.get(att)
        if v is None:
            return False
        v = v.strip().lower()
        if v in ("1", "true"):
            return True
        if v in ("0", "false"):
            return False
        try:
            return int(v) != 0
        except ValueError:
            return False


def _trace(root: ET.Element, keys: list[str]) -> _El:
    e = _El(root)
    for k in keys:
        e = e.child(k)
    return e


# ----------------------------------------------------------------- records


@dataclass
class GridParam:
    filename: str = ""
    build_permission: bool = False
    read_permission: bool = False
    write_permission: bool = False
    nx: int = 0
    ny: int = 0
    nz: int = 0
    x_min: float = 0.0
    x_max: float = 0.0
    y_min: float = 0.0
    y_max: float = 0.0
    z_min: float = 0.0
    z_max: float = 0.0
    # energy axis of the CRE grid (param.cc:760-776)
    E_min: float = 0.0
    E_max: float = 0.0
    E_fact: float = 0.0
    nE: int = 0

    @property
    def full_size(self) -> int:
        return self.nx * self.ny * self.nz

    def box(self) -> capi.HamxBox:
        return capi.HamxBox(self.nx, self.ny, self.nz, self.x_min, self.x_max,
                            self.y_min, self.y_max, self.z_min, self.z_max)


@dataclass
class ObsParam:
    write_permission: bool = False
    nside_dm: int = 0
    nside_fd: int = 0
    nside_mask: int = 0
    total_shell: int = 0
    nside_sync: list = field(default_factory=list)
    nside_shell: list = field(default_factory=list)
    cut_shell: list = field(default_factory=list)
    radii_shell: list = field(default_factory=list)
    gc_r_min: float = 0.0
    gc_r_max: float = 0.0
    gc_z_min: float = 0.0
    gc_z_max: float = 0.0
    oc_r_min: float = 0.0
    oc_r_max: float = 0.0
    oc_r_res: float = 0.0
    do_dm: bool = False
    do_fd: bool = False
    do_sync: list = field(default_factory=list)
    sim_fd_name: str = ""
    sim_dm_name: str = ""
    sim_sync_name: list = field(default_factory=list)
    sim_sync_freq: list = field(default_factory=list)
    sim_sync_freq_str: list = field(default_factory=list)  # XML text, GHz
    do_mask: bool = False
    mask_name: str = ""


def _box(root: ET.Element, name: str, g: GridParam) -> None:
    b = _trace(root, ["grid", name])
    g.nx = b.u("value", "nx")
    g.ny = b.u("value", "ny")
    g.nz = b.u("value", "nz")
    g.x_max = CGS.kpc * b.f("value", "x_max")
    g.x_min = CGS.kpc * b.f("value", "x_min")
    g.y_max = CGS.kpc * b.f("value", "y_max")
    g.y_min = CGS.kpc * b.f("value", "y_min")
    g.z_max = CGS.kpc * b.f("value", "z_max")
    g.z_min = CGS.kpc * b.f("value", "z_min")


# Parameter order of HAMX_BREG_JF12 (include/hamx.h)
JF12_PARAMS = ("b_arm1", "b_arm2", "b_arm3", "b_arm4", "b_arm5", "b_arm6", "b_arm7", "b_ring",
               "h_disk", "w_disk", "Bn", "Bs", "rn", "rs", "wh", "z0", "B0_X", "Xtheta",
               "rpc_X", "r0_X")


# Parameter order of HAMX_BREG_JAFFE (include/hamx.h)
JAFFE_PARAMS = ("quadruple", "bss", "disk_amp", "disk_z0", "halo_amp", "halo_z0", "r_inner",
                "r_scale", "r_peak", "ring", "bar", "ring_amp", "ring_r", "bar_amp", "bar_a",
                "bar_b", "bar_phi0", "arm_num", "arm_r0", "arm_z0", "arm_phi1", "arm_phi2",
                "arm_phi3", "arm_phi4", "arm_amp1", "arm_amp2", "arm_amp3", "arm_amp4",
                "arm_pitch", "comp_c", "comp_d", "comp_r", "comp_p")


# Energy profile of Brnd_jf12 in hamx_es_desc.jf12 order, with the unit each XML
# value is given in (param.cc:375-409)
BRND_JF12_PARAMS = tuple((f"b0_{i}", CGS.muGauss) for i in range(1, 9)) + (
    ("b0_int", CGS.muGauss), ("z0_spiral", CGS.kpc), ("b0_halo", CGS.muGauss),
    ("r0_halo", CGS.kpc), ("z0_halo", CGS.kpc))


def ymw16_params(root: ET.Element) -> list[float]:
    """The <ymw16> block in HAMX_TEREG_YMW16 order and CGS units
    (param.cc:480-641): lengths in pc or kpc to cm, angles of the arms to
    radians and their pitch to its tangent and cosine, the other angles kept
    in degrees."""
    def blk(name):
        return _trace(root, ["thermalelectron", "regular", "ymw16", name])
    w, tk, tn = blk("warp"), blk("thickdisk"), blk("thindisk")
    sa, gc, gn = blk("spiralarm"), blk("galcenter"), blk("gumnebula")
    lb, li = blk("localbubble"), blk("loopi")
    pc, kpc, rad = CGS.pc, CGS.kpc, CGS.rad
    arms = range(5)
    pitch = [sa.f("value", f"pitch_arm_{i}") * rad for i in arms]
    p = [w.f("value", "r_warp") * kpc, w.f("value", "r0") * kpc, w.f("value", "gamma_w"),
         tk.f("value", "ad") * pc, tk.f("value", "bd") * pc, tk.f("value", "n1"),
         tk.f("value", "h1") * pc,
         tn.f("value", "a2") * pc, tn.f("value", "b2") * pc, tn.f("value", "n2"),
         tn.f("value", "k2"),
         sa.f("value", "b2s") * pc]
    p += [sa.f("value", f"ele_arm_{i}") for i in arms]
    p += [sa.f("value", f"wid_arm_{i}") * pc for i in arms]
    p += [sa.f("value", f"rref_arm_{i}") * kpc for i in arms]
    p += [sa.f("value", f"phiref_arm_{i}") * rad for i in arms]
    p += [math.tan(a) for a in pitch] + [math.cos(a) for a in pitch]
    p += [sa.f("value", "aa") * pc, sa.f("value", "ka"), sa.f("value", "ncn"),
          sa.f("value", "thetacn"), sa.f("value", "wcn"), sa.f("value", "nsg"),
          sa.f("value", "thetasg"), sa.f("value", "wsg"),
          gc.f("value", "ngc"), gc.f("value", "agc") * pc, gc.f("value", "hgc") * pc,
          gn.f("value", "ngn"), gn.f("value", "wgn") * pc, gn.f("value", "agn") * pc,
          gn.f("value", "kgn"),
          lb.f("value", "j_lb"),
          lb.f("value", "nlb1"), lb.f("value", "thetalb1"), lb.f("value", "detlb1"),
          lb.f("value", "wlb1") * pc, lb.f("value", "hlb1") * pc,
          lb.f("value", "nlb2"), lb.f("value", "thetalb2"), lb.f("value", "detlb2"),
          lb.f("value", "wlb2") * pc, lb.f("value", "hlb2") * pc,
          li.f("value", "nli"), li.f("value", "rli") * pc, li.f("value", "wli") * pc,
          li.f("value", "detthetali"), li.f("value", "thetali")]
    assert len(p) == 73
    return p


class Param:
    """Parsed parameter set (param.cc:12-28)."""

    def __init__(self, filename: Optional[str] = None, text: Optional[str] = None):
        self.observer = [0.0, 0.0, 0.0]
        self.grid_obs = ObsParam()
        self.grid_breg = GridParam()
        self.grid_brnd = GridParam()
        self.grid_tereg = GridParam()
        self.grid_ternd = GridParam()
        self.grid_cre = GridParam()
        self.breg_type = ""
        self.brnd_type = ""
        self.brnd_method = ""
        self.tereg_type = ""
        self.ternd_type = ""
        self.ternd_method = ""
        self.cre_type = ""
        self.brnd_seed = 0
        self.ternd_seed = 0
        self.breg_lsa: dict = {}
        self.breg_cart: dict = {}
        self.breg_unif: dict = {}
        self.brnd_es: dict = {}
        self.tereg_unif: dict = {}
        self.ternd_dft: dict = {}
        self.cre_ana: dict = {}
        self.cre_unif: dict = {}
        if filename is None and text is None:
            return
        root = ET.parse(filename).getroot() if text is None else ET.fromstring(text)
        if root.tag != "root":
            raise RuntimeError("parameter file has no <root> element")
        o = _trace(root, ["grid", "observer"])
        self.observer = [CGS.kpc * o.f("value", "x"), CGS.kpc * o.f("value", "y"),
                         CGS.kpc * o.f("value", "z")]
        self._obs(root)
        self._breg(root)
        self._brnd(root)
        self._tereg(root)
        self._ternd(root)
        self._cre(root)

    # param.cc:30-144
    def _obs(self, root: ET.Element) -> None:
        g = self.grid_obs
        ob = _trace(root, ["observable"])
        if ob.has("dm"):
            g.write_permission = True
            g.do_dm = ob.b("cue", "dm")
            g.sim_dm_name = ob.s("filename", "dm")
            g.nside_dm = ob.u("nside", "dm")
        if ob.has("faraday"):
            g.write_permission = True
            g.do_fd = ob.b("cue", "faraday")
            g.sim_fd_name = ob.s("filename", "faraday")
            g.nside_fd = ob.u("nside", "faraday")
        syncs = ob.children("sync")
        if syncs:
            g.write_permission = True
            for e in syncs:
                g.do_sync.append(e.b("cue"))
                g.sim_sync_freq.append(e.f("freq") * CGS.GHz)
                g.sim_sync_freq_str.append(e.el.get("freq", "0"))
                g.sim_sync_name.append(e.s("filename"))
                g.nside_sync.append(e.u("nside"))
        else:
            g.do_sync.append(False)
        if not g.write_permission:
            return
        sh = _trace(root, ["grid", "shell"])
        g.oc_r_min = sh.f("value", "oc_r_min") * CGS.kpc
        g.oc_r_max = sh.f("value", "oc_r_max") * CGS.kpc
        g.gc_r_min = sh.f("value", "gc_r_min") * CGS.kpc
        g.gc_r_max = sh.f("value", "gc_r_max") * CGS.kpc
        g.gc_z_min = sh.f("value", "gc_z_min") * CGS.kpc
        g.gc_z_max = sh.f("value", "gc_z_max") * CGS.kpc
        g.oc_r_res = sh.f("value", "oc_r_res") * CGS.kpc
        layer = sh.s("type", "layer")
        if layer == "auto":
            a = _trace(root, ["grid", "shell", "layer", "auto"])
            g.total_shell = a.u("value", "shell_num")
            nside_sim = a.u("value", "nside_sim")
            g.nside_shell = [int(2 ** i) * nside_sim for i in range(g.total_shell)]
            g.radii_shell = [g.oc_r_min]
            for i in range(g.total_shell):
                g.radii_shell.append((g.oc_r_max - g.oc_r_min) *
                                     math.pow(0.5, g.total_shell - i - 1) + g.oc_r_min)
        elif layer == "manual":
            m = _trace(root, ["grid", "shell", "layer", "manual"])
            g.nside_shell = [e.u("value") for e in m.children("nside_sim")]
            g.total_shell = len(g.nside_shell)
            g.cut_shell = [e.f("value") for e in m.children("cut")] + [1.0]
            if len(g.cut_shell) != g.total_shell:
                raise RuntimeError("manual layer: cut count must be nside_sim count - 1")
            g.radii_shell = [g.oc_r_min] + [(g.oc_r_max - g.oc_r_min) * c + g.oc_r_min
                                            for c in g.cut_shell]
        else:
            raise RuntimeError("unsupported layer option")
        mk = _trace(root, ["mask"])
        if mk.b("cue"):
            g.do_mask = True
            g.mask_name = mk.s("filename")
            g.nside_mask = mk.u("nside")

    def _fieldio(self, root: ET.Element, name: str, g: GridParam) -> None:
        fio = _trace(root, ["fieldio"])
        if fio.has(name):
            g.read_permission = fio.b("read", name)
            g.write_permission = fio.b("write", name)
            g.filename = fio.s("filename", name)

    # param.cc:146-343
    def _breg(self, root: ET.Element) -> None:
        g = self.grid_breg
        self._fieldio(root, "breg", g)
        mf = _trace(root, ["magneticfield"])
        g.build_permission = mf.b("cue", "regular")
        if g.build_permission and not g.read_permission:
            self.breg_type = mf.s("type", "regular")
            reg = ["magneticfield", "regular"]
            if self.breg_type == "lsa":
                s = _trace(root, reg + ["lsa"])
                self.breg_lsa = dict(b0=s.f("value", "b0") * CGS.muGauss,
                                     psi0=s.f("value", "psi0") * CGS.rad,
                                     psi1=s.f("value", "psi1") * CGS.rad,
                                     chi0=s.f("value", "chi0") * CGS.rad)
            elif self.breg_type == "cart":
                s = _trace(root, reg + ["cart"])
                self.breg_cart = dict(bx=s.f("value", "bx") * CGS.muGauss,
                                      by=s.f("value", "by") * CGS.muGauss,
                                      bz=s.f("value", "bz") * CGS.muGauss)
            elif self.breg_type == "unif":
                s = _trace(root, reg + ["unif"])
                self.breg_unif = dict(bp=s.f("value", "bp") * CGS.muGauss,
                                      bv=s.f("value", "bv") * CGS.muGauss,
                                      l0=s.f("value", "l0") * CGS.rad)
            elif self.breg_type == "helix":  # param.cc:258-270
                s = _trace(root, reg + ["helix"])
                self.breg_helix = dict(bx=s.f("value", "bx") * CGS.muGauss,
                                       by=s.f("value", "by") * CGS.muGauss,
                                       bz=s.f("value", "bz") * CGS.muGauss,
                                       r_min=s.f("value", "r_min") * CGS.kpc,
                                       r_max=s.f("value", "r_max") * CGS.kpc)
            elif self.breg_type == "jf12":  # param.cc:271-322
                s = _trace(root, reg + ["jf12"])
                mu = [f"b_arm{i}" for i in range(1, 8)] + ["b_ring"]
                self.breg_jf12 = {k: s.f("value", k) * CGS.muGauss for k in mu}
                for k in ("h_disk", "w_disk", "rn", "rs", "wh", "z0", "rpc_X", "r0_X"):
                    self.breg_jf12[k] = s.f("value", k) * CGS.kpc
                for k in ("Bn", "Bs", "B0_X"):
                    self.breg_jf12[k] = s.f("value", k) * CGS.muGauss
                self.breg_jf12["Xtheta"] = s.f("value", "Xtheta") * CGS.rad
            elif self.breg_type == "jaffe":  # param.cc:173-257
                s = _trace(root, reg + ["jaffe"])
                q = {"quadruple": float(s.b("cue", "quadruple")),
                     "bss": float(s.b("cue", "bss")),
                     "ring": float(s.b("cue", "ring")), "bar": float(s.b("cue", "bar"))}
                for k in ("disk_amp", "halo_amp", "ring_amp", "bar_amp"):
                    q[k] = s.f("value", k) * CGS.muGauss
                for k in ("disk_z0", "halo_z0", "r_inner", "r_scale", "r_peak", "ring_r",
                          "bar_a", "bar_b", "arm_r0", "arm_z0", "comp_d", "comp_r"):
                    q[k] = s.f("value", k) * CGS.kpc
                q["bar_phi0"] = s.f("value", "bar_phi0") * CGS.rad
                q["arm_num"] = float(s.u("value", "arm_num"))
                if q["arm_num"] > 4:
                    raise RuntimeError("jaffe: at most 4 spiral arms (arm_phi1..4)")
                for i in range(1, 5):
                    q[f"arm_phi{i}"] = s.f("value", f"arm_phi{i}") * CGS.rad
                    q[f"arm_amp{i}"] = s.f("value", f"arm_amp{i}") * CGS.muGauss
                q["arm_pitch"] = s.f("value", "arm_pitch") * CGS.rad
                q["comp_c"] = s.f("value", "comp_c")
                q["comp_p"] = s.f("value", "comp_p")
                self.breg_jaffe = q
            else:
                raise RuntimeError("unsupported breg model")
        if g.read_permission or g.write_permission:
            _box(root, "box_breg", g)

    # param.cc:345-460
    def _brnd(self, root: ET.Element) -> None:
        g = self.grid_brnd
        self._fieldio(root, "brnd", g)
        mf = _trace(root, ["magneticfield"])
        g.build_permission = mf.b("cue", "random")
        if g.build_permission and not g.read_permission:
            self.brnd_seed = mf.u("seed", "random")
            self.brnd_type = mf.s("type", "random")
            if self.brnd_type == "global":
                gl = _trace(root, ["magneticfield", "random", "global"])
                self.brnd_method = gl.s("type")
                if self.brnd_method == "es":
                    s = _trace(root, ["magneticfield", "random", "global", "es"])
                    self.brnd_es = dict(rms=s.f("value", "rms") * CGS.muGauss,
                                        k0=s.f("value", "k0"), a0=s.f("value", "a0"),
                                        k1=s.f("value", "k1"), a1=s.f("value", "a1"),
                                        rho=s.f("value", "rho"),
                                        r0=s.f("value", "r0") * CGS.kpc,
                                        z0=s.f("value", "z0") * CGS.kpc)
                elif self.brnd_method == "jf12":
                    s = _trace(root, ["magneticfield", "random", "global", "jf12"])
                    self.brnd_es = dict(rms=s.f("value", "rms") * CGS.muGauss,
                                        k0=s.f("value", "k0"), a0=s.f("value", "a0"),
                                        k1=s.f("value", "k1"), a1=s.f("value", "a1"),
                                        rho=s.f("value", "rho"), r0=0.0, z0=0.0)
                    self.brnd_jf12 = [s.f("value", k) * u for k, u in BRND_JF12_PARAMS]
                elif self.brnd_method == "jaffe":
                    _trace(root, ["magneticfield", "random", "global", "jaffe"])
                else:
                    raise RuntimeError("unsupported brnd model")
            elif self.brnd_type == "local":
                lc = _trace(root, ["magneticfield", "random", "local"])
                self.brnd_method = lc.s("type")
                if self.brnd_method != "mhd":
                    raise RuntimeError("unsupported brnd model")
                s = _trace(root, ["magneticfield", "random", "local", "mhd"])
                mu2 = CGS.muGauss * CGS.muGauss
                self.brnd_mhd = dict(pa0=s.f("value", "pa0") * mu2,
                                     pf0=s.f("value", "pf0") * mu2,
                                     ps0=s.f("value", "ps0") * mu2,
                                     k0=s.f("value", "k0"), aa0=s.f("value", "aa0"),
                                     af0=s.f("value", "af0"), as0=s.f("value", "as0"),
                                     k1=s.f("value", "k1"), a1=s.f("value", "a1"),
                                     beta=s.f("value", "beta"), ma=s.f("value", "ma"))
            else:
                raise RuntimeError("unsupported brnd type")
        if g.read_permission or g.write_permission or g.build_permission:
            _box(root, "box_brnd", g)

    # param.cc:462-667
    def _tereg(self, root: ET.Element) -> None:
        g = self.grid_tereg
        self._fieldio(root, "tereg", g)
        te = _trace(root, ["thermalelectron"])
        g.build_permission = te.b("cue", "regular")
        if g.build_permission:
            self.tereg_type = te.s("type", "regular")
            if self.tereg_type == "unif":
                s = _trace(root, ["thermalelectron", "regular", "unif"])
                self.tereg_unif = dict(n0=s.f("value", "n0"),
                                       r0=s.f("value", "r0") * CGS.kpc)
            elif self.tereg_type == "ymw16":
                self.tereg_ymw16 = ymw16_params(root)
            else:
                raise RuntimeError("unsupported tereg model")
        if g.read_permission or g.write_permission:
            _box(root, "box_tereg", g)

    # param.cc:669-716
    def _ternd(self, root: ET.Element) -> None:
        g = self.grid_ternd
        self._fieldio(root, "ternd", g)
        te = _trace(root, ["thermalelectron"])
        g.build_permission = te.b("cue", "random")
        if g.build_permission:
            self.ternd_seed = te.u("seed", "random")
            self.ternd_type = te.s("type", "random")
            if self.ternd_type == "global":
                gl = _trace(root, ["thermalelectron", "random", "global"])
                self.ternd_method = gl.s("type")
                # "dgm" is accepted as another name of the DFT generator
                if self.ternd_method in ("dft", "dgm"):
                    s = _trace(root, ["thermalelectron", "random", "global", "dft"])
                    self.ternd_method = "dft"
                    self.ternd_dft = dict(rms=s.f("value", "rms"), k0=s.f("value", "k0"),
                                          a0=s.f("value", "a0"),
                                          r0=s.f("value", "r0") * CGS.kpc,
                                          z0=s.f("value", "z0") * CGS.kpc)
                else:
                    raise RuntimeError("unsupported ternd model")
            else:
                raise RuntimeError("unsupported ternd type")
        if g.read_permission or g.write_permission or g.build_permission:
            _box(root, "box_ternd", g)

    # param.cc:718-778
    def _cre(self, root: ET.Element) -> None:
        g = self.grid_cre
        self._fieldio(root, "cre", g)
        cr = _trace(root, ["cre"])
        g.build_permission = cr.b("cue")
        if g.build_permission:
            self.cre_type = cr.s("type")
            if self.cre_type == "analytic":
                s = _trace(root, ["cre", "analytic"])
                self.cre_ana = dict(alpha=s.f("value", "alpha"), beta=s.f("value", "beta"),
                                    theta=s.f("value", "theta"),
                                    r0=s.f("value", "r0") * CGS.kpc,
                                    z0=s.f("value", "z0") * CGS.kpc,
                                    E0=s.f("value", "E0") * CGS.GeV,
                                    j0=s.f("value", "j0"))
            elif self.cre_type == "unif":
                s = _trace(root, ["cre", "unif"])
                self.cre_unif = dict(alpha=s.f("value", "alpha"),
                                     r0=s.f("value", "r0") * CGS.kpc,
                                     E0=s.f("value", "E0") * CGS.GeV,
                                     j0=s.f("value", "j0"))
            else:
                raise RuntimeError("unsupported cre model")
        if g.read_permission or g.write_permission:
            b = _trace(root, ["grid", "box_cre"])
            g.E_min = CGS.GeV * b.f("value", "E_min")
            g.E_max = CGS.GeV * b.f("value", "E_max")
            g.nE = b.u("value", "nE")
            g.E_fact = math.log(g.E_max / g.E_min) / (g.nE - 1)
            _box(root, "box_cre", g)

    def cre_desc(self) -> capi.HamxLosDesc:
        """The observer and the analytic / uniform CRE model of the parameter
        set, as hamx_cre_write_grid reads them (CREfield::write_grid)."""
        d = capi.HamxLosDesc()
        d.observer[:] = self.observer
        self._cre_model(d)
        return d

    def _cre_model(self, d: capi.HamxLosDesc) -> None:
        if not self.grid_cre.build_permission:
            d.cre_model = capi.HAMX_CRE_NONE
        elif self.cre_type == "analytic":
            d.cre_model = capi.HAMX_CRE_ANA
            p = self.cre_ana
            d.cre[0:7] = [p["alpha"], p["beta"], p["theta"], p["r0"], p["z0"], p["E0"],
                          p["j0"]]
        elif self.cre_type == "unif":
            d.cre_model = capi.HAMX_CRE_UNIF
            p = self.cre_unif
            d.cre[0:4] = [p["alpha"], p["r0"], p["E0"], p["j0"]]

    # ----------------------------------------------------------- descriptors
    def los_desc(self, schedule: int = capi.HAMX_SCHED_AUTO) -> capi.HamxLosDesc:
        """Descriptor of one write_grid pass for the sync entry that
        grid_obs.*.back() names (integrator.cc:44,105,164)."""
        g = self.grid_obs
        d = capi.HamxLosDesc()
        d.observer[:] = self.observer
        if g.total_shell > capi.HAMX_MAX_SHELL:
            raise RuntimeError(f"at most {capi.HAMX_MAX_SHELL} shells are supported")
        d.n_shell = g.total_shell
        for i, n in enumerate(g.nside_shell):
            d.nside_shell[i] = n
        for i, r in enumerate(g.radii_shell):
            d.radii_shell[i] = r
        d.oc_r_res = g.oc_r_res
        d.gc_r_min, d.gc_r_max = g.gc_r_min, g.gc_r_max
        d.gc_z_min, d.gc_z_max = g.gc_z_min, g.gc_z_max
        d.do_dm, d.do_fd = int(g.do_dm), int(g.do_fd)
        d.do_sync = int(bool(g.do_sync) and g.do_sync[-1])
        d.nside_dm, d.nside_fd = g.nside_dm, g.nside_fd
        if d.do_sync:
            d.nside_sync = g.nside_sync[-1]
            d.sync_freq = g.sim_sync_freq[-1]
        # regular B
        if not self.grid_breg.build_permission:
            d.breg_model = capi.HAMX_BREG_NONE
        elif self.grid_breg.read_permission:
            d.breg_model = capi.HAMX_BREG_GRID
        elif self.breg_type == "lsa":
            d.breg_model = capi.HAMX_BREG_LSA
            p = self.breg_lsa
            d.breg[0:4] = [p["b0"], p["psi0"], p["psi1"], p["chi0"]]
        elif self.breg_type == "cart":
            d.breg_model = capi.HAMX_BREG_CART
            p = self.breg_cart
            d.breg[0:3] = [p["bx"], p["by"], p["bz"]]
        elif self.breg_type == "unif":
            d.breg_model = capi.HAMX_BREG_UNIF
            p = self.breg_unif
            d.breg[0:3] = [p["bp"], p["bv"], p["l0"]]
        elif self.breg_type == "helix":
            d.breg_model = capi.HAMX_BREG_HELIX
            p = self.breg_helix
            d.breg[0:5] = [p["bx"], p["by"], p["bz"], p["r_min"], p["r_max"]]
        elif self.breg_type == "jf12":
            d.breg_model = capi.HAMX_BREG_JF12
            d.breg[0:20] = [self.breg_jf12[k] for k in JF12_PARAMS]
        elif self.breg_type == "jaffe":
            d.breg_model = capi.HAMX_BREG_JAFFE
            d.breg[0:33] = [self.breg_jaffe[k] for k in JAFFE_PARAMS]
        else:
            raise RuntimeError(f"breg model '{self.breg_type}' has no device evaluator")
        # regular TE
        if not self.grid_tereg.build_permission:
            d.tereg_model = capi.HAMX_TEREG_NONE
        elif self.grid_tereg.read_permission:
            d.tereg_model = capi.HAMX_TEREG_GRID
        elif self.tereg_type == "unif":
            d.tereg_model = capi.HAMX_TEREG_UNIF
            d.tereg[0:2] = [self.tereg_unif["n0"], self.tereg_unif["r0"]]
        elif self.tereg_type == "ymw16":
            d.tereg_model = capi.HAMX_TEREG_YMW16
            d.tereg[0:73] = self.tereg_ymw16
        else:
            raise RuntimeError(f"tereg model '{self.tereg_type}' has no device evaluator")
        # CRE: an imported grid takes precedence (integrator.cc:287,356)
        if self.grid_cre.read_permission:
            d.cre_model = capi.HAMX_CRE_GRID
        else:
            self._cre_model(d)
        d.schedule = schedule
        return d

    def es_desc(self) -> capi.HamxEsDesc:
        """Descriptor of Brnd_es::write_grid on box_brnd (brnd_es.cc:76-308)."""
        p = self.brnd_es
        d = capi.HamxEsDesc()
        d.box = self.grid_brnd.box()
        d.observer[:] = self.observer
        d.rms, d.k0, d.k1 = p["rms"], p["k0"], p["k1"]
        d.a0, d.a1, d.rho = p["a0"], p["a1"], p["rho"]
        d.r0, d.z0 = p["r0"], p["z0"]
        self._generator_breg(d)
        if self.brnd_method == "jf12":
            d.jf12[:] = self.brnd_jf12
        return d

    def _generator_breg(self, d) -> None:
        """The regular field a random generator reads (anisotropy direction,
        the field at the observer): the LOS descriptor's model."""
        ld = self.los_desc() if self.grid_obs.write_permission else None
        if ld is not None:
            d.breg_model = ld.breg_model
            d.breg[:] = ld.breg[:]
        elif self.breg_type == "lsa" and self.grid_breg.build_permission:
            d.breg_model = capi.HAMX_BREG_LSA
            q = self.breg_lsa
            d.breg[0:4] = [q["b0"], q["psi0"], q["psi1"], q["chi0"]]
        if d.breg_model == capi.HAMX_BREG_GRID:
            d.breg_box = self.grid_breg.box()

    def mhd_desc(self) -> capi.HamxMhdDesc:
        """Descriptor of Brnd_mhd::write_grid on box_brnd (brnd_mhd.cc:16-167)."""
        d = capi.HamxMhdDesc()
        d.box = self.grid_brnd.box()
        d.observer[:] = self.observer
        for k, v in self.brnd_mhd.items():
            setattr(d, k, v)
        self._generator_breg(d)
        return d

    def dft_desc(self) -> capi.HamxDftDesc:
        """Descriptor of TErnd_dft::write_grid on box_ternd (ternd_dft.cc:35-128)."""
        p = self.ternd_dft
        d = capi.HamxDftDesc()
        d.box = self.grid_ternd.box()
        d.observer[:] = self.observer
        d.rms, d.k0, d.a0 = p["rms"], p["k0"], p["a0"]
        d.r0, d.z0 = p["r0"], p["z0"]
        return d
