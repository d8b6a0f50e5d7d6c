"""Pipeline: the reference's run sequence on the device path.

Mirrors ``Pipeline`` (include/pipeline.h:14-45, source/pipeline/
pipeline.cc:19-190) as driven by ``main`` (main_std.cc:31-38):

  assemble_grid -> assemble_tereg -> assemble_breg -> assemble_ternd
  -> assemble_brnd -> assemble_cre -> assemble_obs

* random grids are generated on the GPU (Brnd_es / TErnd_dft) or imported
  from the reference's binary grid files, and exported in the same format
  (grid_brnd.cc:57-98, grid_ternd.cc:42-75); they stay resident on the device
  between passes;
* assemble_obs runs one write_grid pass per <sync> entry, last-listed first,
  switching DM and FD off after the first pass (pipeline.cc:165-190), and
  exports each pass's maps as raw FP64 RING arrays with DM in pc cm^-3 and FD
  in rad m^-2 (grid_obs.cc:53-83).
"""
from __future__ import annotations

import os
from typing import Optional

import numpy as np

from . import _capi as capi
from .param import Param
from .runtime import Context
from .units import CGS


def _npix(nside: int) -> int:
    return 12 * nside * nside


def export_map(path: str, values: np.ndarray) -> None:
    """Hamio::dump of a Hampix map (hamio.h:60-74): raw FP64, no header."""
    np.ascontiguousarray(values, dtype="<f8").tofile(path)


def sync_names(filename: str) -> tuple[str, str, str]:
    """_I/_Q/_U inserted before the last 4 characters (grid_obs.cc:63-68)."""
    return tuple(filename[:-4] + tag + filename[-4:] for tag in ("_I", "_Q", "_U"))


def import_grid(path: str, n: int, components: int) -> np.ndarray:
    """Grid_brnd / Grid_ternd::import_grid (and the mask map): exactly
    n*components doubles."""
    size = os.path.getsize(path)
    if size != n * components * 8:
        raise RuntimeError(f"grid file {path} has {size} bytes, expected "
                           f"{n * components * 8}")
    return np.fromfile(path, dtype="<f8")


class Pipeline:
    def __init__(self, filename: Optional[str] = None, text: Optional[str] = None,
                 ctx: Optional[Context] = None, export: bool = True):
        self.par = Param(filename, text)
        self.ctx = ctx if ctx is not None else Context(0)
        self._own_ctx = ctx is None
        self.export = export
        self.brnd = None   # device tensors (bx, by, bz) or None
        self.ternd = None  # device tensor or None
        self.maps: dict = {}  # (kind, freq GHz str, nside str, comp) -> np.ndarray

    def close(self) -> None:
        if self._own_ctx and self.ctx is not None:
            self.ctx.close()
        self.ctx = None

    # ----------------------------------------------------------- assemble
    def assemble_grid(self) -> None:
        import torch  # device memory for resident grids
        self._torch = torch
        self._dev = torch.device("cuda", 0)

    def load_mask(self) -> None:
        """Grid_obs::build_grid (grid_obs.cc:44-50): the mask map, raw FP64 at
        nside_mask, when <mask> is given."""
        g = self.par.grid_obs
        self.mask = None
        if g.do_mask:
            n = 12 * g.nside_mask * g.nside_mask
            self.mask = self._torch.from_numpy(import_grid(g.mask_name, n, 1)).to(self._dev)

    def _zeros(self, n: int):
        return self._torch.zeros(n, dtype=self._torch.float64, device=self._dev)

    def assemble_tereg(self) -> None:
        g = self.par.grid_tereg
        if g.build_permission and g.read_permission:
            self.tereg_grid = self._torch.from_numpy(
                import_grid(g.filename, g.full_size, 1)).to(self._dev)

    def assemble_breg(self) -> None:
        g = self.par.grid_breg
        if g.build_permission and g.read_permission:
            raw = import_grid(g.filename, g.full_size, 3).reshape(-1, 3)
            self.breg_grid = tuple(self._torch.from_numpy(np.ascontiguousarray(raw[:, i]))
                                   .to(self._dev) for i in range(3))

    def assemble_ternd(self) -> None:
        """pipeline.cc:87-110; the generator reads brnd_seed (ternd_dft.cc:44)."""
        g = self.par.grid_ternd
        if g.read_permission:
            self.ternd = self._torch.from_numpy(
                import_grid(g.filename, g.full_size, 1)).to(self._dev)
        elif g.build_permission:
            self.ternd = self._zeros(g.full_size)
            self.ctx.ternd_dft(self.par.dft_desc(), self.par.brnd_seed, self.ternd)
        if g.write_permission and self.ternd is not None:
            export_map(g.filename, self.ternd.cpu().numpy())

    def assemble_brnd(self) -> None:
        """pipeline.cc:112-143"""
        g = self.par.grid_brnd
        if g.read_permission:
            raw = import_grid(g.filename, g.full_size, 3).reshape(-1, 3)
            self.brnd = tuple(self._torch.from_numpy(np.ascontiguousarray(raw[:, i]))
                              .to(self._dev) for i in range(3))
        elif g.build_permission:
            kind = (self.par.brnd_type, self.par.brnd_method)
            if kind not in (("global", "es"), ("global", "jf12"), ("local", "mhd")):
                raise RuntimeError(f"random B '{self.par.brnd_type}/{self.par.brnd_method}'"
                                   " has no device generator")
            self.brnd = tuple(self._zeros(g.full_size) for _ in range(3))
            breg = getattr(self, "breg_grid", None)
            if kind[1] == "mhd":
                self.ctx.brnd_mhd(self.par.mhd_desc(), self.par.brnd_seed, *self.brnd,
                                  breg_grid=breg)
            else:
                gen = self.ctx.brnd_jf12 if kind[1] == "jf12" else self.ctx.brnd_es
                gen(self.par.es_desc(), self.par.brnd_seed, *self.brnd, breg_grid=breg)
        if g.write_permission and self.brnd is not None:
            inter = np.stack([a.cpu().numpy() for a in self.brnd], axis=1)
            export_map(g.filename, inter.ravel())

    def assemble_cre(self) -> None:
        """pipeline.cc:145-162: an imported CRE flux grid (CRE_num), or the
        analytic / uniform model, whose grid is written when asked
        (CREfield::write_grid, cre.cc:204-225)."""
        g = self.par.grid_cre
        self.cre_grid = None
        if g.read_permission:
            self.cre_grid = self._torch.from_numpy(
                import_grid(g.filename, g.full_size * g.nE, 1)).to(self._dev)
        elif g.write_permission:
            if not g.build_permission:
                raise RuntimeError("CRE grid output needs a CRE model")
            out = self._zeros(g.full_size * g.nE)
            self.ctx.cre_write_grid(self.par.cre_desc(), g.box(), g.E_min, g.E_fact, g.nE, out)
            export_map(g.filename, out.cpu().numpy())

    def _fields(self) -> dict:
        kw = {}
        if self.brnd is not None:
            kw.update(brnd=self.brnd, brnd_box=self.par.grid_brnd.box())
        if self.ternd is not None:
            kw.update(ternd=self.ternd, ternd_box=self.par.grid_ternd.box())
        if getattr(self, "breg_grid", None) is not None:
            kw.update(breg=self.breg_grid, breg_box=self.par.grid_breg.box())
        if getattr(self, "tereg_grid", None) is not None:
            kw.update(tereg=self.tereg_grid, tereg_box=self.par.grid_tereg.box())
        if getattr(self, "mask", None) is not None:
            kw.update(mask=self.mask, nside_mask=self.par.grid_obs.nside_mask)
        if getattr(self, "cre_grid", None) is not None:
            g = self.par.grid_cre
            kw.update(cre=self.cre_grid, cre_box=g.box(), cre_E=(g.E_min, g.E_fact, g.nE))
        return kw

    def assemble_obs(self) -> None:
        """pipeline.cc:165-190: one pass per sync entry, last listed first."""
        g = self.par.grid_obs
        if not g.write_permission:
            return
        repeat = len(g.do_sync)
        for i in range(repeat):
            if i > 0:
                g.do_dm = False
                g.do_fd = False
            d = self.par.los_desc()
            maps = {}
            if d.do_dm:
                maps["dm"] = self._zeros(_npix(d.nside_dm))
            if d.do_sync:
                for k in ("is", "qs", "us"):
                    maps[k] = self._zeros(_npix(d.nside_sync))
            if d.do_fd or d.do_sync:
                maps["fd"] = self._zeros(_npix(d.nside_fd if d.do_fd else d.nside_sync))
            self.ctx.los_forward(d, maps, **self._fields())
            self._collect(d, maps)
            if g.do_sync[-1]:
                g.nside_sync.pop()
                g.do_sync.pop()
                g.sim_sync_freq.pop()
                g.sim_sync_freq_str.pop()
                g.sim_sync_name.pop()

    def _collect(self, d, maps) -> None:
        """Grid_obs::export_grid (grid_obs.cc:53-83) and the in-memory copy."""
        g = self.par.grid_obs
        if d.do_dm:
            dm = maps["dm"].cpu().numpy() * (CGS.ccm / CGS.pc)
            self.maps[("dm", "nan", str(d.nside_dm), "nan")] = dm
            if self.export:
                export_map(g.sim_dm_name, dm)
        if d.do_sync:
            key = ("sync", g.sim_sync_freq_str[-1], str(d.nside_sync))
            names = sync_names(g.sim_sync_name[-1])
            for comp, k, name in zip("IQU", ("is", "qs", "us"), names):
                v = maps[k].cpu().numpy()
                self.maps[key + (comp,)] = v
                if self.export:
                    export_map(name, v)
        if d.do_fd:
            fd = maps["fd"].cpu().numpy() * (CGS.m * CGS.m)
            self.maps[("fd", "nan", str(d.nside_fd), "nan")] = fd
            if self.export:
                export_map(g.sim_fd_name, fd)

    def run(self) -> dict:
        """main_std.cc:31-38"""
        self.assemble_grid()
        self.assemble_tereg()
        self.assemble_breg()
        self.assemble_ternd()
        self.assemble_brnd()
        self.assemble_cre()
        self.load_mask()
        self.assemble_obs()
        return self.maps
