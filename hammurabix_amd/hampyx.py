"""hampyx: Python front end of the simulator, run in-process on the GPU.

Mirrors the reference's ``hampyx.Hampyx`` interface (hampyx/hampyx.py):

* ``Hampyx(xml_path=None, exe_path=None)`` loads a parameter file (the first
  ``*xml*`` file of the working directory when none is given);
* ``mod_par`` / ``add_par`` / ``del_par`` / ``print_par`` edit the parameter
  tree with the same key-chain arguments and ``ValueError`` messages;
* calling the object runs the simulation and fills ``sim_map`` with the same
  keys: ``('sync', freq, nside, 'I'|'Q'|'U')``, ``('fd', 'nan', nside,
  'nan')``, ``('dm', 'nan', nside, 'nan')``, freq and nside being the XML
  attribute strings.

The reference launches the ``hamx`` executable as a subprocess and reads the
maps back from temporary files; here the parameter tree is handed to the
device pipeline in memory and the maps come back as numpy arrays.
``exe_path`` is accepted for call compatibility and not used.

``differentiable()`` (no counterpart in the reference) runs the same passes
through the adjoint-backed ``torch.autograd.Function`` of ``autograd.py``
and returns the ``sim_map`` entries as float64 device tensors: a backward
pass through any scalar of them reaches the field parameters and the random
grids across every ``<sync>`` pass.
"""
from __future__ import annotations

import os
import xml.etree.ElementTree as et
from typing import Optional

from .pipeline import Pipeline
from .runtime import Context
from .units import CGS


class Hampyx:
    def __init__(self, xml_path: Optional[str] = None, exe_path: Optional[str] = None,
                 ctx: Optional[Context] = None):
        self.wk_dir = os.getcwd()
        self.exe_path = exe_path
        if xml_path is None:
            cands = [s for s in sorted(os.listdir(self.wk_dir)) if "xml" in s]
            cands = [c for c in cands if os.path.isfile(os.path.join(self.wk_dir, c))]
            if not cands:
                raise ValueError("no XML parameter file in the working directory")
            xml_path = os.path.join(self.wk_dir, cands[0])
        if not os.path.isfile(xml_path):
            raise ValueError(f"parameter file {xml_path} not found")
        self.xml_path = os.path.abspath(xml_path)
        self.tree = et.parse(self.xml_path)
        self.sim_map: dict = {}
        self.sim_map_name: dict = {}
        self._ctx = ctx

    def __call__(self, verbose: bool = False) -> None:
        return self.run(verbose)

    def run(self, verbose: bool = False) -> None:
        """Run the pipeline on the current parameter tree."""
        text = et.tostring(self.tree.getroot(), encoding="unicode")
        ctx = self._ctx if self._ctx is not None else Context(0)
        try:
            pipe = Pipeline(text=text, ctx=ctx, export=False)
            maps = pipe.run()
        finally:
            if self._ctx is None:
                ctx.close()
        if verbose:
            for k, v in maps.items():
                print(k, v.shape)
        self.sim_map.update(maps)

    def differentiable(self, params=None, grids_grad: bool = False):
        """The pipeline's passes (pipeline.cc:165-190: the last-listed <sync>
        entry first with DM and FD, the others sync only) as differentiable
        maps.  Returns (maps, params, grids): maps has the sim_map keys and
        units (DM in pc cm^-3, FD in rad m^-2) as float64 device tensors;
        params is the float64 tensor of (b0, psi0, psi1, chi0, alpha, n0) in
        the descriptor's CGS / radian units that every pass reads (created with
        requires_grad from the parameter file when not given); grids is
        (bx, by, bz, te) of the random fields on the device, requiring grad
        when grids_grad.  Models: any regular B and TE model (the gradients
        of b0, psi0, psi1, chi0 are zero unless B is LSA, n0's unless TE is
        the uniform sphere), a mask, the analytic or uniform CRE with a
        constant index, random grids on one box."""
        from .autograd import desc_params, los_maps
        text = et.tostring(self.tree.getroot(), encoding="unicode")
        if self._ctx is None:
            self._ctx = Context(0)  # kept open: the backward pass runs on it
        ctx = self._ctx
        pipe = Pipeline(text=text, ctx=ctx, export=False)
        pipe.assemble_grid()
        pipe.assemble_tereg()
        pipe.assemble_breg()
        pipe.assemble_ternd()
        pipe.assemble_brnd()
        pipe.assemble_cre()
        pipe.load_mask()
        if getattr(pipe, "cre_grid", None) is not None:
            raise ValueError("differentiable(): the CRE grid has no adjoint")
        g = pipe.par.grid_obs
        if not g.do_sync:
            raise ValueError("differentiable(): the adjoint needs a <sync> output")
        d0 = pipe.par.los_desc()
        if params is None:
            params = desc_params(d0).requires_grad_(True)
        brnd = pipe.brnd
        te = pipe.ternd
        if grids_grad:
            brnd = tuple(a.detach().clone().requires_grad_(True) for a in brnd) if brnd else None
            te = te.detach().clone().requires_grad_(True) if te is not None else None
        kw = dict(brnd=brnd, brnd_box=pipe.par.grid_brnd.box() if brnd else None,
                  ternd=te, ternd_box=pipe.par.grid_ternd.box() if te is not None else None)
        # the mask and the regular-field grids go in as they are
        kw.update({k: v for k, v in pipe._fields().items()
                   if k in ("mask", "nside_mask", "breg", "breg_box", "tereg", "tereg_box")})
        out = {}
        n = len(g.do_sync)
        for p in range(n):
            fi = n - 1 - p  # the pass's sync entry
            d = pipe.par.los_desc()
            d.sync_freq = g.sim_sync_freq[fi]
            d.nside_sync = g.nside_sync[fi]
            d.do_sync = 1 if g.do_sync[fi] else 0
            if p > 0:
                d.do_dm = d.do_fd = 0
            maps = los_maps(ctx, d, params, **kw)
            if d.do_sync:
                key = ("sync", g.sim_sync_freq_str[fi], str(d.nside_sync))
                for comp, k in zip("IQU", ("is", "qs", "us")):
                    out[key + (comp,)] = maps[k]
            if d.do_dm:
                out[("dm", "nan", str(d.nside_dm), "nan")] = maps["dm"] * (CGS.ccm / CGS.pc)
            if d.do_fd:
                out[("fd", "nan", str(d.nside_fd), "nan")] = maps["fd"] * (CGS.m * CGS.m)
        self.sim_map.update({k: v.detach().cpu().numpy() for k, v in out.items()})
        grids = (brnd[0], brnd[1], brnd[2], te) if brnd else (None, None, None, te)
        return out, params, grids

    # ------------------------------------------------------ XML editing
    @staticmethod
    def _path(keychain: list) -> str:
        return "." + "".join("/" + k for k in keychain)

    def mod_par(self, keychain=None, attrib=None) -> None:
        if type(attrib) is not dict or type(keychain) is not list:
            raise ValueError("wrong input %s %s" % (keychain, attrib))
        target = self.tree.getroot().find(self._path(keychain))
        if target is None:
            raise ValueError("wrong path %s" % self._path(keychain))
        for k in attrib:
            target.set(k, attrib.get(k))

    def add_par(self, keychain=None, subkey=None, attrib=None) -> None:
        if type(keychain) is not list or type(subkey) is not str:
            raise ValueError("wrong input %s %s %s" % (keychain, subkey, attrib))
        target = self.tree.getroot().find(self._path(keychain))
        if target is None:
            raise ValueError("wrong path %s" % self._path(keychain))
        if attrib is None:
            et.SubElement(target, subkey)
        elif type(attrib) is dict:
            for existed in target.findall(subkey):
                if existed.attrib == attrib:
                    raise ValueError("repeatitive input %s %s %s" % (keychain, subkey, attrib))
            et.SubElement(target, subkey, attrib)
        else:
            raise ValueError("wrong input %s %s %s" % (keychain, subkey, attrib))

    def print_par(self, keychain=None) -> None:
        root = self.tree.getroot()
        if keychain is None:
            for child in root:
                print(child.tag, child.attrib)
            return
        if type(keychain) is not list:
            raise ValueError("wrong input %s" % keychain)
        for target in root.findall(self._path(keychain)):
            print(target.tag, target.attrib)
            for child in target:
                print("|--> ", child.tag, child.attrib)

    def del_par(self, keychain=None, opt=None) -> None:
        if type(keychain) is not list or not keychain:
            raise ValueError("wrong input %s" % keychain)
        root = self.tree.getroot()
        path = self._path(keychain)
        parent_path = self._path(keychain[:-1])
        target = root.find(path)
        parent = root.find(parent_path)
        if target is None or parent is None:
            raise ValueError("wrong path %s" % path)
        if opt is None:
            parent.remove(target)
        elif opt == "all":
            for t in parent.findall(keychain[-1]):
                parent.remove(t)
        else:
            raise ValueError("unsupported option at %s" % keychain)
