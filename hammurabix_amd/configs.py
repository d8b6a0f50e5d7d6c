"""Synthetic workloads written in the reference's XML parameter schema
(templates/params_template.xml).

* ``plain_forward``    -- WMAP-LSA regular field (b0 6 muG, psi0 27 deg,
  psi1 0.9 deg, chi0 25 deg), uniform thermal electrons (n0 0.01 cm^-3,
  r0 3 kpc), analytic CRE (alpha 3, r0 5 kpc, z0 1 kpc), observer at
  (-8.3, 0, 0.006) kpc, DM / FD / sync 23 GHz, one shell at nside 32,
  0-30 kpc at 0.01 kpc.  No random fields.
* ``random_forward``   -- the same regular models plus a global ES random B
  (rms 2 muG, k0 0.1, a0 1.7, r0 8 kpc, z0 1 kpc, seed 1) and a DFT random
  TE (rms 0.001 cm^-3 = 0.1 n0, k0 2, a0 -1.7) on a 256x256x64 box over
  [-20,20]x[-20,20]x[-4,4] kpc; 3 shells at nside 64/128/256, outputs at
  nside 256, sync 23 GHz.

Every function takes overrides so tests can shrink a workload while keeping
its physics.  A grid coarser than 1/(2 k0) kpc has no DFT mode above the TE
cutoff k0 and the generator divides by a zero variance (as the reference's
does), so shrunk grids also pass a smaller ``te_k0``.
"""
from __future__ import annotations

from typing import Optional

_TEMPLATE = """<?xml version="1.0"?>
<root>
  <observable>
    <dm cue="{do_dm}" filename="{out}dm.bin" nside="{nside_out}"/>
    <faraday cue="{do_fd}" filename="{out}fd.bin" nside="{nside_out}"/>
{sync}
  </observable>
  <mask cue="0" filename="mask.bin" nside="32"/>
  <fieldio>
    <breg read="0" write="0" filename="breg.bin"/>
    <brnd read="{brnd_read}" write="0" filename="{brnd_file}"/>
    <tereg read="0" write="0" filename="tereg.bin"/>
    <ternd read="{ternd_read}" write="0" filename="{ternd_file}"/>
    <cre read="0" write="0" filename="cre.bin"/>
  </fieldio>
  <grid>
    <observer>
      <x value="-8.3"/>
      <y value="0"/>
      <z value="0.006"/>
    </observer>
    <box_brnd>
      <nx value="{nx}"/> <ny value="{ny}"/> <nz value="{nz}"/>
      <x_min value="-20.0"/> <x_max value="20.0"/>
      <y_min value="-20.0"/> <y_max value="20.0"/>
      <z_min value="-4.0"/> <z_max value="4.0"/>
    </box_brnd>
    <box_ternd>
      <nx value="{nx}"/> <ny value="{ny}"/> <nz value="{nz}"/>
      <x_min value="-20.0"/> <x_max value="20.0"/>
      <y_min value="-20.0"/> <y_max value="20.0"/>
      <z_min value="-4.0"/> <z_max value="4.0"/>
    </box_ternd>
    <shell>
      <layer type="auto">
        <auto>
          <shell_num value="{shell_num}"/>
          <nside_sim value="{nside_sim}"/>
        </auto>
        <manual>
          <cut value="0.5"/>
          <nside_sim value="32"/>
          <nside_sim value="64"/>
        </manual>
      </layer>
      <oc_r_min value="0.0"/>
      <oc_r_max value="{oc_r_max}"/>
      <gc_r_min value="0.0"/>
      <gc_r_max value="20.0"/>
      <gc_z_min value="-10.0"/>
      <gc_z_max value="10.0"/>
      <oc_r_res value="{oc_r_res}"/>
    </shell>
  </grid>
  <magneticfield>
    <regular cue="1" type="lsa">
      <lsa>
        <b0 value="{b0}"/>
        <psi0 value="{psi0}"/>
        <psi1 value="0.9"/>
        <chi0 value="25.0"/>
      </lsa>
    </regular>
    <random cue="{brnd_on}" type="global" seed="{seed}">
      <global type="es">
        <es>
          <rms value="2.0"/>
          <k0 value="0.1"/>
          <k1 value="0.1"/>
          <a0 value="1.7"/>
          <a1 value="0.0"/>
          <rho value="1.0"/>
          <r0 value="8.0"/>
          <z0 value="1.0"/>
        </es>
      </global>
    </random>
  </magneticfield>
  <thermalelectron>
    <regular cue="1" type="unif">
      <unif>
        <n0 value="0.01"/>
        <r0 value="3.0"/>
      </unif>
    </regular>
    <random cue="{ternd_on}" type="global" seed="{seed}">
      <global type="dft">
        <dft>
          <rms value="0.001"/>
          <k0 value="{te_k0}"/>
          <a0 value="-1.7"/>
          <r0 value="8.0"/>
          <z0 value="1.0"/>
        </dft>
      </global>
    </random>
  </thermalelectron>
  <cre cue="1" type="analytic">
    <analytic>
      <alpha value="3.0"/>
      <beta value="0.0"/>
      <theta value="0.0"/>
      <r0 value="5.0"/>
      <z0 value="1.0"/>
      <E0 value="20.6"/>
      <j0 value="0.0217"/>
    </analytic>
  </cre>
</root>
"""


def _sync_lines(freqs: list[float], nside: int, out: str) -> str:
    return "\n".join(
        f'    <sync cue="1" freq="{f}" filename="{out}sync_{f}.bin" nside="{nside}"/>'
        for f in freqs)


def xml(*, shell_num: int, nside_sim: int, nside_out: int, oc_r_max: float = 30.0,
        oc_r_res: float = 0.01, freqs: Optional[list[float]] = None,
        random: bool = False, grid: tuple[int, int, int] = (256, 256, 64),
        seed: int = 1, b0: float = 6.0, psi0: float = 27.0, te_k0: float = 2.0,
        do_dm: bool = True,
        do_fd: bool = True, brnd_file: str = "", ternd_file: str = "",
        out: str = "") -> str:
    freqs = [23.0] if freqs is None else freqs
    return _TEMPLATE.format(
        do_dm=int(do_dm), do_fd=int(do_fd), nside_out=nside_out,
        sync=_sync_lines(freqs, nside_out, out), out=out,
        brnd_read=int(bool(brnd_file)), ternd_read=int(bool(ternd_file)),
        brnd_file=brnd_file or "brnd.bin", ternd_file=ternd_file or "ternd.bin",
        nx=grid[0], ny=grid[1], nz=grid[2], shell_num=shell_num, nside_sim=nside_sim,
        oc_r_max=oc_r_max, oc_r_res=oc_r_res, b0=b0, psi0=psi0,
        brnd_on=int(random), ternd_on=int(random), seed=seed, te_k0=te_k0)


def plain_forward(**kw) -> str:
    """One shell at nside 32, 0-30 kpc at 0.01 kpc, analytic fields."""
    args = dict(shell_num=1, nside_sim=32, nside_out=32)
    args.update(kw)
    return xml(**args)


def random_forward(**kw) -> str:
    """3 shells at nside 64/128/256 with ES random B and DFT random TE."""
    args = dict(shell_num=3, nside_sim=64, nside_out=256, random=True)
    args.update(kw)
    return xml(**args)
