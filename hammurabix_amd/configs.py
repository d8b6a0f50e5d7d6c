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
  <mask cue="{mask_on}" filename="{mask_file}" nside="{nside_mask}"/>
  <fieldio>
    <breg read="0" write="{breg_write}" filename="{breg_file}"/>
    <brnd read="{brnd_read}" write="0" filename="{brnd_file}"/>
    <tereg read="0" write="{tereg_write}" filename="{tereg_file}"/>
    <ternd read="{ternd_read}" write="0" filename="{ternd_file}"/>
    <cre read="{cre_read}" write="{cre_write}" filename="{cre_file}"/>
  </fieldio>
  <grid>
    <observer>
      <x value="-8.3"/>
      <y value="0"/>
      <z value="0.006"/>
    </observer>
    <box_breg>
      <nx value="{rg_nx}"/> <ny value="{rg_ny}"/> <nz value="{rg_nz}"/>
      <x_min value="-20.0"/> <x_max value="20.0"/>
      <y_min value="-20.0"/> <y_max value="20.0"/>
      <z_min value="-4.0"/> <z_max value="4.0"/>
    </box_breg>
    <box_tereg>
      <nx value="{rg_nx}"/> <ny value="{rg_ny}"/> <nz value="{rg_nz}"/>
      <x_min value="-20.0"/> <x_max value="20.0"/>
      <y_min value="-20.0"/> <y_max value="20.0"/>
      <z_min value="-4.0"/> <z_max value="4.0"/>
    </box_tereg>
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
    <box_cre>
      <nx value="{cre_nx}"/> <ny value="{cre_ny}"/> <nz value="{cre_nz}"/> <nE value="{cre_nE}"/>
      <x_min value="-20.0"/> <x_max value="20.0"/>
      <y_min value="-20.0"/> <y_max value="20.0"/>
      <z_min value="-4.0"/> <z_max value="4.0"/>
      <E_min value="{cre_E_min}"/> <E_max value="{cre_E_max}"/>
    </box_cre>
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
    <random cue="{brnd_on}" type="{brnd_type}" seed="{seed}">
      <local type="mhd">
        <mhd>
          <pa0 value="1"/> <pf0 value="1"/> <ps0 value="1"/>
          <k0 value="0.3"/> <aa0 value="1.7"/> <af0 value="1.5"/> <as0 value="1.7"/>
          <k1 value="0.1"/> <a1 value="0.0"/> <beta value="0.1"/> <ma value="0.5"/>
        </mhd>
      </local>
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
    <regular cue="1" type="{tereg}">
      <ymw16>
        <warp> <r_warp value="8.4"/> <gamma_w value="0.14"/> <r0 value="8.3"/> </warp>
        <thickdisk> <ad value="2500"/> <bd value="15000"/> <n1 value="0.01132"/> <h1 value="1673"/> </thickdisk>
        <thindisk> <a2 value="1200"/> <b2 value="4000"/> <n2 value="0.404"/> <k2 value="1.54"/> </thindisk>
        <spiralarm>
          <b2s value="4000"/>
          <ele_arm_0 value="0.135"/> <ele_arm_1 value="0.129"/> <ele_arm_2 value="0.103"/>
          <ele_arm_3 value="0.116"/> <ele_arm_4 value="0.0057"/>
          <wid_arm_0 value="300"/> <wid_arm_1 value="500"/> <wid_arm_2 value="300"/>
          <wid_arm_3 value="500"/> <wid_arm_4 value="300"/>
          <rref_arm_0 value="3.35"/> <rref_arm_1 value="3.707"/> <rref_arm_2 value="3.56"/>
          <rref_arm_3 value="3.670"/> <rref_arm_4 value="8.21"/>
          <phiref_arm_0 value="44.4"/> <phiref_arm_1 value="120.0"/> <phiref_arm_2 value="218.6"/>
          <phiref_arm_3 value="330.3"/> <phiref_arm_4 value="55.1"/>
          <pitch_arm_0 value="11.43"/> <pitch_arm_1 value="9.84"/> <pitch_arm_2 value="10.38"/>
          <pitch_arm_3 value="10.54"/> <pitch_arm_4 value="2.77"/>
          <aa value="11680"/> <ka value="5.015"/> <ncn value="2.4"/> <wcn value="8.2"/>
          <thetacn value="109"/> <nsg value="0.626"/> <wsg value="20"/> <thetasg value="75.8"/>
        </spiralarm>
        <galcenter> <ngc value="6.2"/> <agc value="160"/> <hgc value="35"/> </galcenter>
        <gumnebula> <ngn value="1.84"/> <wgn value="15.1"/> <agn value="125.8"/> <kgn value="1.4"/> </gumnebula>
        <localbubble>
          <j_lb value="0.480"/>
          <nlb1 value="1.094"/> <detlb1 value="28.4"/> <wlb1 value="14.2"/> <hlb1 value="112.9"/>
          <thetalb1 value="195.4"/>
          <nlb2 value="2.33"/> <detlb2 value="14.7"/> <wlb2 value="15.6"/> <hlb2 value="43.6"/>
          <thetalb2 value="278.2"/>
        </localbubble>
        <loopi> <nli value="1.907"/> <rli value="80.0"/> <wli value="15.0"/> <detthetali value="30.0"/>
          <thetali value="40.0"/> </loopi>
      </ymw16>
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
  <cre cue="1" type="{cre_type}">
    <analytic>
      <alpha value="3.0"/>
      <beta value="0.0"/>
      <theta value="0.0"/>
      <r0 value="5.0"/>
      <z0 value="1.0"/>
      <E0 value="20.6"/>
      <j0 value="0.0217"/>
    </analytic>
    <unif>
      <alpha value="3.0"/>
      <r0 value="5.0"/>
      <E0 value="20.6"/>
      <j0 value="0.0217"/>
    </unif>
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
        out: str = "", mask_file: str = "", nside_mask: int = 32,
        cre_file: str = "", cre_read: bool = False, cre_write: bool = False,
        cre_grid: tuple[int, int, int, int] = (40, 40, 8, 32),
        cre_E: tuple[float, float] = (0.1, 1000.0), tereg: str = "unif",
        brnd_type: str = "global", breg_file: str = "", tereg_file: str = "",
        regular_grid: tuple[int, int, int] = (80, 80, 16), cre_type: str = "analytic") -> str:
    """The parameter file; cre_read / cre_write with cre_file switch on the
    CRE grid I/O on box_cre (cre_grid = nx, ny, nz, nE; cre_E = E_min, E_max
    in GeV); tereg picks the regular thermal electron model ("unif" or
    "ymw16", the template's YMW16 values); brnd_type "local" switches the
    random B to the MHD generator (k0 0.3 /kpc, the template's other
    values); breg_file / tereg_file switch on the export of the regular B /
    TE grid on box_breg / box_tereg (regular_grid = nx, ny, nz of both
    boxes); cre_type "unif" switches the CRE to the uniform sphere (alpha 3,
    r0 5 kpc)."""
    freqs = [23.0] if freqs is None else freqs
    return _TEMPLATE.format(
        do_dm=int(do_dm), do_fd=int(do_fd), nside_out=nside_out,
        sync=_sync_lines(freqs, nside_out, out), out=out,
        brnd_read=int(bool(brnd_file)), ternd_read=int(bool(ternd_file)),
        brnd_file=brnd_file or "brnd.bin", ternd_file=ternd_file or "ternd.bin",
        nx=grid[0], ny=grid[1], nz=grid[2], shell_num=shell_num, nside_sim=nside_sim,
        oc_r_max=oc_r_max, oc_r_res=oc_r_res, b0=b0, psi0=psi0,
        brnd_on=int(random), ternd_on=int(random), seed=seed, te_k0=te_k0,
        mask_on=int(bool(mask_file)), mask_file=mask_file or "mask.bin", nside_mask=nside_mask,
        cre_read=int(cre_read), cre_write=int(cre_write), cre_file=cre_file or "cre.bin",
        cre_nx=cre_grid[0], cre_ny=cre_grid[1], cre_nz=cre_grid[2], cre_nE=cre_grid[3],
        cre_E_min=cre_E[0], cre_E_max=cre_E[1], tereg=tereg, brnd_type=brnd_type,
        breg_write=int(bool(breg_file)), breg_file=breg_file or "breg.bin",
        tereg_write=int(bool(tereg_file)), tereg_file=tereg_file or "tereg.bin",
        rg_nx=regular_grid[0], rg_ny=regular_grid[1], rg_nz=regular_grid[2],
        cre_type=cre_type)


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


# ---- the 64-instance batch (many realisations and parameter sets)

BATCH_FREQS_GHZ = (0.408, 1.4, 23.0, 30.0)
BATCH_TOP_NSIDE = (64, 128, 256)
# (b0 [muG] ~ U[4, 8], psi0 [deg] ~ U[20, 35]) of instance i, drawn once with
# numpy.random.default_rng(20251202) (b0 for all 64 instances, then psi0)
# and frozen here.
BATCH_B0_PSI0 = (
    (6.1138, 33.4704), (4.9177, 31.7859), (6.5748, 28.1820), (7.0530, 30.1541),
    (6.6507, 29.8894), (5.7448, 32.9300), (6.5333, 27.8276), (6.4442, 29.9588),
    (4.1436, 27.6601), (6.8107, 31.4726), (7.7355, 34.6906), (6.9297, 24.7631),
    (6.4880, 29.0623), (6.5368, 27.3158), (6.6774, 21.8906), (4.0631, 25.8351),
    (7.3027, 22.9290), (7.5530, 23.1883), (6.5636, 34.6284), (5.6773, 34.5441),
    (6.7436, 33.3368), (6.6236, 32.5872), (6.1060, 27.4111), (4.4917, 26.5854),
    (4.0906, 30.2557), (5.0300, 33.1263), (6.4519, 20.4097), (5.3218, 29.0560),
    (5.1073, 32.0901), (7.7029, 24.2940), (6.4919, 28.8323), (5.6668, 33.2728),
    (6.2141, 34.4645), (7.5456, 28.6275), (6.4639, 21.5399), (6.7816, 28.3655),
    (4.6443, 21.3638), (7.2567, 28.8985), (4.5609, 33.4839), (5.4024, 20.9197),
    (4.3294, 32.8582), (6.9287, 27.2931), (5.4564, 27.5291), (4.8879, 26.4520),
    (5.0011, 34.9021), (4.7695, 27.1010), (6.3286, 28.1691), (5.1859, 33.6316),
    (5.9223, 29.4615), (6.2254, 25.5421), (5.8939, 25.6044), (7.7351, 34.5893),
    (7.5156, 22.0591), (4.2985, 23.4984), (7.2770, 30.1869), (6.7702, 32.1974),
    (4.1267, 26.2203), (7.1226, 27.2274), (4.4570, 26.4148), (5.5693, 30.9069),
    (5.0068, 34.0949), (5.3785, 24.2213), (6.0916, 23.6613), (6.5287, 23.7561),
)


def batch_instance(i: int, **kw) -> str:
    """Instance i (0-based) of the batch: random fields with seed i + 1, top
    nside BATCH_TOP_NSIDE[i % 3] over 3 automatic shells (nside_sim = top/4),
    the first 1 + i % 4 frequencies of BATCH_FREQS_GHZ and the frozen
    (b0, psi0) draw."""
    top = BATCH_TOP_NSIDE[i % 3]
    b0, psi0 = BATCH_B0_PSI0[i]
    args = dict(shell_num=3, nside_sim=top // 4, nside_out=top, random=True, seed=i + 1,
                b0=b0, psi0=psi0, freqs=list(BATCH_FREQS_GHZ[:1 + i % 4]))
    args.update(kw)
    return xml(**args)


# ---- the large batched adjoint (512x512x128 grids, 4 shells up to nside 1024)

LARGE_GRID = (512, 512, 128)


def large_instance(i: int, **kw) -> str:
    """Instance i (0-based) of the large batch: ES random B and DFT random TE
    with seed i + 1 on a 512x512x128 box over the same extent, 4 automatic
    shells at nside 128/256/512/1024 (750/750/1500/3000 steps of 0.005 kpc,
    0-30 kpc), DM, FD and sync at 0.408, 1.4, 23 and 30 GHz on nside 1024,
    the default LSA / TE / CRE parameters."""
    args = dict(shell_num=4, nside_sim=128, nside_out=1024, oc_r_res=0.005, random=True,
                grid=LARGE_GRID, seed=i + 1, freqs=list(BATCH_FREQS_GHZ))
    args.update(kw)
    return xml(**args)
