"""hamx_los_forward_static, the allocation-free variant of hamx_los_forward,
called through the C-ABI and compared with hamx_los_forward on the same
device inputs: two calls in a row, the second on a smaller sky with fewer
shells and no mask, so that the fixed scratch keeps the first call's
contents while the second runs.  FP64 maps must agree within 1e-12 relative
L2 (the static variant gathers the random grids where they lie instead of
packed, and may take the generic march, so the sums round differently)."""
import ctypes as C
import os

import numpy as np
import pytest

from hammurabix_amd import Param, _capi as capi, configs
from hammurabix_amd.units import CGS
from oracle_lib import npix, rel_l2

pytestmark = pytest.mark.gpu

TOL = 1e-12
MAPS = ("dm", "fd", "is", "qs", "us")


def torch_dev():
    import torch
    return torch, torch.device("cuda:0")


def device_grids(shape, seed):
    torch, dev = torch_dev()
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    b = tuple(torch.from_numpy(rng.normal(0, 2e-6, n)).to(dev) for _ in range(3))
    te = torch.from_numpy(rng.normal(0.0, 0.003, n)).to(dev)
    box = capi.HamxBox(*shape, -20 * CGS.kpc, 20 * CGS.kpc, -20 * CGS.kpc, 20 * CGS.kpc,
                       -4 * CGS.kpc, 4 * CGS.kpc)
    return b, te, box


def empty_maps(d):
    torch, dev = torch_dev()
    z = lambda ns: torch.full((npix(ns),), float("nan"), dtype=torch.float64, device=dev)  # noqa
    return {"dm": z(d.nside_dm), "fd": z(d.nside_fd), "is": z(d.nside_sync),
            "qs": z(d.nside_sync), "us": z(d.nside_sync)}


def call_static(ctx, d, maps, b, te, box, mask=None, nside_mask=0):
    torch, _ = torch_dev()
    f = capi.HamxFields()
    f.brnd_x, f.brnd_y, f.brnd_z = (a.data_ptr() for a in b)
    f.brnd_box = box
    f.ternd = te.data_ptr()
    f.ternd_box = box
    if mask is not None:
        f.mask = mask.data_ptr()
        f.nside_mask = nside_mask
    m = capi.HamxMaps(*[maps[k].data_ptr() for k in MAPS])
    torch.cuda.synchronize()
    return ctx.lib.hamx_los_forward_static(ctx.handle, C.byref(d), C.byref(f), C.byref(m),
                                           capi.HAMX_IN_DEVICE | capi.HAMX_OUT_DEVICE)


def check_pass(ctx, d, b, te, box, label, mask=None, nside_mask=0):
    got = empty_maps(d)
    rc = call_static(ctx, d, got, b, te, box, mask, nside_mask)
    assert rc == capi.HAMX_OK, ctx.lib.hamx_last_error(ctx.handle).decode()
    ref = empty_maps(d)
    ctx.los_forward(d, ref, brnd=b, brnd_box=box, ternd=te, ternd_box=box, mask=mask,
                    nside_mask=nside_mask)
    for k in MAPS:
        g, r = got[k].cpu().numpy(), ref[k].cpu().numpy()
        assert np.all(np.isfinite(g)), f"{label} {k}"
        err = rel_l2(g, r)
        print(f"{label} {k}: rel L2 static vs hamx_los_forward {err:.3e}", flush=True)
        assert err <= TOL, f"{label} {k}: {err:.3e}"


def test_static_forward_matches_forward_on_two_calls(ctx):
    torch, dev = torch_dev()
    print("case: static forward, call 1 (3 shells 16/32/64, mask, 32x32x16 grids)", flush=True)
    d1 = Param(text=configs.random_forward(nside_sim=16, nside_out=64, oc_r_res=0.05,
                                           grid=(32, 32, 16))).los_desc()
    b1, te1, box1 = device_grids((32, 32, 16), 1)
    piv = np.fromfile(os.path.join(os.path.dirname(__file__), "golden", "mask_map_nside8.bin"),
                      dtype="<f8")
    mask = torch.from_numpy(np.ascontiguousarray(piv)).to(dev)
    check_pass(ctx, d1, b1, te1, box1, "call 1", mask=mask, nside_mask=8)
    print("case: static forward, call 2 (1 shell at nside 8, 16x16x8 grids)", flush=True)
    d2 = Param(text=configs.random_forward(nside_sim=8, nside_out=16, oc_r_res=0.05,
                                           grid=(16, 16, 8), shell_num=1)).los_desc()
    b2, te2, box2 = device_grids((16, 16, 8), 2)
    check_pass(ctx, d2, b2, te2, box2, "call 2")


def test_static_forward_refuses_what_it_cannot_hold(ctx):
    torch, dev = torch_dev()
    d = Param(text=configs.random_forward(nside_sim=128, nside_out=64, oc_r_res=1.0,
                                          grid=(16, 16, 8))).los_desc()
    assert d.nside_shell[2] == 512 > capi.HAMX_STATIC_MAX_NSIDE
    b, te, box = device_grids((16, 16, 8), 3)
    rc = call_static(ctx, d, empty_maps(d), b, te, box)
    assert rc == capi.HAMX_ERR_ARG
    assert b"nside" in ctx.lib.hamx_last_error(ctx.handle)
    # host buffers (no device flags) are refused before anything runs
    m = capi.HamxMaps()
    rc = ctx.lib.hamx_los_forward_static(ctx.handle, C.byref(d), None, C.byref(m), 0)
    assert rc == capi.HAMX_ERR_ARG
    # the context stays usable
    d2 = Param(text=configs.random_forward(nside_sim=4, nside_out=8, oc_r_res=0.1,
                                           grid=(16, 16, 8))).los_desc()
    check_pass(ctx, d2, b, te, box, "after refusals")
