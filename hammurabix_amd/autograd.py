"""Differentiable LOS maps: torch.autograd over hamx_los_forward /
hamx_los_adjoint.

``los_maps(ctx, desc, params, brnd=..., ternd=...)`` returns the output maps
of one write_grid pass (the forward of integrator.cc:23-134) as float64
device tensors; calling ``backward`` on any scalar of them runs the adjoint
and fills ``params.grad`` (d/d of b0, psi0, psi1, chi0 of the LSA field,
alpha of the CRE model, n0 of the uniform thermal electrons, in the CGS /
radian units of the descriptor) and the random-grid tensors' ``.grad``.
Any regular B and TE model of the forward and a mask may be given; the CRE
is the analytic or uniform model with a constant index.

The forward and the adjoint run on the library's stream and are synchronous
with respect to the caller.  The forward keeps its tape (FD checkpoints) in
the context and the backward uses it; if another pass has replaced the tape
in between, the backward re-runs the forward to record it again.
"""
from __future__ import annotations

import ctypes as C
from typing import Any, Optional

import torch

from . import _capi as capi
from .runtime import Context, HamxError

MAP_KEYS = ("dm", "fd", "is", "qs", "us")
# descriptor slots of the differentiable parameters, in HAMX_NPARAM order
PARAM_SLOTS = (("breg", 0), ("breg", 1), ("breg", 2), ("breg", 3), ("cre", 0),
               ("tereg", 0))


def desc_params(desc: capi.HamxLosDesc) -> torch.Tensor:
    """The six differentiable parameters of a descriptor as a float64 tensor."""
    return torch.tensor([getattr(desc, a)[j] for a, j in PARAM_SLOTS], dtype=torch.float64)


def _with_params(desc: capi.HamxLosDesc, params: torch.Tensor) -> capi.HamxLosDesc:
    d = capi.HamxLosDesc()
    C.memmove(C.byref(d), C.byref(desc), C.sizeof(d))
    vals = params.detach().to("cpu", torch.float64).tolist()
    for (a, j), v in zip(PARAM_SLOTS, vals):
        getattr(d, a)[j] = v
    return d


def map_sizes(desc: capi.HamxLosDesc) -> dict[str, int]:
    n = {}
    if desc.do_dm:
        n["dm"] = 12 * desc.nside_dm ** 2
    if desc.do_fd:
        n["fd"] = 12 * desc.nside_fd ** 2
    if desc.do_sync:
        for k in ("is", "qs", "us"):
            n[k] = 12 * desc.nside_sync ** 2
    return n


class _LosFunction(torch.autograd.Function):
    @staticmethod
    def forward(fctx, params, bx, by, bz, te, ctx, desc, brnd_box, ternd_box, keys, fixed):
        d = _with_params(desc, params)
        dev = torch.device("cuda", torch.cuda.current_device())
        sizes = map_sizes(d)
        maps = {k: torch.empty(sizes[k], dtype=torch.float64, device=dev) for k in keys}
        brnd = (bx, by, bz) if bx is not None else None
        ctx.los_forward(d, maps, brnd=brnd, brnd_box=brnd_box, ternd=te,
                        ternd_box=ternd_box, keep_tape=True, **fixed)
        fctx.ctx, fctx.desc, fctx.keys, fctx.fixed = ctx, d, keys, fixed
        fctx.brnd_box, fctx.ternd_box = brnd_box, ternd_box
        fctx.save_for_backward(params, bx, by, bz, te)
        return tuple(maps[k] for k in keys)

    @staticmethod
    def backward(fctx, *grads):
        params, bx, by, bz, te = fctx.saved_tensors
        d = fctx.desc
        sizes = map_sizes(d)
        dev = torch.device("cuda", torch.cuda.current_device())
        adj = {}
        for k, g in zip(fctx.keys, grads):
            adj[k] = (g.to(dev, torch.float64).contiguous() if g is not None else
                      torch.zeros(sizes[k], dtype=torch.float64, device=dev))
        dparams = torch.zeros(capi.HAMX_NPARAM, dtype=torch.float64, device=dev)
        want_b = bx is not None and any(fctx.needs_input_grad[1:4])
        want_t = te is not None and fctx.needs_input_grad[4]
        n = bx.numel() if bx is not None else 0
        dbrnd = tuple(torch.zeros(n, dtype=torch.float64, device=dev)
                      for _ in range(3)) if want_b else None
        # the TE gradient lives on the B grid's box (one packed record per vertex)
        dternd = torch.zeros(n, dtype=torch.float64, device=dev) if want_t else None
        brnd = (bx, by, bz) if bx is not None else None
        kw = dict(brnd=brnd, brnd_box=fctx.brnd_box, ternd=te, ternd_box=fctx.ternd_box,
                  dbrnd=dbrnd, dternd=dternd, **fctx.fixed)
        try:  # the tape of this forward, unless another pass has replaced it
            fctx.ctx.los_adjoint(d, adj, dparams, use_tape=True, **kw)
        except HamxError as e:
            if e.code != capi.HAMX_ERR_ARG:
                raise
            fctx.ctx.los_adjoint(d, adj, dparams, **kw)
        gp = dparams.to(params.device, params.dtype) if fctx.needs_input_grad[0] else None
        gb = dbrnd if dbrnd is not None else (None, None, None)
        return (gp, gb[0], gb[1], gb[2], dternd, None, None, None, None, None, None)


def los_maps(ctx: Context, desc: capi.HamxLosDesc, params: torch.Tensor,
             brnd: Optional[tuple[torch.Tensor, torch.Tensor, torch.Tensor]] = None,
             brnd_box: Optional[capi.HamxBox] = None, ternd: Optional[torch.Tensor] = None,
             ternd_box: Optional[capi.HamxBox] = None,
             **fixed: Any) -> dict[str, torch.Tensor]:
    """Differentiable write_grid pass.  `params` (6 float64 values, any device)
    replaces the descriptor's b0, psi0, psi1, chi0, alpha and n0; grids are
    float64 device tensors.  With a TE grid the B grid must be given on the
    same box (the adjoint keeps one record per vertex).  `fixed` holds the
    los_forward field arguments that are not differentiated: mask /
    nside_mask, the regular-field grids breg / breg_box and tereg /
    tereg_box.  The gradients of b0, psi0, psi1 and chi0 are zero unless the
    regular B model is LSA, n0's unless the TE model is the uniform sphere."""
    if ternd is not None and brnd is None:
        raise ValueError("a differentiable TE grid needs the random B grid on the same box")
    allowed = {"mask", "nside_mask", "breg", "breg_box", "tereg", "tereg_box"}
    if set(fixed) - allowed:
        raise ValueError(f"los_maps: unknown field arguments {sorted(set(fixed) - allowed)}")
    keys = tuple(k for k in MAP_KEYS if k in map_sizes(desc))
    bx, by, bz = brnd if brnd is not None else (None, None, None)
    outs = _LosFunction.apply(params, bx, by, bz, ternd, ctx, desc, brnd_box, ternd_box,
                              keys, fixed)
    return dict(zip(keys, outs))


def gaussian_loglike(maps: dict[str, torch.Tensor], data: dict[str, torch.Tensor],
                     sigma: dict[str, Any]) -> torch.Tensor:
    """-1/2 sum over the given maps of ((map - data) / sigma)^2."""
    total = None
    for k, obs in data.items():
        r = (maps[k] - obs) / sigma[k]
        t = -0.5 * torch.sum(r * r)
        total = t if total is None else total + t
    return total
