"""Python handle on the C-ABI: a device context and typed calls.

Arrays may be numpy arrays (host memory) or torch tensors on the GPU
(device memory); all inputs of one call must share a residency, and so must
all outputs.  Every call is synchronous and raises :class:`HamxError` on a
negative return code, carrying the library's message -- the Python analogue
of the reference's ``std::runtime_error``.
"""
from __future__ import annotations

import ctypes as C
from typing import Any, Optional

import numpy as np

from . import _capi as capi


class HamxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hamx error {code}: {msg}")
        self.code = code


def _is_torch(a: Any) -> bool:
    return hasattr(a, "data_ptr") and hasattr(a, "is_cuda")


def _ptr(a: Any) -> tuple[Optional[int], Optional[bool]]:
    """(address, on_device) of an array, (None, None) for None."""
    if a is None:
        return None, None
    if _is_torch(a):
        if not a.is_contiguous():
            raise ValueError("tensor must be contiguous")
        import torch
        if a.dtype != torch.float64:  # int64 / complex64 would be read as doubles
            raise ValueError(f"tensor must hold float64 values, not {a.dtype}")
        return a.data_ptr(), bool(a.is_cuda)
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous or a.dtype != np.float64:
            raise ValueError("numpy array must be C-contiguous float64")
        return a.ctypes.data, False
    raise TypeError(f"unsupported array type {type(a)!r}")


def _order_after_torch(arrays: list[Any]) -> None:
    """The library runs on its own HIP stream.  Before it reads or writes
    torch CUDA tensors, torch's pending work on the device (zero-fills,
    copies, elementwise ops on those tensors) must be complete, or the two
    streams race; the library synchronises its own stream before every call
    returns, so torch work issued after a call sees its results."""
    for a in arrays:
        if a is not None and _is_torch(a) and a.is_cuda:
            import torch
            torch.cuda.synchronize(a.device)
            return


def _residency(arrays: list[Any], what: str) -> bool:
    res = {_ptr(a)[1] for a in arrays if a is not None}
    if len(res) > 1:
        raise ValueError(f"all {what} arrays must share one residency (host or device)")
    return bool(res.pop()) if res else False


def box(nx: int, ny: int, nz: int, x_min: float, x_max: float, y_min: float,
        y_max: float, z_min: float, z_max: float) -> capi.HamxBox:
    return capi.HamxBox(nx, ny, nz, x_min, x_max, y_min, y_max, z_min, z_max)


def _fields(brnd, brnd_box, ternd, ternd_box, breg, breg_box, tereg, tereg_box, mask,
            nside_mask, cre, cre_box, cre_E):
    """hamx_fields of a LOS call and the arrays it points at."""
    f = capi.HamxFields()
    inputs = []
    if cre is not None:
        f.cre = _ptr(cre)[0]
        f.cre_box = cre_box
        f.cre_E_min, f.cre_E_fact, f.cre_nE = cre_E
        inputs.append(cre)
    if mask is not None:
        f.mask = _ptr(mask)[0]
        f.nside_mask = nside_mask
        inputs.append(mask)
    if brnd is not None:
        f.brnd_x, f.brnd_y, f.brnd_z = (_ptr(a)[0] for a in brnd)
        f.brnd_box = brnd_box
        inputs += list(brnd)
    if ternd is not None:
        f.ternd = _ptr(ternd)[0]
        f.ternd_box = ternd_box
        inputs.append(ternd)
    if breg is not None:
        f.breg_x, f.breg_y, f.breg_z = (_ptr(a)[0] for a in breg)
        f.breg_box = breg_box
        inputs += list(breg)
    if tereg is not None:
        f.tereg = _ptr(tereg)[0]
        f.tereg_box = tereg_box
        inputs.append(tereg)
    return f, inputs


class Context:
    """One HIP stream plus the library's device workspace on `device`."""

    def __init__(self, device: int = 0):
        self.lib = capi.load()
        h = C.c_void_p()
        rc = self.lib.hamx_open(device, C.byref(h))
        if rc != capi.HAMX_OK:
            raise HamxError(rc, f"hamx_open(device={device}) failed (no usable GPU?)")
        self.handle = h

    def close(self) -> None:
        if self.handle:
            self.lib.hamx_close(self.handle)
            self.handle = None

    def __enter__(self) -> "Context":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int) -> None:
        if rc != capi.HAMX_OK:
            msg = self.lib.hamx_last_error(self.handle)
            raise HamxError(rc, msg.decode() if msg else "")

    def sync(self) -> None:
        self._check(self.lib.hamx_sync(self.handle))

    def profile(self, enable: bool = True) -> None:
        """Start (clearing earlier records) or stop kernel timing."""
        self._check(self.lib.hamx_profile(self.handle, int(enable)))

    def kernel_times(self) -> list[dict]:
        """Per-launch records [{kernel, shell, work, ms}] since profile()."""
        n = C.c_int()
        self._check(self.lib.hamx_profile_read(self.handle, None, 0, C.byref(n)))
        arr = (capi.HamxKernelTime * max(n.value, 1))()
        self._check(self.lib.hamx_profile_read(self.handle, arr, n.value, C.byref(n)))
        return [dict(kernel=r.kernel.decode(), shell=r.shell, work=r.work, ms=r.ms)
                for r in arr[:n.value]]

    # ------------------------------------------------------------------ LOS
    def los_forward(self, desc: capi.HamxLosDesc, maps: dict[str, Any],
                    brnd: Optional[tuple[Any, Any, Any]] = None,
                    brnd_box: Optional[capi.HamxBox] = None,
                    ternd: Any = None, ternd_box: Optional[capi.HamxBox] = None,
                    breg: Optional[tuple[Any, Any, Any]] = None,
                    breg_box: Optional[capi.HamxBox] = None,
                    tereg: Any = None, tereg_box: Optional[capi.HamxBox] = None,
                    keep_tape: bool = False, mask: Any = None, nside_mask: int = 0,
                    cre: Any = None, cre_box: Optional[capi.HamxBox] = None,
                    cre_E: Optional[tuple[float, float, int]] = None) -> None:
        """One Integrator::write_grid pass into `maps` (keys dm, fd, is, qs, us).
        keep_tape leaves the pass's tape in the context for los_adjoint; mask
        (a map at nside_mask) restricts the integration to pixels where its
        duplicate at each shell's nside is 1.0; cre is the CRE flux grid of
        HAMX_CRE_GRID on cre_box with the energy axis cre_E = (E_min, E_fact,
        nE), stored [x][y][z][E]."""
        f, inputs = _fields(brnd, brnd_box, ternd, ternd_box, breg, breg_box, tereg, tereg_box,
                            mask, nside_mask, cre, cre_box, cre_E)
        outs = [maps.get(k) for k in ("dm", "fd", "is", "qs", "us")]
        m = capi.HamxMaps(*[_ptr(a)[0] for a in outs])
        flags = 0
        if _residency(inputs, "input"):
            flags |= capi.HAMX_IN_DEVICE
        if _residency(outs, "output"):
            flags |= capi.HAMX_OUT_DEVICE
        if keep_tape:
            flags |= capi.HAMX_KEEP_TAPE
        _order_after_torch(inputs + outs)
        self._check(self.lib.hamx_los_forward(self.handle, C.byref(desc), C.byref(f),
                                              C.byref(m), flags))

    def los_adjoint(self, desc: capi.HamxLosDesc, adj: dict[str, Any], dparams: Any,
                    brnd: Optional[tuple[Any, Any, Any]] = None,
                    brnd_box: Optional[capi.HamxBox] = None,
                    ternd: Any = None, ternd_box: Optional[capi.HamxBox] = None,
                    dbrnd: Optional[tuple[Any, Any, Any]] = None,
                    dternd: Any = None, use_tape: bool = False,
                    breg: Optional[tuple[Any, Any, Any]] = None,
                    breg_box: Optional[capi.HamxBox] = None,
                    tereg: Any = None, tereg_box: Optional[capi.HamxBox] = None,
                    mask: Any = None, nside_mask: int = 0) -> None:
        """Reverse mode of los_forward: `adj` holds dL/d(maps) (keys dm, fd, is,
        qs, us; missing maps are zero); writes dL/d(b0, psi0, psi1, chi0,
        alpha, n0) into `dparams` and dL/d(grid) into `dbrnd` / `dternd`.
        The field arguments are los_forward's.  use_tape takes the tape of the
        last los_forward(keep_tape=True) with the same descriptor and arrays
        instead of re-running the forward."""
        f, fin = _fields(brnd, brnd_box, ternd, ternd_box, breg, breg_box, tereg, tereg_box,
                         mask, nside_mask, None, None, None)
        inputs = [adj.get(k) for k in ("dm", "fd", "is", "qs", "us")] + fin
        m = capi.HamxMaps(*[_ptr(a)[0] for a in inputs[:5]])
        outs = [dparams] + (list(dbrnd) if dbrnd is not None else []) + [dternd]
        flags = 0
        if _residency(inputs, "input"):
            flags |= capi.HAMX_IN_DEVICE
        if _residency(outs, "output"):
            flags |= capi.HAMX_OUT_DEVICE
        if use_tape:
            flags |= capi.HAMX_USE_TAPE
        gx, gy, gz = (_ptr(a)[0] for a in dbrnd) if dbrnd is not None else (None,) * 3
        _order_after_torch(inputs + outs)
        self._check(self.lib.hamx_los_adjoint(self.handle, C.byref(desc), C.byref(f),
                                              C.byref(m), _ptr(dparams)[0], gx, gy, gz,
                                              _ptr(dternd)[0], flags))

    # -------------------------------------------------------------- batch
    def batch(self, instances: list[dict], precision: int = capi.HAMX_PREC_F32,
              tune: bool = True, adjoint: bool = False, lanes: bool = False) -> "Batch":
        """Plan a batched LOS run (hamx_batch_create).  Each instance is a dict:
        desc (HamxLosDesc), freqs (1-4 Hz values, <sync> list order), maps
        (dm / fd tensors or None, is / qs / us lists of tensors, one per
        frequency), and optionally brnd (3 tensors), ternd and box.  With
        adjoint=True an instance may also hold adj (dL/d(maps) in the layout
        of maps) and grads (dparams: 6 values; dbrnd: 3 grid tensors;
        dternd), filled by Batch.run_adjoint().  Every array is a float64
        CUDA tensor; addresses are fixed for the plan's lifetime.  lanes=True
        takes the FP32 forward with one ray per lane instead of timing the
        launch shapes."""
        return Batch(self, instances, precision, tune, adjoint, lanes)

    # -------------------------------------------------------- random fields
    def brnd_es(self, desc: capi.HamxEsDesc, seed: int, bx: Any, by: Any, bz: Any,
                breg_grid: Optional[tuple[Any, Any, Any]] = None) -> None:
        """Brnd_es::write_grid (brnd_es.cc:76-308) into bx, by, bz [G]."""
        self._brnd(self.lib.hamx_brnd_es, desc, seed, bx, by, bz, breg_grid)

    def brnd_jf12(self, desc: capi.HamxEsDesc, seed: int, bx: Any, by: Any, bz: Any,
                  breg_grid: Optional[tuple[Any, Any, Any]] = None) -> None:
        """Brnd_jf12::write_grid (brnd_jf12.cc:127-385): the ES steps with the JF12
        energy profile taken from desc.jf12."""
        self._brnd(self.lib.hamx_brnd_jf12, desc, seed, bx, by, bz, breg_grid)

    def brnd_mhd(self, desc: capi.HamxMhdDesc, seed: int, bx: Any, by: Any, bz: Any,
                 breg_grid: Optional[tuple[Any, Any, Any]] = None) -> None:
        """Brnd_mhd::write_grid (brnd_mhd.cc:16-167): Alfven, fast and slow
        modes about the regular field at the observer."""
        self._brnd(self.lib.hamx_brnd_mhd, desc, seed, bx, by, bz, breg_grid)

    def _brnd(self, fn, desc, seed, bx, by, bz, breg_grid) -> None:
        flags = capi.HAMX_OUT_DEVICE if _residency([bx, by, bz], "output") else 0
        if breg_grid is not None:
            desc.breg_x, desc.breg_y, desc.breg_z = (_ptr(a)[0] for a in breg_grid)
            if _residency(list(breg_grid), "input"):
                flags |= capi.HAMX_IN_DEVICE
        _order_after_torch([bx, by, bz] + (list(breg_grid) if breg_grid is not None else []))
        self._check(fn(self.handle, C.byref(desc), seed,
                       _ptr(bx)[0], _ptr(by)[0], _ptr(bz)[0], flags))

    def cre_write_grid(self, desc: capi.HamxLosDesc, box: capi.HamxBox, E_min: float,
                       E_fact: float, nE: int, out: Any) -> None:
        """CREfield::write_grid (cre.cc:204-225) of the analytic / uniform model
        in `desc` into `out` ([x][y][z][E], box_size(box) * nE values)."""
        flags = capi.HAMX_OUT_DEVICE if _residency([out], "output") else 0
        _order_after_torch([out])
        self._check(self.lib.hamx_cre_write_grid(self.handle, C.byref(desc), C.byref(box),
                                                 E_min, E_fact, nE, _ptr(out)[0], flags))

    def sync_kernels(self, x: Any, F: Any, G: Any) -> None:
        """The device synchrotron kernels F(x) and G(x) for 0 < x <= 100."""
        n = len(x)
        flags = capi.HAMX_IN_DEVICE if _residency([x], "input") else 0
        if _residency([F, G], "output"):
            flags |= capi.HAMX_OUT_DEVICE
        _order_after_torch([x, F, G])
        self._check(self.lib.hamx_sync_kernels(self.handle, _ptr(x)[0], n, _ptr(F)[0],
                                               _ptr(G)[0], flags))

    def ternd_dft(self, desc: capi.HamxDftDesc, seed: int, te: Any) -> None:
        flags = capi.HAMX_OUT_DEVICE if _residency([te], "output") else 0
        _order_after_torch([te])
        self._check(self.lib.hamx_ternd_dft(self.handle, C.byref(desc), seed,
                                            _ptr(te)[0], flags))

    def fft3d(self, data: Any, nx: int, ny: int, nz: int, sign: int) -> None:
        """In-place FFT of interleaved complex128 data stored as float64 pairs."""
        addr, dev = _ptr(data)
        flags = capi.HAMX_IN_DEVICE | capi.HAMX_OUT_DEVICE if dev else 0
        _order_after_torch([data])
        self._check(self.lib.hamx_fft3d(self.handle, addr, nx, ny, nz, sign, flags))

    def gauss_pairs(self, seed: int, stream: int, first: int, count: int,
                    out: Any) -> None:
        addr, dev = _ptr(out)
        flags = capi.HAMX_OUT_DEVICE if dev else 0
        _order_after_torch([out])
        self._check(self.lib.hamx_gauss_pairs(self.handle, seed, stream, first, count,
                                              addr, flags))


def los_samples(desc: capi.HamxLosDesc) -> int:
    """Geometric LOS samples of one write_grid pass (host arithmetic only)."""
    lib = capi.load()
    n = C.c_uint64()
    rc = lib.hamx_los_samples(C.byref(desc), C.byref(n))
    if rc != capi.HAMX_OK:
        raise HamxError(rc, "invalid LOS descriptor")
    return int(n.value)


class Batch:
    """A planned batch: run() replays the captured graph."""

    def __init__(self, ctx: Context, instances: list[dict], precision: int, tune: bool,
                 adjoint: bool = False, lanes: bool = False):
        self.ctx = ctx
        self.keep = []
        arr = (capi.HamxBatchInst * len(instances))()
        for x, spec in zip(arr, instances):
            C.memmove(C.byref(x.los), C.byref(spec["desc"]), C.sizeof(capi.HamxLosDesc))
            freqs = list(spec["freqs"])
            if not 1 <= len(freqs) <= capi.HAMX_MAX_FREQ:
                raise ValueError("1 to 4 frequencies per instance")
            x.n_freq = len(freqs)
            for i, f in enumerate(freqs):
                x.freq[i] = f
            arrays = []
            brnd = spec.get("brnd")
            if brnd is not None:
                x.brnd_x, x.brnd_y, x.brnd_z = (_ptr(a)[0] for a in brnd)
                arrays += list(brnd)
            te = spec.get("ternd")
            if te is not None:
                x.ternd = _ptr(te)[0]
                arrays.append(te)
            if spec.get("box") is not None:
                x.box = spec["box"]
            m = spec["maps"]
            x.dm = _ptr(m.get("dm"))[0]
            x.fd = _ptr(m.get("fd"))[0]
            arrays += [m.get("dm"), m.get("fd")]
            for key, field in (("is", x.is_), ("qs", x.qs), ("us", x.us)):
                lst = m[key]
                if len(lst) != len(freqs):
                    raise ValueError(f"maps[{key!r}] needs one map per frequency")
                for i, a in enumerate(lst):
                    field[i] = _ptr(a)[0]
                arrays += list(lst)
            adj = spec.get("adj") or {}
            x.adj_dm = _ptr(adj.get("dm"))[0]
            x.adj_fd = _ptr(adj.get("fd"))[0]
            arrays += [adj.get("dm"), adj.get("fd")]
            for key, field in (("is", x.adj_is), ("qs", x.adj_qs), ("us", x.adj_us)):
                lst = adj.get(key) or [None] * len(freqs)
                if len(lst) != len(freqs):
                    raise ValueError(f"adj[{key!r}] needs one map per frequency")
                for i, a in enumerate(lst):
                    field[i] = _ptr(a)[0]
                arrays += list(lst)
            gr = spec.get("grads") or {}
            x.dparams = _ptr(gr.get("dparams"))[0]
            db = gr.get("dbrnd")
            if db is not None:
                x.dbrnd_x, x.dbrnd_y, x.dbrnd_z = (_ptr(a)[0] for a in db)
                arrays += list(db)
            x.dternd = _ptr(gr.get("dternd"))[0]
            arrays += [gr.get("dparams"), gr.get("dternd")]
            for a in arrays:
                if a is not None and not _ptr(a)[1]:
                    raise ValueError("batch arrays must be device tensors")
            self.keep += [a for a in arrays if a is not None]
        flags = capi.HAMX_IN_DEVICE | capi.HAMX_OUT_DEVICE
        if not tune:
            flags |= capi.HAMX_BATCH_NO_TUNE
        if adjoint:
            flags |= capi.HAMX_BATCH_ADJOINT
        if lanes:
            flags |= capi.HAMX_BATCH_LANES
        h = C.c_void_p()
        _order_after_torch(self.keep)
        ctx._check(ctx.lib.hamx_batch_create(ctx.handle, arr, len(instances), precision,
                                             flags, C.byref(h)))
        self.handle = h
        self._arr = arr

    def run(self) -> None:
        _order_after_torch(self.keep)
        self.ctx._check(self.ctx.lib.hamx_batch_run(self.handle))

    def run_adjoint(self) -> None:
        _order_after_torch(self.keep)
        self.ctx._check(self.ctx.lib.hamx_batch_run_adjoint(self.handle))

    def info(self) -> dict:
        i = capi.HamxBatchInfo()
        self.ctx._check(self.ctx.lib.hamx_batch_get_info(self.handle, C.byref(i)))
        return dict(samples_geo=i.samples_geo, samples_ref=i.samples_ref, levels=i.levels,
                    variant=i.variant, shape=i.shape.decode(),
                    variant_ms=list(i.variant_ms[:i.n_variants]))

    def close(self) -> None:
        if self.handle:
            self.ctx.lib.hamx_batch_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
