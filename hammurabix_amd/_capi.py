"""ctypes binding of the C-ABI declared in include/hamx.h.

The structures below mirror the C structs field by field; the library is the
in-tree ``hammurabix_amd/libhamx.so`` built by ``hammurabix_amd._build``.
Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HAMX_ABI_VERSION = 2
HAMX_MAX_SHELL = 16

HAMX_OK = 0
HAMX_ERR_ARG = -1
HAMX_ERR_UNSUPPORTED = -2
HAMX_ERR_DEVICE = -3
HAMX_ERR_NOMEM = -4
HAMX_ERR_NSIDE = -5

HAMX_IN_DEVICE = 1
HAMX_OUT_DEVICE = 2
HAMX_KEEP_TAPE = 8
HAMX_USE_TAPE = 16

HAMX_STATIC_MAX_NSIDE = 256
HAMX_STATIC_SCRATCH_BYTES = 40 * 1024 * 1024

(HAMX_BREG_NONE, HAMX_BREG_LSA, HAMX_BREG_CART, HAMX_BREG_UNIF, HAMX_BREG_GRID,
 HAMX_BREG_HELIX, HAMX_BREG_JF12, HAMX_BREG_JAFFE) = range(8)
HAMX_BREG_NPARAM = 40
HAMX_TEREG_NONE, HAMX_TEREG_UNIF, HAMX_TEREG_GRID, HAMX_TEREG_YMW16 = range(4)
HAMX_TEREG_NPARAM = 80
HAMX_CRE_NONE, HAMX_CRE_ANA, HAMX_CRE_UNIF, HAMX_CRE_GRID = range(4)
HAMX_SCHED_AUTO, HAMX_SCHED_THREAD, HAMX_SCHED_WAVE = range(3)

_dp = C.POINTER(C.c_double)


class HamxBox(C.Structure):
    _fields_ = [("nx", C.c_uint64), ("ny", C.c_uint64), ("nz", C.c_uint64),
                ("x_min", C.c_double), ("x_max", C.c_double),
                ("y_min", C.c_double), ("y_max", C.c_double),
                ("z_min", C.c_double), ("z_max", C.c_double)]


class HamxFields(C.Structure):
    _fields_ = [("brnd_x", C.c_void_p), ("brnd_y", C.c_void_p), ("brnd_z", C.c_void_p),
                ("brnd_box", HamxBox),
                ("ternd", C.c_void_p), ("ternd_box", HamxBox),
                ("breg_x", C.c_void_p), ("breg_y", C.c_void_p), ("breg_z", C.c_void_p),
                ("breg_box", HamxBox),
                ("tereg", C.c_void_p), ("tereg_box", HamxBox),
                ("mask", C.c_void_p), ("nside_mask", C.c_uint32), ("reserved", C.c_uint32),
                ("cre", C.c_void_p), ("cre_box", HamxBox),
                ("cre_E_min", C.c_double), ("cre_E_fact", C.c_double),
                ("cre_nE", C.c_uint32), ("reserved2", C.c_uint32)]


class HamxLosDesc(C.Structure):
    _fields_ = [("observer", C.c_double * 3),
                ("n_shell", C.c_uint32),
                ("nside_shell", C.c_uint32 * HAMX_MAX_SHELL),
                ("radii_shell", C.c_double * (HAMX_MAX_SHELL + 1)),
                ("oc_r_res", C.c_double),
                ("gc_r_min", C.c_double), ("gc_r_max", C.c_double),
                ("gc_z_min", C.c_double), ("gc_z_max", C.c_double),
                ("do_dm", C.c_int32), ("do_fd", C.c_int32), ("do_sync", C.c_int32),
                ("nside_dm", C.c_uint32), ("nside_fd", C.c_uint32), ("nside_sync", C.c_uint32),
                ("sync_freq", C.c_double),
                ("breg_model", C.c_int32), ("breg", C.c_double * HAMX_BREG_NPARAM),
                ("tereg_model", C.c_int32), ("tereg", C.c_double * HAMX_TEREG_NPARAM),
                ("cre_model", C.c_int32), ("cre", C.c_double * 8),
                ("schedule", C.c_int32), ("reserved", C.c_int32)]


class HamxMaps(C.Structure):
    _fields_ = [("dm", C.c_void_p), ("fd", C.c_void_p), ("is_", C.c_void_p),
                ("qs", C.c_void_p), ("us", C.c_void_p)]


class HamxEsDesc(C.Structure):
    _fields_ = [("box", HamxBox), ("observer", C.c_double * 3),
                ("rms", C.c_double), ("k0", C.c_double), ("k1", C.c_double),
                ("a0", C.c_double), ("a1", C.c_double), ("rho", C.c_double),
                ("r0", C.c_double), ("z0", C.c_double),
                ("breg_model", C.c_int32), ("reserved", C.c_int32),
                ("breg", C.c_double * HAMX_BREG_NPARAM),
                ("breg_x", C.c_void_p), ("breg_y", C.c_void_p), ("breg_z", C.c_void_p),
                ("breg_box", HamxBox), ("jf12", C.c_double * 13)]


class HamxMhdDesc(C.Structure):
    _fields_ = [("box", HamxBox), ("observer", C.c_double * 3),
                ("pa0", C.c_double), ("pf0", C.c_double), ("ps0", C.c_double),
                ("k0", C.c_double), ("k1", C.c_double),
                ("aa0", C.c_double), ("af0", C.c_double), ("as0", C.c_double),
                ("a1", C.c_double), ("beta", C.c_double), ("ma", C.c_double),
                ("breg_model", C.c_int32), ("reserved", C.c_int32),
                ("breg", C.c_double * HAMX_BREG_NPARAM),
                ("breg_x", C.c_void_p), ("breg_y", C.c_void_p), ("breg_z", C.c_void_p),
                ("breg_box", HamxBox)]


class HamxDftDesc(C.Structure):
    _fields_ = [("box", HamxBox), ("observer", C.c_double * 3),
                ("rms", C.c_double), ("k0", C.c_double), ("a0", C.c_double),
                ("r0", C.c_double), ("z0", C.c_double)]


# C function name -> (restype, argtypes)
_PROTOS = {
    "hamx_abi_version": (C.c_int, []),
    "hamx_open": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "hamx_close": (None, [C.c_void_p]),
    "hamx_last_error": (C.c_char_p, [C.c_void_p]),
    "hamx_sync": (C.c_int, [C.c_void_p]),
    "hamx_dev_alloc": (C.c_int, [C.c_void_p, C.c_size_t, C.POINTER(C.c_void_p)]),
    "hamx_dev_free": (C.c_int, [C.c_void_p, C.c_void_p]),
    "hamx_memcpy_h2d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "hamx_memcpy_d2h": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]),
    "hamx_los_forward": (C.c_int, [C.c_void_p, C.POINTER(HamxLosDesc),
                                   C.POINTER(HamxFields), C.POINTER(HamxMaps), C.c_uint32]),
    "hamx_los_forward_static": (C.c_int, [C.c_void_p, C.POINTER(HamxLosDesc),
                                          C.POINTER(HamxFields), C.POINTER(HamxMaps),
                                          C.c_uint32]),
    "hamx_los_samples": (C.c_int, [C.POINTER(HamxLosDesc), C.POINTER(C.c_uint64)]),
    "hamx_brnd_es": (C.c_int, [C.c_void_p, C.POINTER(HamxEsDesc), C.c_uint64,
                               C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "hamx_cre_write_grid": (C.c_int, [C.c_void_p, C.POINTER(HamxLosDesc), C.POINTER(HamxBox),
                                      C.c_double, C.c_double, C.c_uint32, C.c_void_p,
                                      C.c_uint32]),
    "hamx_sync_kernels": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p,
                                    C.c_void_p, C.c_uint32]),
    "hamx_brnd_mhd": (C.c_int, [C.c_void_p, C.POINTER(HamxMhdDesc), C.c_uint64,
                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "hamx_brnd_jf12": (C.c_int, [C.c_void_p, C.POINTER(HamxEsDesc), C.c_uint64,
                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "hamx_ternd_dft": (C.c_int, [C.c_void_p, C.POINTER(HamxDftDesc), C.c_uint64,
                                 C.c_void_p, C.c_uint32]),
    "hamx_fft3d": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64,
                             C.c_uint64, C.c_int, C.c_uint32]),
    "hamx_gauss_pairs": (C.c_int, [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint64,
                                   C.c_uint64, C.c_void_p, C.c_uint32]),
}

LIB_NAME = "libhamx.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

_lib = None


def lib_path() -> str:
    return LIB_PATH


def load() -> C.CDLL:
    """Load the in-tree HIP library; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build the HIP library first "
            "(python -c 'import __graft_entry__ as g; g.build()')")
    # One HIP runtime per process: torch ships its own libamdhip64 (same
    # soname, libamdhip64.so.7).  Loading torch first lets the library's
    # dependency resolve to that already-mapped runtime, so torch tensors and
    # the library share one device context.  Without torch the system ROCm
    # runtime is used.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.hamx_abi_version() != HAMX_ABI_VERSION:
        raise RuntimeError("libhamx.so ABI version mismatch")
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return list(_PROTOS)


class HamxKernelTime(C.Structure):
    _fields_ = [("kernel", C.c_char * 32), ("shell", C.c_uint32), ("reserved", C.c_uint32),
                ("work", C.c_uint64), ("ms", C.c_double)]


_PROTOS["hamx_profile"] = (C.c_int, [C.c_void_p, C.c_int])
_PROTOS["hamx_los_adjoint"] = (C.c_int, [C.c_void_p, C.POINTER(HamxLosDesc),
                                         C.POINTER(HamxFields), C.POINTER(HamxMaps),
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_uint32])
HAMX_NPARAM = 6
PARAM_NAMES = ("b0", "psi0", "psi1", "chi0", "alpha", "n0")
_PROTOS["hamx_profile_read"] = (C.c_int, [C.c_void_p, C.POINTER(HamxKernelTime), C.c_int,
                                          C.POINTER(C.c_int)])

# ---- batched LOS
HAMX_MAX_FREQ = 4
HAMX_PREC_F64 = 0
HAMX_PREC_F32 = 1
HAMX_BATCH_NO_TUNE = 4
HAMX_BATCH_ADJOINT = 32
HAMX_BATCH_LANES = 64


class HamxBatchInst(C.Structure):
    _fields_ = [("los", HamxLosDesc), ("n_freq", C.c_uint32), ("reserved", C.c_uint32),
                ("freq", C.c_double * HAMX_MAX_FREQ),
                ("brnd_x", C.c_void_p), ("brnd_y", C.c_void_p), ("brnd_z", C.c_void_p),
                ("ternd", C.c_void_p), ("box", HamxBox),
                ("dm", C.c_void_p), ("fd", C.c_void_p),
                ("is_", C.c_void_p * HAMX_MAX_FREQ), ("qs", C.c_void_p * HAMX_MAX_FREQ),
                ("us", C.c_void_p * HAMX_MAX_FREQ),
                ("adj_dm", C.c_void_p), ("adj_fd", C.c_void_p),
                ("adj_is", C.c_void_p * HAMX_MAX_FREQ), ("adj_qs", C.c_void_p * HAMX_MAX_FREQ),
                ("adj_us", C.c_void_p * HAMX_MAX_FREQ),
                ("dparams", C.c_void_p), ("dbrnd_x", C.c_void_p), ("dbrnd_y", C.c_void_p),
                ("dbrnd_z", C.c_void_p), ("dternd", C.c_void_p)]


class HamxBatchInfo(C.Structure):
    _fields_ = [("samples_geo", C.c_uint64), ("samples_ref", C.c_uint64),
                ("levels", C.c_uint32), ("variant", C.c_uint32), ("n_variants", C.c_uint32),
                ("reserved", C.c_uint32), ("variant_ms", C.c_double * 16),
                ("shape", C.c_char * 64)]


_PROTOS["hamx_batch_create"] = (C.c_int, [C.c_void_p, C.POINTER(HamxBatchInst), C.c_uint32,
                                          C.c_uint32, C.c_uint32, C.POINTER(C.c_void_p)])
_PROTOS["hamx_batch_run"] = (C.c_int, [C.c_void_p])
_PROTOS["hamx_batch_run_adjoint"] = (C.c_int, [C.c_void_p])
_PROTOS["hamx_batch_get_info"] = (C.c_int, [C.c_void_p, C.POINTER(HamxBatchInfo)])
_PROTOS["hamx_batch_destroy"] = (None, [C.c_void_p])
