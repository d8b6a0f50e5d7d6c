"""ctypes loader of the CPU oracle (oracle/liboracle.so), the checker the HIP
path is compared against.  Builds it with make when it is missing."""
import ctypes as C
import os
import subprocess

import numpy as np

from hammurabix_amd import _capi as capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

_dp = C.POINTER(C.c_double)
_vp = C.c_void_p

PROTOS = {
    "orc_pix2ang": (None, [C.c_uint64, C.c_uint64, _dp, _dp]),
    "orc_pix2ang_map": (None, [C.c_uint64, _vp]),
    "orc_interpolate": (C.c_double, [_vp, C.c_uint64, C.c_double, C.c_double]),
    "orc_accumulate": (None, [_vp, C.c_uint64, _vp, C.c_uint64]),
    "orc_shell_ref": (None, [C.c_double, C.c_double, C.c_double,
                             C.POINTER(C.c_uint64), _dp]),
    "orc_los_versor": (None, [C.c_double, C.c_double, _vp]),
    "orc_los_parproj": (C.c_double, [_vp, _vp]),
    "orc_los_perproj": (C.c_double, [_vp, _vp]),
    "orc_sync_ipa": (C.c_double, [_vp, C.c_double, C.c_double]),
    "orc_temp_convert": (C.c_double, [C.c_double, C.c_double]),
    "orc_check_upper": (C.c_int, [C.c_double, C.c_double]),
    "orc_check_lower": (C.c_int, [C.c_double, C.c_double]),
    "orc_breg": (None, [C.c_int, _vp, _vp, _vp]),
    "orc_tereg_unif": (C.c_double, [_vp, _vp, _vp]),
    "orc_read_grid3": (None, [_vp, _vp, _vp, C.POINTER(capi.HamxBox), _vp, _vp]),
    "orc_read_grid1": (C.c_double, [_vp, C.POINTER(capi.HamxBox), _vp]),
    "orc_sync_emissivity": (None, [C.POINTER(capi.HamxLosDesc), _vp, C.c_double, _dp, _dp]),
    "orc_write_grid": (C.c_int, [C.POINTER(capi.HamxLosDesc), C.POINTER(capi.HamxFields),
                                 C.POINTER(capi.HamxMaps), C.c_int]),
    "orc_radial_integration": (None, [C.POINTER(capi.HamxLosDesc), C.POINTER(capi.HamxFields),
                                      C.c_uint32, C.c_double, C.c_double, C.c_double, _vp]),
    "orc_philox4x32_10": (None, [_vp, _vp, _vp]),
    "orc_gauss_pairs": (None, [C.c_uint64, C.c_uint32, C.c_uint64, C.c_uint64, _vp]),
    "orc_fft3d": (None, [_vp, C.c_uint64, C.c_uint64, C.c_uint64, C.c_int]),
    "orc_variance": (C.c_double, [_vp, C.c_uint64]),
    "orc_brnd_es": (C.c_int, [C.POINTER(capi.HamxEsDesc), C.c_uint64, _vp, _vp, _vp]),
    "orc_sync_kernels": (None, [_vp, C.c_uint64, _vp, _vp]),
    "orc_sync_emissivity_grid": (None, [C.POINTER(capi.HamxLosDesc), C.POINTER(capi.HamxFields),
                                        _vp, C.c_double, _dp, _dp]),
    "orc_tereg_ymw16": (C.c_double, [_vp, _vp]),
    "orc_cre_read_grid_num": (C.c_double, [C.POINTER(capi.HamxFields), _vp, C.c_uint64]),
    "orc_cre_write_grid": (C.c_int, [C.POINTER(capi.HamxLosDesc), C.POINTER(capi.HamxBox),
                                     C.c_double, C.c_double, C.c_uint32, _vp]),
    "orc_brnd_jf12": (C.c_int, [C.POINTER(capi.HamxEsDesc), C.c_uint64, _vp, _vp, _vp]),
    "orc_jf12_profile": (C.c_double, [_vp, _vp]),
    "orc_brnd_mhd": (C.c_int, [C.POINTER(capi.HamxMhdDesc), C.c_uint64, _vp, _vp, _vp]),
    "orc_mhd_terms": (None, [C.c_double, C.c_double, C.c_double, _vp]),
    "orc_mhd_spectrum": (C.c_double, [C.POINTER(capi.HamxMhdDesc), C.c_double, C.c_double,
                                      C.c_double]),
    "orc_ternd_dft": (C.c_int, [C.POINTER(capi.HamxDftDesc), C.c_uint64, _vp]),
    "orc_es_spectrum": (C.c_double, [C.POINTER(capi.HamxEsDesc), C.c_double]),
    "orc_dft_spectrum": (C.c_double, [C.POINTER(capi.HamxDftDesc), C.c_double]),
    "orc_ring2xyf": (None, [C.c_uint64, C.c_uint64, C.POINTER(C.c_int64),
                            C.POINTER(C.c_int64), C.POINTER(C.c_int)]),
    "orc_xyf2ring": (C.c_uint64, [C.c_uint64, C.c_int64, C.c_int64, C.c_int]),
    "orc_mask_duplicate": (None, [_vp, C.c_uint64, C.c_uint64, _vp]),
}

_lib = None


def load_oracle():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_LIB):
        subprocess.run(["make", "-C", ORACLE_DIR], check=True, capture_output=True)
    lib = C.CDLL(ORACLE_LIB)
    for name, (res, args) in PROTOS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def ptr(a):
    """Address of a numpy array as a ctypes pointer that keeps the array
    alive for the duration of the call (temporaries included)."""
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def fields_struct(brnd=None, brnd_box=None, ternd=None, ternd_box=None,
                  breg=None, breg_box=None, tereg=None, tereg_box=None,
                  mask=None, nside_mask=0, cre=None, cre_box=None, cre_E=None):
    f = capi.HamxFields()
    if cre is not None:
        f.cre = ptr(cre)
        f.cre_box = cre_box
        f.cre_E_min, f.cre_E_fact, f.cre_nE = cre_E
    if mask is not None:
        f.mask = ptr(mask)
        f.nside_mask = nside_mask
    if brnd is not None:
        f.brnd_x, f.brnd_y, f.brnd_z = (ptr(a) for a in brnd)
        f.brnd_box = brnd_box
    if ternd is not None:
        f.ternd = ptr(ternd)
        f.ternd_box = ternd_box
    if breg is not None:
        f.breg_x, f.breg_y, f.breg_z = (ptr(a) for a in breg)
        f.breg_box = breg_box
    if tereg is not None:
        f.tereg = ptr(tereg)
        f.tereg_box = tereg_box
    return f


def npix(nside):
    return 12 * nside * nside


def alloc_maps(desc):
    """Zeroed host maps sized for a descriptor: dict and HamxMaps."""
    m = {}
    if desc.do_dm:
        m["dm"] = np.zeros(npix(desc.nside_dm))
    if desc.do_sync:
        for k in ("is", "qs", "us"):
            m[k] = np.zeros(npix(desc.nside_sync))
    if desc.do_fd or desc.do_sync:
        m["fd"] = np.zeros(npix(desc.nside_fd if desc.do_fd else desc.nside_sync))
    hm = capi.HamxMaps(ptr(m.get("dm")), ptr(m.get("fd")), ptr(m.get("is")),
                       ptr(m.get("qs")), ptr(m.get("us")))
    return m, hm


def write_grid(desc, fields=None, nthreads=0, **grids):
    """Oracle write_grid pass -> dict of maps."""
    lib = load_oracle()
    f = fields if fields is not None else fields_struct(**grids)
    m, hm = alloc_maps(desc)
    rc = lib.orc_write_grid(C.byref(desc), C.byref(f), C.byref(hm), nthreads)
    if rc != 0:
        raise RuntimeError(f"oracle write_grid failed: {rc}")
    return m


def rel_l2(a, b):
    """||a - b||_2 / ||b||_2 (b the reference)."""
    nb = np.linalg.norm(b)
    if nb == 0:
        return float(np.linalg.norm(a))
    return float(np.linalg.norm(a - b) / nb)
