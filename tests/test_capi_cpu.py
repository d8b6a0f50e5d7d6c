"""C-ABI checks that need no GPU: every function include/hamx.h declares is
exported by libhamx.so, the ctypes structs have the C layout, host-only
entry points compute the reference's sample counts, and a context on a
machine without a usable GPU fails with an error code (no crash)."""
import ctypes as C
import os
import re
import subprocess

import pytest

from hammurabix_amd import Param, _capi as capi, configs, los_samples

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hamx.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*(hamx_\w+)\s*\(",
                                 text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    names = declared_functions()
    assert len(names) >= 14
    out = subprocess.run(["nm", "-D", "--defined-only", capi.lib_path()],
                         capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (hamx_\w+)$", out, flags=re.M))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert getattr(lib, n) is not None
    assert set(capi.exported_symbols()) == set(names)


def test_struct_layout_matches_c(tmp_path):
    src = tmp_path / "sz.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "hamx.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu %zu\\n\","
        "sizeof(hamx_box), sizeof(hamx_fields), sizeof(hamx_los_desc),"
        "sizeof(hamx_maps), sizeof(hamx_es_desc), sizeof(hamx_dft_desc),"
        "offsetof(hamx_los_desc, cre), offsetof(hamx_es_desc, breg_box));"
        "printf(\"%zu %zu %zu %zu %zu\\n\", sizeof(hamx_batch_inst), sizeof(hamx_batch_info),"
        "offsetof(hamx_batch_inst, box), offsetof(hamx_batch_inst, us),"
        "sizeof(hamx_kernel_time));return 0;}\n")
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    c_sizes = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                              check=True).stdout.split()]
    py = [C.sizeof(capi.HamxBox), C.sizeof(capi.HamxFields), C.sizeof(capi.HamxLosDesc),
          C.sizeof(capi.HamxMaps), C.sizeof(capi.HamxEsDesc), C.sizeof(capi.HamxDftDesc),
          capi.HamxLosDesc.cre.offset, capi.HamxEsDesc.breg_box.offset,
          C.sizeof(capi.HamxBatchInst), C.sizeof(capi.HamxBatchInfo),
          capi.HamxBatchInst.box.offset, capi.HamxBatchInst.us.offset,
          C.sizeof(capi.HamxKernelTime)]
    assert c_sizes == py


def test_sample_counts_of_the_workloads():
    # plain forward: 12288 pixels x 3000 steps
    d = Param(text=configs.plain_forward()).los_desc()
    assert los_samples(d) == 12288 * 3000
    # random forward: nside 64/128/256 with 750/750/1500 steps
    d = Param(text=configs.random_forward()).los_desc()
    assert los_samples(d) == 49152 * 750 + 196608 * 750 + 786432 * 1500


def test_open_without_gpu_is_an_error_not_a_crash():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is visible")
    except ImportError:
        pass
    lib = capi.load()
    h = C.c_void_p()
    rc = lib.hamx_open(0, C.byref(h))
    assert rc == capi.HAMX_ERR_DEVICE
    assert not h.value
    assert lib.hamx_last_error(None) == b"null hamx context"
