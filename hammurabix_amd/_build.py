"""Build the in-tree HIP library (gfx950) and the C++ host adapter.

``python -m hammurabix_amd._build`` compiles

* ``hammurabix_amd/libhamx.so``   -- every ``csrc/*.hip`` translation unit,
  hipcc --offload-arch=gfx950, FP64 parity flags (no FP contraction);
* ``hammurabix_amd/host/hamx``     -- the ``hamx <params.xml>`` driver and
  ``hammurabix_amd/libhamx_host.so`` (the C++ Param / Pipeline / Integrator
  mirror over the C-ABI), when the host sources exist.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
HOST = os.path.join(PKG, "host")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libhamx.so")
HOST_LIB = os.path.join(PKG, "libhamx_host.so")
HOST_EXE = os.path.join(HOST, "hamx")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17",
    # FP64 parity with the CPU restatement: evaluate expressions as written
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # keep LLVM's machine loop-invariant code motion from hoisting every
    # 64-bit constant out of the sample loops: in the FP64 marches the hoisted
    # constants filled the scalar registers, which the allocator then spilled
    # into vector-register lanes and read back with one VALU instruction each
    # (measured: LOS march 6.37 -> 5.91 ms, reverse march 20.8 -> 19.9 ms,
    # batched forward 83.5 -> 75.8 ms)
    "-mllvm", "-disable-machine-licm",
    "-Wall",
]

# Translation units whose FP64 arithmetic may be contracted into fused
# multiply-adds (a later -ffp-contract wins): the plain and reverse LOS
# marches, whose decisions (radial and z limits, the LSA window, the TE / CRE
# spheres, grid cells) can then differ from the reference's only for a sample
# within an ulp of a boundary, as with their other rewrites (measured: march
# 5.66 -> 5.37 ms, reverse march 19.9 -> 18.9 ms; the parity tests unchanged).
# The generators stay as written: their spectra give no power to a mode whose
# |k| ties a band edge exactly, and a contracted |k|^2 rounds such a tie
# differently (a 48 x 48 x 6 kpc box on 32 x 32 x 16 vertices holds modes with
# |k| = 0.5/kpc exactly; with k0 = 0.5/kpc its JF12 random field then differed
# from the oracle by 4e-2).
CONTRACT_SRCS = {"hamx_los.hip", "hamx_adj.hip"}

# The LOS march under LLVM's max-ILP machine scheduler (measured 5.31 -> 5.23 ms
# per march; the reverse march measured 1.4% slower with it and keeps the default).
ILP_SRCS = {"hamx_los.hip"}


def _run(cmd: list[str]) -> None:
    print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = (res.stdout or "") + (res.stderr or "")
    # hipcc's host pass does not know the packed-fp32 target feature
    lines = [ln for ln in out.splitlines() if "packed-fp32-ops" not in ln]
    if lines:
        print("\n".join(lines), flush=True)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")


def build_hip() -> str:
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    if not srcs:
        raise RuntimeError("no HIP sources found")
    objs = []
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    for s in srcs:
        o = os.path.join(objdir, os.path.basename(s) + ".o")
        base = os.path.basename(s)
        extra = ["-ffp-contract=fast"] if base in CONTRACT_SRCS else []
        if base in ILP_SRCS:
            extra += ["-mllvm", "-amdgpu-sched-strategy=max-ilp"]
        _run([HIPCC, *HIP_FLAGS, *extra, "-I", INCLUDE, "-I", CSRC, "-c", s, "-o", o])
        objs.append(o)
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_host() -> list[str]:
    srcs = sorted(glob.glob(os.path.join(HOST, "*.cc")))
    if not srcs:
        return []
    lib_srcs = [s for s in srcs if not s.endswith("main.cc")]
    # -I CSRC: the field models shared with the device code (hamx_models.h)
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-I", INCLUDE, "-I", HOST,
             "-I", CSRC]
    out = []
    tmp = HOST_LIB + ".tmp"
    _run(["g++", *flags, "-shared", "-o", tmp, *lib_srcs,
          "-L", PKG, "-lhamx", f"-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    out.append(HOST_LIB)
    main = os.path.join(HOST, "main.cc")
    if os.path.exists(main):
        _run(["g++", *flags, "-o", HOST_EXE, main, "-L", PKG, "-lhamx_host", "-lhamx",
              f"-Wl,-rpath,$ORIGIN/.."])
        out.append(HOST_EXE)
    return out


def build_all() -> list[str]:
    built = [build_hip()]
    built += build_host()
    return built


if __name__ == "__main__":
    for p in build_all():
        print("built", p)
    sys.exit(0)
