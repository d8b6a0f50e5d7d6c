"""hammurabix_amd -- MI355X (gfx950) line-of-sight integrator and Gaussian
random-field generators of hammurabiX behind a C-ABI (include/hamx.h).

Layers:
  _capi     ctypes declarations of the C-ABI (libhamx.so, built in-tree)
  runtime   Context: device stream + typed, synchronous calls
  param     Param: the reference's XML schema parsed into CGS units and the
            C-ABI descriptors built from it
  pipeline  Pipeline: the reference's assemble_* sequence and multi-frequency
            write_grid passes, with map / grid export in the reference formats
  configs   synthetic workloads in the same XML schema
"""
from . import _capi
from .param import Param
from .runtime import Context, HamxError, box, los_samples

__all__ = ["Context", "HamxError", "Param", "box", "los_samples", "_capi"]
