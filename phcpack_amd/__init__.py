"""MI355X-native multiple-double power-series path (PHCpack hot path).

HIP kernels for gfx950 behind the C-ABI of include/phcmd.h, loaded with
ctypes.  Submodules:
  series   - series products and polynomial evaluation/differentiation
  tracker  - batched predictor-corrector path tracking of homotopies
"""
from ._lib import DeviceArray, PhcError, load  # noqa: F401

__all__ = ["DeviceArray", "PhcError", "load"]
