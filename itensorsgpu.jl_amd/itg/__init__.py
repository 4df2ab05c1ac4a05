"""itg — host-side mirror of the ITensors/ITensorsGPU dense-storage API over
libitg, an MI355X (gfx950) HIP library with a C ABI (include/itg.h).

Layers: `_lib` (ctypes binding), `ops` (storage-level calls, one per C entry
point), `itensor` (Index/ITensor with roc/cpu conversion, contraction,
factorisations), `mps` (MPS/MPO, ProjMPO, davidson, dmrg).
"""
from . import _lib, ops  # noqa: F401
from ._lib import ArgumentError, DimensionMismatch, ITGError  # noqa: F401
