"""biofem_amd: host side of the MI355X element internal-force library.

The compute path is libbiofem_hip.so (HIP kernels for gfx950 behind the C ABI in
include/biofem_hip.h); this package generates the synthetic meshes, mirrors the
reference's FE/FInternalAct interface and drives batched and served evaluation.
"""
from . import _lib
from .fem import ActParams, ActStrainFE
from .mesh import Mesh, flat_membrane

__all__ = ["ActParams", "ActStrainFE", "Mesh", "flat_membrane", "_lib"]
