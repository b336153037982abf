"""MI355X-native beta-surfel adjoint renderer (HIP for gfx950).

Public pieces:
  pale      -- drop-in for the reference's pybind11 module ``pale``
               (``pale.Renderer``), backed by the HIP C-ABI library.
  scene     -- seeded synthetic scenes of the benchmark configurations.
  _abi      -- ctypes mirror of include/pale_abi.h.
"""
from . import _abi, scene  # noqa: F401

__all__ = ["_abi", "scene"]
