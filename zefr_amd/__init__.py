"""zefr_amd: MI355X-native (gfx950, HIP) implementation of ZEFR's FRSolver::compute_residual
path -- direct flux reconstruction on quads and hexes, Euler / Navier-Stokes with
Rusanov + LDG interface fluxes, RK44 -- behind the reference's Zefr / zefr API.

Submodules:
  zefr   -- mirror of the reference's SWIG module (Zefr, InputStruct, initialize, ...)
  mesh   -- Gmsh 2.2 writer for periodic quad / hex boxes
  _lib   -- ctypes binding of libzefr_amd.so (include/zefr_amd.h)
"""

from . import mesh  # noqa: F401
from ._lib import build, lib  # noqa: F401
