"""dipole_forces_ext on the MI355X: the HOOMD-facing package of the
mutual-polarization solver.

Same import surface as the reference package
("CPU - HOOMD simulation test/dipole_forces_ext/__init__.py"):
  compute_forces, __version__      -- truncated dipole forces (HIP kernel)
  dipole_forces_ext.ewald          -- spectral-Ewald field on the GPU
  dipole_forces_ext.capacitance    -- mutual-polarization GMRES on the GPU
  dipole_forces_ext.hoomd_force    -- hoomd.md.force.Custom wrappers (needs HOOMD)

There is no CPU path: without a ROCm GPU and the built libmpol.so every entry
point raises.
"""
from __future__ import annotations

import os as _os
import sys as _sys

_PKG_ROOT = _os.path.dirname(_os.path.dirname(_os.path.abspath(__file__)))
if _PKG_ROOT not in _sys.path:
    _sys.path.insert(0, _PKG_ROOT)

from ._forces import compute_forces  # noqa: E402

__version__ = "0.5.1"

__all__ = ["compute_forces", "__version__"]
