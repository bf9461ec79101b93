"""MI355X spectral-Ewald mutual-polarization library (HIP kernels behind a C ABI).

Submodules:
  _lib     -- ctypes binding of libmpol.so (include/mpol.h)
  context  -- Context: one geometry on one GPU (tables, precompute, matvec, GMRES)
  build    -- hipcc build of libmpol.so for gfx950
  autotune -- launch-shape and Ewald-grid autotuner
  batch    -- many independent configurations on concurrent streams
  frames   -- the frames of a trajectory on concurrent streams
"""
from .context import (Context, SolveResult, default_tables, real_space_table, real_space_table_field,
                      real_space_table_quad, real_space_table_quad_field, real_space_table_sizes,
                      real_space_table_sizes_field, require_device, FP32, FP64)

__all__ = ["Context", "SolveResult", "default_tables", "real_space_table", "real_space_table_field",
           "real_space_table_quad", "real_space_table_quad_field", "real_space_table_sizes",
           "real_space_table_sizes_field", "require_device", "FP32", "FP64"]
