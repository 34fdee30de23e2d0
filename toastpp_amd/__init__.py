"""toastpp_amd: MI355X (gfx950) implementation of Toast++'s diffusion forward /
adjoint solve, boundary projection and Jacobian.

The compute path is libtoastpp_amd.so (hand-written HIP kernels behind the
C-ABI in include/toastpp_amd.h); this package is its Python surface:
  Mesh       -- toast.Mesh (src/python/toast/mesh.py)
  FwdSolver  -- libstoast TFwdSolver<T> (src/libstoast/fwdsolver.h)
  Solution   -- optical -> FEM coefficients (cmua, ckappa, c2a)
  Raster     -- toast.Raster with a Raster_Pixel basis (bdim <= gdim)
  ReadNim, WriteNim -- nodal image files (toast.ReadNim / WriteNim)
  Sysmat_CW, Jacobian, JacobianCW, Gradient -- toast's module-level
             functions, on Mesh / Raster objects
The command-line forward model (fwdfem, fwdfem_cw of src/supertoast) is
built as bin/fwdfem and bin/fwdfem_cw on the same library.
"""
from ._lib import TmiError, NotConverged  # noqa: F401
from .fwdsolver import FwdSolver, Solution  # noqa: F401
from .mesh import Gradient, Jacobian, JacobianCW, Mesh, ReadNim, Sysmat_CW, WriteNim  # noqa: F401
from .raster import Raster  # noqa: F401

__all__ = ["Mesh", "FwdSolver", "Solution", "Raster", "ReadNim", "WriteNim", "Sysmat_CW", "Jacobian", "JacobianCW",
           "Gradient", "TmiError", "NotConverged"]
