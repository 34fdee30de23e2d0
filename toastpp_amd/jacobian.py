"""GenerateJacobian (src/libstoast/jacobian.cc:109-126) on the GPU library.

jacobian(mesh, raster, dphi, aphi, proj) follows toast.Mesh.Jacobian
(src/python/toast/mesh.py:262-289, toastmodule.cc:1433-1509): dphi is
(nlen x nQ) and aphi (nlen x nM) complex, node-major; proj holds the nQM
linear projections in link-list order; the result is the (2 nQM x 2 slen)
float64 matrix [dlnA/dcmua | dlnA/dckappa ; dphase/dcmua | dphase/dckappa].
raster None (or -1) selects the mesh (nodal) basis.
"""
import numpy as np

from . import _lib as L


def _c128(a, shape, name):
    a = np.ascontiguousarray(a, np.complex128)
    if a.shape != shape:
        raise ValueError("%s: expected shape %s, got %s" % (name, shape, a.shape))
    return a


def jacobian(mesh, raster, dphi, aphi, proj, scale="log"):
    n = mesh.NodeCount()
    nq, nm, nqm = mesh.qm_dims()
    dphi = _c128(dphi, (n, nq), "dphi")
    aphi = _c128(aphi, (n, nm), "aphi")
    proj = _c128(np.ravel(proj), (nqm,), "proj")
    dscale = L.DATA_LOG if scale == "log" else L.DATA_LIN
    if raster is None or (isinstance(raster, int) and raster < 0):
        J = np.empty((2 * nqm, 2 * n), np.float64)
        L.call("tmi_jacobian_nodal", mesh.handle_ptr(), nq, nm, dphi.view(np.float64), aphi.view(np.float64),
               proj.view(np.float64), dscale, J)
        return J
    return raster.Jacobian(dphi, aphi, proj, dscale)
