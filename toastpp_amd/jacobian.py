"""GenerateJacobian (src/libstoast/jacobian.cc:109-126) on the GPU library.

jacobian(mesh, raster, dphi, aphi, proj) follows toast.Mesh.Jacobian
(src/python/toast/mesh.py:262-289, toastmodule.cc:1433-1509): dphi is
(nlen x nQ) and aphi (nlen x nM) complex, node-major; proj holds the nQM
linear projections in link-list order; the result is the (2 nQM x 2 slen)
float64 matrix [dlnA/dcmua | dlnA/dckappa ; dphase/dcmua | dphase/dckappa].
raster None (or -1) selects the mesh (nodal) basis.  jacobian_cw is the CW
(real) counterpart, GenerateJacobian_cw (jacobian.cc:132-186).
"""
import ctypes as C

import numpy as np

from . import _lib as L


def _c128(a, shape, name):
    a = np.ascontiguousarray(a, np.complex128)
    if a.shape != shape:
        raise ValueError("%s: expected shape %s, got %s" % (name, shape, a.shape))
    return a


def _f64(a, shape, name):
    a = np.asarray(a)
    if np.iscomplexobj(a):
        if np.any(a.imag != 0):
            raise ValueError("%s must be real (CW fields)" % name)
        a = a.real
    a = np.ascontiguousarray(a, np.float64)
    if a.shape != shape:
        raise ValueError("%s: expected shape %s, got %s" % (name, shape, a.shape))
    return a


def jacobian_cw(mesh, raster, dphi, aphi, proj=None, scale="log", mua=True, kappa=False):
    """GenerateJacobian_cw (jacobian.cc:132-186): real fields dphi (nlen x nQ)
    and aphi (nlen x nM), node-major; proj: the nQM real linear projections
    (needed for log data).  Returns Jmua and/or Jkap (nQM x nsol) -- nsol the
    node count (raster None) or the raster's solution voxels -- as a tuple
    (Jmua, Jkap) with None for the block not asked for."""
    n = mesh.NodeCount()
    nq, nm, nqm = mesh.qm_dims()
    dphi = _f64(dphi, (n, nq), "dphi")
    aphi = _f64(aphi, (n, nm), "aphi")
    dscale = L.DATA_LOG if scale == "log" else L.DATA_LIN
    pj = None
    if proj is not None:
        pj = _f64(np.ravel(proj), (nqm,), "proj")
    elif dscale == L.DATA_LOG:
        raise ValueError("log data scaling needs the projections")
    nodal = raster is None or (isinstance(raster, int) and raster < 0)
    nsol = n if nodal else raster.SLen()
    Jm = np.empty((nqm, nsol)) if mua else None
    Jk = np.empty((nqm, nsol)) if kappa else None
    ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    if nodal:
        L.call("tmi_jacobian_cw", mesh.handle_ptr(), nq, nm, dphi, aphi, ptr(pj), dscale, ptr(Jm), ptr(Jk))
    else:
        L.call("tmi_jacobian_raster_cw", raster.handle_ptr(), nq, nm, dphi, aphi, ptr(pj), dscale, ptr(Jm), ptr(Jk))
    return Jm, Jk


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
