"""GenerateJacobian (src/libstoast/jacobian.cc:109-126) on the GPU library.

jacobian(mesh, raster, dphi, aphi, proj) follows toast.Mesh.Jacobian
(src/python/toast/mesh.py:262-289, toastmodule.cc:1433-1509): dphi is
(nlen x nQ) and aphi (nlen x nM) complex, node-major; proj holds the nQM
linear projections in link-list order; the result is the (2 nQM x 2 slen)
float64 matrix [dlnA/dcmua | dlnA/dckappa ; dphase/dcmua | dphase/dckappa].
raster None (or -1) selects the mesh (nodal) basis, RASTER_NDBASIS;
"element" (or -2) the element basis, RASTER_ELBASIS (jacobian.h:33-35), with
one column per element.  Instead of proj the measurement vectors may be
given (mvec, nM x nlen): the projections are then computed from them, the
overload GenerateJacobian(raster, mesh, mvec, dphi, aphi, ...)
(jacobian.cc:86-101).  jacobian_cw is the CW (real) counterpart,
GenerateJacobian_cw (jacobian.cc:132-186).
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


def _basis(mesh, raster):
    """(basis code, raster handle or None, solution length)"""
    if raster is None or (isinstance(raster, int) and raster == -1):
        return L.BASIS_NODAL, None, mesh.NodeCount()
    if (isinstance(raster, str) and raster == "element") or (isinstance(raster, int) and raster == -2):
        return L.BASIS_ELEMENT, None, mesh.ElementCount()
    if isinstance(raster, (int, str)):
        raise ValueError("raster must be None / -1 (nodal basis), 'element' / -2 (element basis) or a Raster")
    return L.BASIS_RASTER, raster.handle_ptr(), raster.SLen()


def _mvec_csr(mvec, shape, dtype):
    from scipy import sparse

    m = sparse.csr_matrix(mvec)
    if m.shape != shape:
        raise ValueError("mvec: expected shape %s, got %s" % (shape, m.shape))
    vals = np.ascontiguousarray(m.data, dtype=dtype)
    return (np.ascontiguousarray(m.indptr, np.int32), np.ascontiguousarray(m.indices, np.int32),
            vals.view(np.float64) if dtype == np.complex128 else vals)


def jacobian_cw(mesh, raster, dphi, aphi, proj=None, scale="log", mua=True, kappa=False, mvec=None):
    """GenerateJacobian_cw (jacobian.cc:132-186): real fields dphi (nlen x nQ)
    and aphi (nlen x nM), node-major; proj: the nQM real linear projections
    (needed for log data), or mvec (nM x nlen) to compute them from.  Returns
    Jmua and/or Jkap (nQM x nsol) -- nsol the node count (raster None), the
    element count (raster 'element') or the raster's solution voxels -- as a
    tuple (Jmua, Jkap) with None for the block not asked for."""
    n = mesh.NodeCount()
    nq, nm, nqm = mesh.qm_dims()
    dphi = _f64(dphi, (n, nq), "dphi")
    aphi = _f64(aphi, (n, nm), "aphi")
    dscale = L.DATA_LOG if scale == "log" else L.DATA_LIN
    basis, rh, nsol = _basis(mesh, raster)
    Jm = np.empty((nqm, nsol)) if mua else None
    Jk = np.empty((nqm, nsol)) if kappa else None
    ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
    if mvec is not None:
        if proj is not None:
            raise ValueError("give either proj or mvec")
        rp, ci, val = _mvec_csr(mvec, (nm, n), np.float64)
        L.call("tmi_jacobian_cw_mvec", mesh.handle_ptr(), rh, basis, nq, nm, rp, ci, val, dphi, aphi, dscale,
               ptr(Jm), ptr(Jk))
        return Jm, Jk
    pj = None
    if proj is not None:
        pj = _f64(np.ravel(proj), (nqm,), "proj")
    elif dscale == L.DATA_LOG:
        raise ValueError("log data scaling needs the projections")
    if basis == L.BASIS_NODAL:
        L.call("tmi_jacobian_cw", mesh.handle_ptr(), nq, nm, dphi, aphi, ptr(pj), dscale, ptr(Jm), ptr(Jk))
    elif basis == L.BASIS_ELEMENT:
        L.call("tmi_jacobian_cw_element", mesh.handle_ptr(), nq, nm, dphi, aphi, ptr(pj), dscale, ptr(Jm), ptr(Jk))
    else:
        L.call("tmi_jacobian_raster_cw", rh, nq, nm, dphi, aphi, ptr(pj), dscale, ptr(Jm), ptr(Jk))
    return Jm, Jk


def jacobian(mesh, raster, dphi, aphi, proj=None, scale="log", mvec=None):
    n = mesh.NodeCount()
    nq, nm, nqm = mesh.qm_dims()
    dphi = _c128(dphi, (n, nq), "dphi")
    aphi = _c128(aphi, (n, nm), "aphi")
    dscale = L.DATA_LOG if scale == "log" else L.DATA_LIN
    basis, rh, nsol = _basis(mesh, raster)
    if mvec is not None:
        if proj is not None:
            raise ValueError("give either proj or mvec")
        rp, ci, val = _mvec_csr(mvec, (nm, n), np.complex128)
        J = np.empty((2 * nqm, 2 * nsol), np.float64)
        L.call("tmi_jacobian_mvec", mesh.handle_ptr(), rh, basis, nq, nm, rp, ci, val, dphi.view(np.float64),
               aphi.view(np.float64), dscale, J)
        return J
    if proj is None:
        raise ValueError("the Jacobian needs the projections (proj) or the measurement vectors (mvec)")
    proj = _c128(np.ravel(proj), (nqm,), "proj")
    if basis == L.BASIS_RASTER:
        return raster.Jacobian(dphi, aphi, proj, dscale)
    J = np.empty((2 * nqm, 2 * nsol), np.float64)
    L.call("tmi_jacobian_nodal" if basis == L.BASIS_NODAL else "tmi_jacobian_element", mesh.handle_ptr(), nq, nm,
           dphi.view(np.float64), aphi.view(np.float64), proj.view(np.float64), dscale, J)
    return J
