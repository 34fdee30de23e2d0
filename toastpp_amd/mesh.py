"""toast.Mesh surface (src/python/toast/mesh.py:7-327) on the GPU library.

Array layouts follow the reference module (toastmodule.cc): nodes (n x d)
float64, elements (m x d+1) int32 0-based, Qvec/Mvec as scipy CSC (n x nQ),
Fields -> (slen x nQ) complex128, Jacobian -> (2 nQM x 2 slen) float64.
The reference's integer handles (ObjectManager, src/common/objmgr_imp.hpp)
become Python objects holding a library handle; a cleared object has
handle 0 and every call on it raises.
"""
import ctypes as C

import numpy as np
from scipy import sparse

from . import _lib as L
from .fwdsolver import FwdSolver, Solution


class Mesh:
    """Finite-element mesh (Tri3 in 2D, Tet4 in 3D), device resident."""

    def __init__(self, name=None, data=None):
        self._h = C.c_void_p()
        if name is not None:
            self.Read(name)
        elif data is not None:
            self.Make(data[0], data[1], data[2] if len(data) > 2 else None)

    # ----------------------------------------------------------- handles
    def handle_ptr(self):
        if not self._h.value:
            raise L.TmiError(L.TMI_ERR_INVALID_ARG, "mesh has no data (handle 0)")
        return self._h

    def Handle(self):
        return int(self._h.value or 0)

    def __del__(self):
        self.Clear()

    def Clear(self):
        if getattr(self, "_h", None) is not None and self._h.value and L._lib is not None:
            L._lib.tmi_mesh_free(self._h)
        self._h = C.c_void_p()

    # ---------------------------------------------------------- creation
    def Read(self, name):
        self.Clear()
        L.call("tmi_mesh_read", name.encode(), C.byref(self._h))
        return self.Handle()

    def Make(self, nlist, elist, eltp=None):
        nlist = np.ascontiguousarray(nlist, np.float64)
        elist = np.ascontiguousarray(elist, np.int32)
        dim = nlist.shape[1]
        if eltp is not None:
            want = L.ELID_TET4 if dim == 3 else L.ELID_TRI3
            if np.any(np.asarray(eltp) != want):
                raise L.TmiError(L.TMI_ERR_UNSUPPORTED, "only Tri3 (2D) and Tet4 (3D) elements are supported")
        if elist.shape[1] != dim + 1:
            raise ValueError("elist must have %d columns" % (dim + 1))
        self.Clear()
        L.call("tmi_mesh_make", dim, nlist.shape[0], nlist, elist.shape[0], elist, C.byref(self._h))
        return self.Handle()

    # ----------------------------------------------------------- queries
    def _dims(self):
        d, n, e, b = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        L.call("tmi_mesh_dims", self.handle_ptr(), C.byref(d), C.byref(n), C.byref(e), C.byref(b))
        return d.value, n.value, e.value, b.value

    def NodeCount(self):
        return self._dims()[1]

    def ElementCount(self):
        return self._dims()[2]

    def Dim(self):
        return self._dims()[0]

    def BoundaryCount(self):
        return self._dims()[3]

    def Data(self):
        d, n, e, _ = self._dims()
        x = np.zeros((n, d))
        el = np.zeros((e, d + 1), np.int32)
        L.call("tmi_mesh_get_nodes", self.handle_ptr(), x)
        L.call("tmi_mesh_get_elements", self.handle_ptr(), el)
        tp = np.full(e, L.ELID_TET4 if d == 3 else L.ELID_TRI3, np.int32)
        return x, el, tp

    def BoundingBox(self):
        x = self.Data()[0]
        return np.vstack([x.min(axis=0), x.max(axis=0)])

    def ElementSize(self):
        s = np.zeros(self.ElementCount())
        L.call("tmi_mesh_element_size", self.handle_ptr(), s)
        return s

    def BoundaryFlags(self):
        b = np.zeros(self.NodeCount(), np.int32)
        L.call("tmi_mesh_get_bnd", self.handle_ptr(), b)
        return b

    def SparseStructure(self):
        """Mesh::SparseRowStructure: (rowptr, colidx) of the system matrix."""
        nnz = C.c_int()
        L.call("tmi_mesh_sparse_structure", self.handle_ptr(), None, None, C.byref(nnz))
        rp = np.zeros(self.NodeCount() + 1, np.int32)
        ci = np.zeros(nnz.value, np.int32)
        L.call("tmi_mesh_sparse_structure", self.handle_ptr(), L.ptr(rp), L.ptr(ci), C.byref(nnz))
        return rp, ci

    def ReadNim(self, name, idx=-1):
        """Nodal image `idx` of a NIM file (mesh.py:206-209): its length must
        be the node count."""
        nim = ReadNim(name, idx)
        if nim.size != self.NodeCount():
            raise ValueError("NIM image has %d values, the mesh %d nodes" % (nim.size, self.NodeCount()))
        return nim

    # ---------------------------------------------------------------- QM
    def ReadQM(self, name):
        L.call("tmi_mesh_read_qm", self.handle_ptr(), name.encode())

    def SetQM(self, Q, M, links=None, pull_m=False):
        """Q (nQ x d), M (nM x d) positions; links: list of detector lists per
        source (default: every source sees every detector)."""
        Q = np.ascontiguousarray(Q, np.float64)
        M = np.ascontiguousarray(M, np.float64)
        if links is None:
            links = [list(range(M.shape[0]))] * Q.shape[0]
        lptr = np.zeros(Q.shape[0] + 1, np.int32)
        lptr[1:] = np.cumsum([len(l) for l in links])
        lidx = np.ascontiguousarray(np.concatenate([np.asarray(l, np.int32) for l in links]) if len(links) else
                                    np.zeros(0, np.int32), np.int32)
        L.call("tmi_mesh_set_qm", self.handle_ptr(), Q.shape[0], L.ptr(Q), M.shape[0], L.ptr(M), lptr, L.ptr(lidx),
               1 if pull_m else 0)

    def qm_dims(self):
        q, m, qm = C.c_int(), C.c_int(), C.c_int()
        L.call("tmi_mesh_qm_dims", self.handle_ptr(), C.byref(q), C.byref(m), C.byref(qm))
        return q.value, m.value, qm.value

    def QMPositions(self):
        nq, nm, nqm = self.qm_dims()
        d = self.Dim()
        Q, M = np.zeros((nq, d)), np.zeros((nm, d))
        lp, li = np.zeros(nq + 1, np.int32), np.zeros(nqm, np.int32)
        L.call("tmi_mesh_get_qm", self.handle_ptr(), L.ptr(Q), L.ptr(M), L.ptr(lp), L.ptr(li))
        return Q, M, lp, li

    def _profiles(self, pts, width, neumann, scale=None):
        pts = np.ascontiguousarray(np.atleast_2d(pts), np.float64)
        out = np.zeros((pts.shape[0], self.NodeCount()))
        sc = None if scale is None else np.ascontiguousarray(scale, np.float64)
        L.call("tmi_qvec_gaussian", self.handle_ptr(), pts.shape[0], pts, float(width), 1 if neumann else 0,
               L.ptr(sc), out)
        return out

    def _points(self, pts, neumann):
        pts = np.ascontiguousarray(np.atleast_2d(pts), np.float64)
        out = np.zeros((pts.shape[0], self.NodeCount()))
        L.call("tmi_qvec_point", self.handle_ptr(), pts.shape[0], pts, 1 if neumann else 0, out)
        return out

    def Qvec(self, type="Neumann", shape="Gaussian", width=1):
        """Source vectors as a CSC (n x nQ) matrix (toastmodule.cc:1098-1170);
        shape 'Gaussian' (source.cc:4-41) or 'Point' (source.cc:157-196)."""
        t = type.lower()
        if t not in ("neumann", "isotropic"):
            raise ValueError("invalid source type %r" % type)
        Q = self.QMPositions()[0]
        if shape.lower() == "gaussian":
            q = self._profiles(Q, width, t == "neumann")
        elif shape.lower() == "point":
            q = self._points(Q, t == "neumann")
        else:
            raise L.TmiError(L.TMI_ERR_UNSUPPORTED, "source profile %r is not supported" % shape)
        return sparse.csc_matrix(sparse.csr_matrix(q.astype(np.complex128)).T)

    def Mvec(self, shape="Gaussian", width=1, ref=1):
        """Measurement vectors scaled by c0/(2 n A(n)) (toastmodule.cc:1189-1259).
        As in the reference (CalcMvec has no point case), 'Point' gives zero
        vectors."""
        n = self.NodeCount()
        if shape.lower() == "point":
            return sparse.csc_matrix((n, self.qm_dims()[1]), dtype=np.complex128)
        if shape.lower() != "gaussian":
            raise L.TmiError(L.TMI_ERR_UNSUPPORTED, "measurement profile %r is not supported" % shape)
        refv = np.broadcast_to(np.asarray(ref, np.float64), (n,)).copy()
        c2a = Solution(np.zeros(n), np.ones(n), refv).c2a
        M = self.QMPositions()[1]
        m = self._profiles(M, width, True, c2a)
        return sparse.csc_matrix(sparse.csr_matrix(m.astype(np.complex128)).T)

    # ------------------------------------------------------------ solves
    def Sysmat(self, mua, mus, ref, freq):
        """Complex FEM system matrix (CSR) (toastmodule.cc:1040-1096)."""
        fws = FwdSolver(self, "BICGSTAB", 1e-10, complex=True)
        fws.Allocate()
        fws.AssembleSystemMatrix(Solution(mua, mus, ref), freq * 2.0 * np.pi * 1e-6)
        rp, ci, val = fws.SystemMatrix()
        n = self.NodeCount()
        return sparse.csr_matrix((val, ci, rp), shape=(n, n))

    def Sysmat_CW(self, mua, mus, ref):
        fws = FwdSolver(self, "BICGSTAB", 1e-10, complex=False)
        fws.Allocate()
        fws.AssembleSystemMatrix(Solution(mua, mus, ref), 0.0)
        rp, ci, val = fws.SystemMatrix()
        n = self.NodeCount()
        return sparse.csr_matrix((val, ci, rp), shape=(n, n))

    def Fields(self, raster, qvec, mua, mus, ref, freq, solver="DIRECT", tol=1e-12, mode="mixed"):
        """Nodal photon density fields, (slen x nQ) complex (mesh.py:226-260).
        The reference always solves directly; here 'DIRECT' means an
        FP64-accurate GPU solve (relative residual 1e-14)."""
        q = sparse.csc_matrix(qvec)   # (n x nQ): its CSC arrays are the CSR of (nQ x n)
        qrows = sparse.csr_matrix((q.data, q.indices, q.indptr), shape=(q.shape[1], q.shape[0]))
        fws = FwdSolver(self, solver, tol, complex=True)
        fws.SetMode(mode)
        fws.Allocate()
        fws.Reset(Solution(mua, mus, ref), freq * 2.0 * np.pi * 1e-6)
        phi = fws.CalcFields(qrows)
        self.last_stats = fws.stats
        out = phi.T.copy()
        if raster is not None:
            out = raster.MeshToSol(out)
        return out

    def Jacobian(self, raster, dphi, aphi, proj):
        """Frequency-domain Jacobian, log data (toastmodule.cc:1433-1509)."""
        from .jacobian import jacobian
        return jacobian(self, raster, dphi, aphi, proj)

    def JacobianCW(self, raster, qvec, mvec, mua, mus, ref, solver="bicgstab", tol=1e-12, mode="mixed"):
        """CW Jacobian of log intensity with respect to mua on the raster
        basis (toast.JacobianCW, mesh.py:351-363 -> CalcJacobianCW,
        toastmodule.cc:1630-1680): real fields for qvec and mvec at frequency
        0, linear projections, GenerateJacobian_cw with DATA_LOG.  Returns
        (nQM x slen), or (nQM x nlen) with raster None."""
        from .jacobian import jacobian_cw
        q = sparse.csc_matrix(qvec)
        m = sparse.csc_matrix(mvec)
        for name, v in (("qvec", q), ("mvec", m)):
            if np.any(np.iscomplex(v.data)):
                raise ValueError("%s must be real or not have any nonzero complex elements" % name)
        q, m = q.real, m.real
        qrows = sparse.csr_matrix((q.data, q.indices, q.indptr), shape=(q.shape[1], q.shape[0]))
        mrows = sparse.csr_matrix((m.data, m.indices, m.indptr), shape=(m.shape[1], m.shape[0]))
        fws = FwdSolver(self, solver, tol, complex=False)
        fws.SetMode(mode)
        fws.SetPrecon("DIAG")
        fws.Allocate()
        fws.Reset(Solution(mua, mus, ref), 0.0)
        dphi = fws.CalcFields(qrows)
        aphi = fws.CalcFields(mrows)
        proj = fws.ProjectAll(mrows, dphi, scale="lin")
        Jm, _ = jacobian_cw(self, raster, dphi.T, aphi.T, np.real(proj), scale="log")
        return Jm

    def Gradient(self, raster, qvec, mvec, mua, mus, ref, freq, data, sd, solver="DIRECT", tol=1e-12, mode="mixed"):
        """Gradient of the data misfit sum(((data - proj) / sd)^2) with respect
        to cmua and ckappa in the raster's solution basis (toast.Gradient,
        mesh.py:369-376 -> GetGradient / AddDataGradient,
        toastmodule.cc:1745-1941): forward fields, log projections, one
        adjoint source per source from the weighted measurement vectors,
        all adjoint fields in one block solve, then the PMDF sums on the
        grid.  Returns (2 slen)."""
        q = sparse.csc_matrix(qvec)
        m = sparse.csc_matrix(mvec)
        qrows = sparse.csr_matrix((q.data, q.indices, q.indptr), shape=(q.shape[1], q.shape[0]))
        mrows = sparse.csr_matrix((m.data, m.indices, m.indptr), shape=(m.shape[1], m.shape[0]))
        fws = FwdSolver(self, solver, tol, complex=True)
        fws.SetMode(mode)
        fws.SetPrecon("DIAG")
        fws.Allocate()
        fws.Reset(Solution(mua, mus, ref), freq * 2.0 * np.pi * 1e-6)
        direct = solver.upper() == "DIRECT"

        def solve(rows):
            # DIRECT asks for the FP64 floor (relative residual 1e-14); a
            # column that stagnates just above it is as exact as FP64 allows
            phi = fws.CalcFields(rows, allow_notconv=direct)
            worst = max(s[2] for s in fws.stats)
            if direct and worst > 1e-11:
                raise L.NotConverged(L.TMI_ERR_NOT_CONVERGED, "relative residual %.3g above 1e-11" % worst)
            return phi

        dphi = solve(qrows)
        proj = fws.ProjectAllReal(mrows, dphi)          # [ln amp ; phase], link order
        cproj = fws.ProjectAll(mrows, dphi, scale="lin")
        data = np.ascontiguousarray(data, np.float64)
        sd = np.ascontiguousarray(sd, np.float64)
        nq, nm, nqm = self.qm_dims()
        if data.shape != (2 * nqm,) or sd.shape != (2 * nqm,):
            raise ValueError("data and sd must hold 2 nQM values")
        _, _, lp, li = self.QMPositions()
        # adjoint source weights (toastmodule.cc:1772-1808): per link, the
        # amplitude and phase terms over the linear projection
        b_mod = (data[:nqm] - proj[:nqm]) / sd[:nqm]
        b_arg = (data[nqm:] - proj[nqm:]) / sd[nqm:]
        rp, ip = cproj.real, cproj.imag
        dn = 1.0 / (rp * rp + ip * ip)
        ta = -2.0 * b_mod / (1.0 * sd[:nqm])
        tp = -2.0 * b_arg / (1.0 * sd[nqm:])
        wa = ta * rp * dn - 1j * (ta * ip * dn)
        wp = -tp * ip * dn - 1j * (tp * rp * dn)
        rows = np.repeat(np.arange(nq), np.diff(lp))
        W = sparse.csr_matrix((wa, (rows, li)), shape=(nq, nm)) + sparse.csr_matrix((wp, (rows, li)), shape=(nq, nm))
        wq = sparse.csr_matrix(W @ mrows)
        aphi = solve(wq)
        return raster.DataGradient(dphi.T, aphi.T)

    def JacobianOptical(self, raster, qvec, mvec, mua, mus, ref, freq=0, solver="bicgstab", tol=1e-12,
                        mode="mixed"):
        """Fields, adjoint fields, projections and Jacobian from optical
        parameters (toastmodule.cc:1514-1627).  mode selects the GPU solve
        ('mixed', 'fp64', 'fp64_block'), as in Fields."""
        from .jacobian import jacobian
        q = sparse.csc_matrix(qvec)
        m = sparse.csc_matrix(mvec)
        qrows = sparse.csr_matrix((q.data, q.indices, q.indptr), shape=(q.shape[1], q.shape[0]))
        mrows = sparse.csr_matrix((m.data, m.indices, m.indptr), shape=(m.shape[1], m.shape[0]))
        fws = FwdSolver(self, solver, tol, complex=True)
        fws.SetMode(mode)
        fws.Allocate()
        fws.Reset(Solution(mua, mus, ref), freq * 2.0 * np.pi * 1e-6)
        dphi = fws.CalcFields(qrows)
        aphi = fws.CalcFields(mrows)
        proj = fws.ProjectAll(mrows, dphi, scale="lin")
        return jacobian(self, raster, dphi.T, aphi.T, proj)


def ReadNim(nimname, idx=-1):
    """Image `idx` (0-based; -1 or past the end: the last) of a NIM file
    (toast.ReadNim, mesh.py:338-339; toastmodule.cc:905-940)."""
    size, nimg = C.c_int(), C.c_int()
    name = str(nimname).encode()
    L.call("tmi_read_nim", name, int(idx), C.byref(size), C.byref(nimg), None)
    img = np.zeros(size.value)
    L.call("tmi_read_nim", name, int(idx), None, None, L.ptr(img))
    return img


def WriteNim(nimname, meshname, nim):
    """One-image NIM file, values "%.12e" (toast.WriteNim, mesh.py:341-342;
    toastmodule.cc:942-972)."""
    v = np.ascontiguousarray(np.asarray(nim, np.float64).ravel())
    L.call("tmi_write_nim", str(nimname).encode(), str(meshname).encode(), v.size, 1, v, L.TMI_NIM_TOASTMOD)


# module-level forms of toast's functions (src/python/toast/mesh.py:344-376),
# taking a Mesh (and Raster) where the reference takes integer handles


def Sysmat_CW(mesh, mua, mus, ref, freq=0):
    """Real CW system matrix as scipy CSR (toast.Sysmat_CW, mesh.py:344-346;
    the reference ignores freq)."""
    return mesh.Sysmat_CW(mua, mus, ref)


def Jacobian(mesh, raster, dphi, aphi, proj):
    """Jacobian from given fields and projections (toast.Jacobian, mesh.py:348-349)."""
    return mesh.Jacobian(raster, dphi, aphi, proj)


def JacobianCW(mesh, raster, qvec, mvec, mua, mus, ref, solver="bicgstab", tol=1e-12):
    """CW Jacobian (toast.JacobianCW, mesh.py:352-366)."""
    return mesh.JacobianCW(raster, qvec, mvec, mua, mus, ref, solver=solver, tol=tol)


def Gradient(mesh, raster, qvec, mvec, mua, mus, ref, freq, data, sd):
    """Data misfit gradient (toast.Gradient, mesh.py:369-376)."""
    return mesh.Gradient(raster, qvec, mvec, mua, mus, ref, freq, data, sd)
