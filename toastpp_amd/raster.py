"""toast.Raster surface (src/python/toast/raster.py) on the GPU library.

Raster(mesh, grd, intgrd=None) builds Raster_Pixel (raster.cc:18-100,
raster_px.cc:10-54) on the GPU: basis grid grd, intermediate grid intgrd
(default grd), the element holding each grid point, the mesh -> grid map B,
the grid -> basis map G (Grid2LinPixMatrix) and its converse GI
(LinPix2GridMatrix) when the two grids differ, and the solution voxels.
Bases in map strings: M mesh nodes, G intermediate grid, B basis grid, S
solution voxels (basis voxels with mesh support).  Maps into the mesh basis
use the reference's NodeMapMatrix (mesh.cc:2695-2853), built on the GPU on
first use.
"""
import ctypes as C

import numpy as np
from scipy import sparse

from . import _lib as L


class Raster:
    def __init__(self, mesh, grd, intgrd=None):
        self._h = C.c_void_p()
        self.Make(mesh, grd, intgrd)

    # ------------------------------------------------------------ handle
    def Make(self, mesh, grd, intgrd=None):
        grd = np.ascontiguousarray(grd, np.int32)
        intgrd = grd.copy() if intgrd is None else np.ascontiguousarray(intgrd, np.int32)
        if len(grd) != mesh.Dim() or len(intgrd) != mesh.Dim():
            raise ValueError("grid dimensions %d/%d do not match the mesh dimension %d"
                             % (len(grd), len(intgrd), mesh.Dim()))
        self.Clear()
        self.mesh = mesh
        self.grd = grd
        self.intgrd = intgrd
        L.call("tmi_raster_create_basis", mesh.handle_ptr(), grd, intgrd, None, C.byref(self._h))
        g, s, b, hg = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        L.call("tmi_raster_dims", self._h, C.byref(g), C.byref(s))
        L.call("tmi_raster_basis_len", self._h, C.byref(b), C.byref(hg))
        self._glen, self._slen, self._blen, self._has_g = g.value, s.value, b.value, bool(hg.value)
        dim = len(grd)
        self.bbmin, self.bbmax = np.zeros(dim), np.zeros(dim)
        L.call("tmi_raster_bbox", self._h, self.bbmin, self.bbmax)

    def Handle(self):
        return int(self._h.value or 0)

    def handle_ptr(self):
        if not self._h.value:
            raise L.TmiError(L.TMI_ERR_INVALID_ARG, "raster has no data (handle 0)")
        return self._h

    def Clear(self):
        if getattr(self, "_h", None) is not None and self._h.value and L._lib is not None:
            L._lib.tmi_raster_free(self._h)
        self._h = C.c_void_p()

    def __del__(self):
        self.Clear()

    # ------------------------------------------------------------- sizes
    def GLen(self):
        return self._glen

    def BLen(self):
        return self._blen

    def SLen(self):
        return self._slen

    def Elref(self):
        e = np.zeros(self._glen, np.int32)
        L.call("tmi_raster_get_elref", self.handle_ptr(), e.ctypes.data_as(C.c_void_p), None)
        return e

    def Sol2Basis(self):
        s = np.zeros(self._slen, np.int32)
        L.call("tmi_raster_get_elref", self.handle_ptr(), None, s.ctypes.data_as(C.c_void_p))
        return s

    def Basis2Sol(self):
        b = np.full(self._blen, -1, np.int32)
        b[self.Sol2Basis()] = np.arange(self._slen, dtype=np.int32)
        return b

    def BasisPoints(self):
        """Basis voxel coordinates (blen x dim), x fastest (BasisVoxelPositions)."""
        dim = len(self.grd)
        axes = [self.bbmin[d] + np.arange(self.grd[d]) * ((self.bbmax[d] - self.bbmin[d]) / (self.grd[d] - 1))
                for d in range(dim)]
        grids = np.meshgrid(*axes[::-1], indexing="ij")
        return np.stack([g.ravel() for g in grids[::-1]], axis=1)

    def SolutionPoints(self):
        return self.BasisPoints()[self.Sol2Basis()]

    # ------------------------------------------------------------ mapping
    def _bmap(self):
        nn = len(self.grd) + 1
        cols = np.zeros(self._glen * nn, np.int32)
        w = np.zeros(self._glen * nn)
        L.call("tmi_raster_get_bmap", self.handle_ptr(), cols.ctypes.data_as(C.c_void_p),
               w.ctypes.data_as(C.c_void_p))
        return cols.reshape(self._glen, nn), w.reshape(self._glen, nn)

    def MeshToGrid(self, mvec):
        """Map_MeshToGrid (raster.cc:188-199) of one or more node-major
        columns (n or n x k), complex or real."""
        mvec = np.asarray(mvec)
        real = not np.iscomplexobj(mvec)
        one = mvec.ndim == 1
        m = np.ascontiguousarray(mvec.reshape(mvec.shape[0], -1), np.complex128)
        out = np.zeros((self._glen, m.shape[1]), np.complex128)
        L.call("tmi_raster_map_mesh_to_grid", self.handle_ptr(), m.shape[1], m.view(np.float64),
               out.view(np.float64))
        if real:
            out = out.real.copy()
        return out[:, 0] if one else out

    def _cols(self, vec, rows):
        vec = np.asarray(vec)
        real = not np.iscomplexobj(vec)
        one = vec.ndim == 1
        if vec.shape[0] != rows:
            raise ValueError("expected %d rows, got %d" % (rows, vec.shape[0]))
        return np.ascontiguousarray(vec.reshape(rows, -1), np.complex128), real, one

    def GridToBasis(self, gvec):
        """Map_GridToBasis (raster_px.cc:102-114): G applied on the GPU (a
        copy when the grids coincide)."""
        g, real, one = self._cols(gvec, self._glen)
        out = np.zeros((self._blen, g.shape[1]), np.complex128)
        L.call("tmi_raster_map_grid_to_basis", self.handle_ptr(), g.shape[1], g.view(np.float64),
               out.view(np.float64))
        if real:
            out = out.real.copy()
        return out[:, 0] if one else out

    def BasisToGrid(self, bvec):
        """Map_BasisToGrid (raster_px.cc:118-130): GI applied on the GPU."""
        b, real, one = self._cols(bvec, self._blen)
        out = np.zeros((self._glen, b.shape[1]), np.complex128)
        L.call("tmi_raster_map_basis_to_grid", self.handle_ptr(), b.shape[1], b.view(np.float64),
               out.view(np.float64))
        if real:
            out = out.real.copy()
        return out[:, 0] if one else out

    def MeshToBasis(self, mvec):
        return self.GridToBasis(self.MeshToGrid(mvec))

    def MeshToSol(self, mvec):
        return self.MeshToBasis(mvec)[self.Sol2Basis()]

    def _nodemap(self):
        nc = 1 << len(self.grd)
        n = self.mesh.NodeCount()
        cols = np.zeros(n * nc, np.int32)
        w = np.zeros(n * nc)
        L.call("tmi_raster_get_nodemap", self.handle_ptr(), cols.ctypes.data_as(C.c_void_p),
               w.ctypes.data_as(C.c_void_p))
        return cols.reshape(n, nc), w.reshape(n, nc)

    def GridToMesh(self, gvec):
        """Map_GridToMesh (raster.cc:212-224) of one or more grid columns
        (glen or glen x k), complex or real."""
        gvec = np.asarray(gvec)
        real = not np.iscomplexobj(gvec)
        one = gvec.ndim == 1
        g = np.ascontiguousarray(gvec.reshape(gvec.shape[0], -1), np.complex128)
        out = np.zeros((self.mesh.NodeCount(), g.shape[1]), np.complex128)
        L.call("tmi_raster_map_grid_to_mesh", self.handle_ptr(), g.shape[1], g.view(np.float64),
               out.view(np.float64))
        if real:
            out = out.real.copy()
        return out[:, 0] if one else out

    # the bases in the order the maps chain them: M <-> G <-> B <-> S
    _CHAIN = "MGBS"

    def _step(self, src, tgt, v):
        if src == "M" and tgt == "G":
            return self.MeshToGrid(v)
        if src == "G" and tgt == "M":
            return self.GridToMesh(v)
        if src == "G" and tgt == "B":
            return self.GridToBasis(v)
        if src == "B" and tgt == "G":
            return self.BasisToGrid(v)
        if src == "B" and tgt == "S":
            return v[self.Sol2Basis()]
        out = np.zeros((self._blen,) + v.shape[1:], v.dtype)   # S -> B
        out[self.Sol2Basis()] = v
        return out

    def Map(self, mapstr, srcvec):
        """Map a field between bases: mapstr 'S->T' with S, T in M, G, B, S
        (Raster::Map_*, raster.cc:188-388, raster_px.cc:102-176)."""
        src, tgt = mapstr.replace(" ", "").split("->")
        if src not in self._CHAIN or tgt not in self._CHAIN:
            raise ValueError("invalid map string %r" % mapstr)
        v = np.asarray(srcvec).ravel()
        i, j = self._CHAIN.index(src), self._CHAIN.index(tgt)
        if i == j:
            return v.copy()
        step = 1 if j > i else -1
        for k in range(i, j, step):
            v = self._step(self._CHAIN[k], self._CHAIN[k + step], v)
        return v

    def GetMatrix(self, mapstr):
        """Sparse map matrices (toast_raster_matrix, toastmodule.cc:416-507):
        M->G (B), G->M (BI), G->B (G), B->G (GI), M->B (C = G B), B->M
        (CI = BI GI), B->S (D); identity for G<->B when the grids coincide."""
        src, tgt = mapstr.replace(" ", "").split("->")
        sizes = {"S": self._slen, "G": self._glen, "B": self._blen, "M": self.mesh.NodeCount()}
        if src == tgt:
            return sparse.identity(sizes[src], format="csr")
        if src in ("G", "B") and tgt in ("G", "B") and not self._has_g:
            return sparse.identity(self._glen, format="csr")
        if src == "M" and tgt == "G":
            cols, w = self._bmap()
            keep = cols >= 0
            rows = np.repeat(np.arange(self._glen), cols.shape[1]).reshape(cols.shape)
            return sparse.csr_matrix((w[keep], (rows[keep], cols[keep])), shape=(self._glen, sizes["M"]))
        if src == "G" and tgt == "M":
            cols, w = self._nodemap()
            keep = cols >= 0
            rows = np.repeat(np.arange(sizes["M"]), cols.shape[1]).reshape(cols.shape)
            return sparse.csr_matrix((w[keep], (rows[keep], cols[keep])), shape=(sizes["M"], self._glen))
        if src == "G" and tgt == "B":
            return self._csr("tmi_raster_get_gmap", self._blen, self._glen)
        if src == "B" and tgt == "G":
            return self._csr("tmi_raster_get_gimap", self._glen, self._blen)
        if src == "M" and tgt == "B":
            B = self.GetMatrix("M->G")
            return B if not self._has_g else sparse.csr_matrix(self.GetMatrix("G->B") @ B)
        if src == "B" and tgt == "M":
            BI = self.GetMatrix("G->M")
            return BI if not self._has_g else sparse.csr_matrix(BI @ self.GetMatrix("B->G"))
        if src == "B" and tgt == "S":
            s2b = self.Sol2Basis()
            return sparse.csr_matrix((np.ones(self._slen), (np.arange(self._slen), s2b)),
                                     shape=(self._slen, self._blen))
        raise L.TmiError(L.TMI_ERR_UNSUPPORTED, "matrix %s is not provided" % mapstr)

    def _csr(self, fn, rows, cols):
        nnz = C.c_int()
        rp = np.zeros(rows + 1, np.int32)
        L.call(fn, self.handle_ptr(), L.ptr(rp), None, None, C.byref(nnz))
        ci = np.zeros(max(nnz.value, 1), np.int32)
        val = np.zeros(max(nnz.value, 1))
        L.call(fn, self.handle_ptr(), L.ptr(rp), L.ptr(ci), L.ptr(val), C.byref(nnz))
        n = nnz.value
        return sparse.csr_matrix((val[:n], ci[:n], rp), shape=(rows, cols))

    # ----------------------------------------------------------- Jacobian
    def DataGradient(self, dphi, aphi, grad=None):
        """AddDataGradient's raster part (toastmodule.cc:1813-1836): dphi and
        aphi (nlen x nq) complex, column q of aphi the adjoint field of source
        q; returns grad (2 slen: cmua, ckappa) decremented by the PMDF sums
        of every source."""
        dphi = np.ascontiguousarray(dphi, np.complex128)
        aphi = np.ascontiguousarray(aphi, np.complex128)
        if dphi.shape != aphi.shape or dphi.shape[0] != self.mesh.NodeCount():
            raise ValueError("direct and adjoint fields must both be (nlen x nq)")
        g = np.zeros(2 * self._slen) if grad is None else np.array(grad, np.float64)
        L.call("tmi_raster_data_gradient", self.handle_ptr(), dphi.shape[1], dphi.view(np.float64),
               aphi.view(np.float64), g)
        return g

    def Jacobian(self, dphi, aphi, proj, dscale=L.DATA_LOG):
        """GenerateJacobian_grid (jacobian.cc:293-373): (2 nQM x 2 slen)."""
        nq, nm, nqm = self.mesh.qm_dims()
        J = np.empty((2 * nqm, 2 * self._slen), np.float64)
        L.call("tmi_jacobian_raster", self.handle_ptr(), nq, nm, dphi.view(np.float64), aphi.view(np.float64),
               proj.view(np.float64), dscale, J)
        return J
