"""toast.Raster surface (src/python/toast/raster.py) on the GPU library.

Raster(mesh, grd) builds the regular-grid basis of Raster_Pixel with the
basis grid equal to the intermediate grid (raster.cc:18-100,
raster_px.cc:10-54) on the GPU: the element holding each grid point, the
mesh -> grid map B and the solution mask.  Bases in map strings: M mesh
nodes, G intermediate grid, B basis grid (= G here), S solution voxels
(basis voxels inside the mesh).  Maps into the mesh basis use the
reference's NodeMapMatrix (mesh.cc:2695-2853), built on the GPU on first use.
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
        if intgrd is not None and not np.array_equal(np.asarray(intgrd, np.int32), grd):
            raise L.TmiError(L.TMI_ERR_UNSUPPORTED, "only an intermediate grid equal to the basis grid is supported")
        if len(grd) != mesh.Dim():
            raise ValueError("grid dimension %d does not match the mesh dimension %d" % (len(grd), mesh.Dim()))
        self.Clear()
        self.mesh = mesh
        self.grd = grd
        L.call("tmi_raster_create", mesh.handle_ptr(), grd, None, C.byref(self._h))
        g, s = C.c_int(), C.c_int()
        L.call("tmi_raster_dims", self._h, C.byref(g), C.byref(s))
        self._glen, self._slen = g.value, s.value
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
        return self._glen

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
        b = np.full(self._glen, -1, np.int32)
        b[self.Sol2Basis()] = np.arange(self._slen, dtype=np.int32)
        return b

    def BasisPoints(self):
        """Grid point coordinates (blen x dim), x fastest (GenerateVoxelPositions)."""
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

    def MeshToSol(self, mvec):
        g = self.MeshToGrid(mvec)
        return g[self.Sol2Basis()]

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

    def Map(self, mapstr, srcvec):
        """Map a field between bases: mapstr 'S->T' with S, T in M, G, B, S."""
        src, tgt = mapstr.replace(" ", "").split("->")
        v = np.asarray(srcvec).ravel()
        if tgt == "M":
            if src == "M":
                return v.copy()
            if src in ("G", "B"):
                return self.GridToMesh(v)
            if src == "S":
                return self.GridToMesh(self.Map("S->B", v))
        if src == "M":
            g = self.MeshToGrid(v)
            return g[self.Sol2Basis()] if tgt == "S" else g
        if src in ("G", "B"):
            return v.copy() if tgt in ("G", "B") else v[self.Sol2Basis()]
        if src == "S":
            if tgt == "S":
                return v.copy()
            out = np.zeros(self._glen, v.dtype)
            out[self.Sol2Basis()] = v
            return out
        raise ValueError("invalid map string %r" % mapstr)

    def GetMatrix(self, mapstr):
        src, tgt = mapstr.replace(" ", "").split("->")
        sizes = {"S": self._slen, "G": self._glen, "B": self._glen, "M": self.mesh.NodeCount()}
        if src in ("G", "B") and tgt in ("G", "B"):
            return sparse.identity(self._glen)
        if src == "M" and tgt in ("G", "B"):
            cols, w = self._bmap()
            keep = cols >= 0
            rows = np.repeat(np.arange(self._glen), cols.shape[1]).reshape(cols.shape)
            return sparse.csr_matrix((w[keep], (rows[keep], cols[keep])), shape=(self._glen, sizes["M"]))
        if src in ("G", "B") and tgt == "M":
            cols, w = self._nodemap()
            keep = cols >= 0
            rows = np.repeat(np.arange(sizes["M"]), cols.shape[1]).reshape(cols.shape)
            return sparse.csr_matrix((w[keep], (rows[keep], cols[keep])), shape=(sizes["M"], self._glen))
        if src in ("G", "B") and tgt == "S":
            s2b = self.Sol2Basis()
            return sparse.csr_matrix((np.ones(self._slen), (np.arange(self._slen), s2b)),
                                     shape=(self._slen, self._glen))
        raise L.TmiError(L.TMI_ERR_UNSUPPORTED, "matrix %s is not provided" % mapstr)

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
