"""ctypes front end of the FP64 CPU oracle (oracle/toast_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and only as the checker.  The product package
toastpp_amd never imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
U8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="C_CONTIGUOUS")
P = C.c_void_p


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        sig = {
            "om_mesh_load": (P, [C.c_char_p]),
            "om_mesh_make": (P, [C.c_int, C.c_int, D, C.c_int, I]),
            "om_mesh_free": (None, [P]),
            "om_setup": (None, [P]),
            "om_dim": (C.c_int, [P]), "om_nlen": (C.c_int, [P]), "om_elen": (C.c_int, [P]),
            "om_get_nodes": (None, [P, D]), "om_get_elements": (None, [P, I]),
            "om_get_bnd": (None, [P, I]), "om_get_regions": (None, [P, I]),
            "om_get_size": (None, [P, D]), "om_get_bndside": (None, [P, U8]),
            "om_get_intbff": (None, [P, D]), "om_get_intdd": (None, [P, D]),
            "om_csr_nnz": (C.c_int, [P]), "om_csr": (None, [P, I, I]),
            "om_assemble": (None, [P, I, I, D, D, D, C.c_double, D]),
            "om_assemble_ff": (None, [P, I, I, D]),
            "om_assemble_el": (None, [P, I, I, D, D, D, C.c_double, D]),
            "om_jacobian_el": (None, [P, C.c_int, C.c_int, D, D, I, I, D, C.c_int, C.c_int, D]),
            "om_el_intfff": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
            "om_el_intfdd": (C.c_double, [P, C.c_int, C.c_int, C.c_int, C.c_int]),
            "om_el_intpff": (C.c_double, [P, C.c_int, C.c_int, C.c_int, D]),
            "om_el_bndintpff": (C.c_double, [P, C.c_int, C.c_int, C.c_int, D]),
            "om_a_keijzer": (C.c_double, [C.c_double]),
            "om_coeffs": (None, [C.c_int, D, D, D, D, D, D]),
            "om_qvec_gaussian": (None, [P, D, C.c_double, C.c_int, D]),
            "om_qm_load": (P, [P, C.c_char_p]),
            "om_qm_nq": (C.c_int, [P]), "om_qm_nm": (C.c_int, [P]), "om_qm_nqm": (C.c_int, [P]),
            "om_qm_get": (None, [P, D, D, I, I]), "om_qm_free": (None, [P]),
            "om_pull_to_boundary": (C.c_int, [P, D, D]),
            "om_bbox": (None, [P, D, D]),
            "om_qvec_point": (None, [P, D, C.c_int, D]),
            "om_raster_elref": (None, [P, I, D, D, I]),
            "om_raster_bmap": (None, [P, I, D, D, I, I, D]),
            "om_image_gradient": (None, [C.c_int, I, D, D, I, D]),
            "om_jacobian_grid": (None, [P, I, D, D, I, C.c_int, C.c_int, D, D, I, I, D, C.c_int, D]),
            "om_pcg": (C.c_int, [C.c_int, I, I, D, D, D, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double)]),
            "om_bicgstab_c": (C.c_int, [C.c_int, I, I, D, D, D, C.c_double, C.c_int, C.c_int, C.POINTER(C.c_double)]),
            "om_relres_c": (C.c_double, [C.c_int, I, I, D, D, D]),
            "om_project": (None, [C.c_int, C.c_int, C.c_int, D, D, I, I, C.c_int, D]),
            "om_jacobian_nodal": (None, [P, C.c_int, C.c_int, D, D, I, I, D, C.c_int, D]),
            "om_jacobian_grid_cw": (None, [P, I, D, D, I, C.c_int, C.c_int, D, D, I, I, P, C.c_int, P, P]),
            "om_data_gradient": (None, [P, I, D, D, I, C.c_int, D, D, D]),
            "om_raster_nodemap": (None, [P, I, D, D, I, I, D]),
            "om_subsample_linpixel": (None, [C.c_int, I, I, P, D, D]),
            "om_grid2linpix": (None, [C.c_int, I, I, P, I, P, P]),
            "om_linpix2grid": (None, [C.c_int, I, I, P, I, P, P]),
            "om_raster_slen": (C.c_long, [C.c_int, I, I, D, D, I, P]),
            "om_jacobian_grid_b": (None, [P, I, I, D, D, I, C.c_int, C.c_int, D, D, I, I, D, C.c_int, D]),
            "om_jacobian_grid_cw_b": (None, [P, I, I, D, D, I, C.c_int, C.c_int, D, D, I, I, P, C.c_int, P, P]),
            "om_data_gradient_b": (None, [P, I, I, D, D, I, C.c_int, D, D, D]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def _cplx(a):
    """complex128 array -> interleaved float64 view (contiguous)"""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)


class Mesh:
    """Oracle mesh: Mesh::operator>> + Mesh::Setup (mesh.cc:96-158, 1787-1825)."""

    def __init__(self, path=None, nodes=None, elems=None):
        L = lib()
        if path is not None:
            self.h = L.om_mesh_load(path.encode())
            if not self.h:
                raise IOError("oracle: cannot read mesh " + path)
        else:
            nodes = _c(nodes, np.float64)
            elems = _c(elems, np.int32)
            self.h = L.om_mesh_make(nodes.shape[1], nodes.shape[0], nodes, elems.shape[0], elems)
        L.om_setup(self.h)
        self.dim = L.om_dim(self.h)
        self.n = L.om_nlen(self.h)
        self.elen = L.om_elen(self.h)
        self.nn = self.dim + 1

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.om_mesh_free(self.h)
            self.h = None

    def nodes(self):
        x = np.zeros((self.n, self.dim)); lib().om_get_nodes(self.h, x); return x

    def elements(self):
        e = np.zeros((self.elen, self.nn), np.int32); lib().om_get_elements(self.h, e); return e

    def bnd(self):
        b = np.zeros(self.n, np.int32); lib().om_get_bnd(self.h, b); return b

    def regions(self):
        r = np.zeros(self.n, np.int32); lib().om_get_regions(self.h, r); return r

    def size(self):
        s = np.zeros(self.elen); lib().om_get_size(self.h, s); return s

    def bndside(self):
        b = np.zeros((self.elen, self.nn), np.uint8); lib().om_get_bndside(self.h, b); return b

    def intbff(self):
        b = np.zeros((self.elen, self.nn, self.nn)); lib().om_get_intbff(self.h, b); return b

    def intdd(self):
        b = np.zeros((self.elen, self.nn, self.nn)); lib().om_get_intdd(self.h, b); return b

    def csr(self):
        L = lib()
        nnz = L.om_csr_nnz(self.h)
        rp = np.zeros(self.n + 1, np.int32); ci = np.zeros(nnz, np.int32)
        L.om_csr(self.h, rp, ci)
        return rp, ci

    def sysmat(self, cmua, ckap, c2a, omega):
        rp, ci = self.csr()
        val = np.zeros(2 * len(ci))
        lib().om_assemble(self.h, rp, ci, _c(cmua, np.float64), _c(ckap, np.float64), _c(c2a, np.float64),
                          float(omega), val)
        return rp, ci, val.view(np.complex128)

    def sysmat_el(self, cmua, ckap, c2a, omega):
        """AssembleSystemMatrix with elbasis = true: elen coefficients each."""
        rp, ci = self.csr()
        val = np.zeros(2 * len(ci))
        lib().om_assemble_el(self.h, rp, ci, _c(cmua, np.float64), _c(ckap, np.float64), _c(c2a, np.float64),
                             float(omega), val)
        return rp, ci, val.view(np.complex128)

    def massmat(self):
        rp, ci = self.csr()
        val = np.zeros(len(ci))
        lib().om_assemble_ff(self.h, rp, ci, val)
        return rp, ci, val

    def qvec_gaussian(self, centres, width, neumann=True):
        centres = np.atleast_2d(_c(centres, np.float64))
        out = np.zeros((centres.shape[0], self.n))
        for i, c in enumerate(centres):
            q = np.zeros(self.n)
            lib().om_qvec_gaussian(self.h, _c(c, np.float64), float(width), 1 if neumann else 0, q)
            out[i] = q
        return out

    def read_qm(self, path):
        L = lib()
        h = L.om_qm_load(self.h, path.encode())
        if not h:
            raise IOError("oracle: cannot read QM " + path)
        nq, nm, nqm = L.om_qm_nq(h), L.om_qm_nm(h), L.om_qm_nqm(h)
        Q = np.zeros((nq, self.dim)); M = np.zeros((nm, self.dim))
        lp = np.zeros(nq + 1, np.int32); li = np.zeros(nqm, np.int32)
        L.om_qm_get(h, Q, M, lp, li)
        L.om_qm_free(h)
        return Q, M, lp, li


def a_keijzer(n):
    return lib().om_a_keijzer(float(n))


def coeffs(mua, mus, ref):
    mua, mus, ref = (_c(a, np.float64) for a in (mua, mus, ref))
    n = len(mua)
    cm, ck, c2 = np.zeros(n), np.zeros(n), np.zeros(n)
    lib().om_coeffs(n, mua, mus, ref, cm, ck, c2)
    return cm, ck, c2


def pcg(rp, ci, val, b, tol, maxit=0, diag=False, x0=None):
    n = len(rp) - 1
    x = np.zeros(n) if x0 is None else _c(x0, np.float64).copy()
    err = C.c_double(0)
    it = lib().om_pcg(n, _c(rp, np.int32), _c(ci, np.int32), _c(val, np.float64), _c(b, np.float64), x,
                      float(tol), int(maxit), 1 if diag else 0, C.byref(err))
    return x, it, err.value


def bicgstab(rp, ci, val, b, tol, maxit=0, diag=False, x0=None):
    n = len(rp) - 1
    x = np.zeros(n, np.complex128) if x0 is None else np.array(x0, np.complex128)
    xv = x.view(np.float64)
    err = C.c_double(0)
    it = lib().om_bicgstab_c(n, _c(rp, np.int32), _c(ci, np.int32), _cplx(val), _cplx(b), xv,
                             float(tol), int(maxit), 1 if diag else 0, C.byref(err))
    return x, it, err.value


def gmres(rp, ci, val, b, tol, maxit=0, diag=False, x0=None, restart=20):
    """GMRES(restart) of gmres_imp.hpp:41-230, one right-hand side, as
    IterativeSolve runs it (gsmatrix_imp.hpp:477-478, 590-593): left
    diagonal preconditioning (Mb = D^-1 b), modified Gram-Schmidt with
    H(j,k) = sum_n w_n conj(v_k,n), the reference's rotation (for complex H
    r = sqrt(H(j,j)^2 + H(j,j+1)^2) without conjugation), back substitution
    and x += y_l v_l for l = j..0; the cycle limit is maxit (n+1 if 0).
    Returns (x, cycles, ||M^-1 (b - A x)|| / ||M^-1 b||)."""
    from scipy import sparse
    n = len(rp) - 1
    cplx = np.iscomplexobj(val) or np.iscomplexobj(b)
    dt = np.complex128 if cplx else np.float64
    A = sparse.csr_matrix((np.asarray(val, dt), np.asarray(ci), np.asarray(rp)), shape=(n, n))
    idg = 1.0 / A.diagonal() if diag else None
    pre = (lambda v: v * idg) if diag else (lambda v: v)
    b = np.asarray(b, dt)
    x = np.zeros(n, dt) if x0 is None else np.array(x0, dt)
    max_cycle = maxit if maxit else n + 1
    Mb = pre(b)
    nb = np.sqrt(np.sum(np.abs(Mb) ** 2))
    v0 = Mb.copy() if np.sqrt(np.sum(np.abs(x) ** 2)) < tol * n else Mb - pre(A @ x)
    cycle = 0
    R = restart
    while True:
        nv = np.sqrt(np.sum(np.abs(v0) ** 2))
        if not (nv > tol * nb and cycle < max_cycle):
            break
        s = np.zeros(R + 1, dt)
        s[0] = nv
        V = [v0 / nv]
        H = np.zeros((R, R + 1), dt)
        c1 = np.zeros(R, dt)
        c2 = np.zeros(R, dt)
        j = -1
        cycle += 1
        while abs(s[j + 1]) > tol * nb and j < R - 1 and cycle < max_cycle:
            j += 1
            w = pre(A @ V[j])
            for k in range(j + 1):
                h = np.sum(w * np.conj(V[k]))
                H[j, k] = h
                w = w - h * V[k]
            t = np.sqrt(np.sum(np.abs(w) ** 2))
            H[j, j + 1] = t
            V.append(w * (1.0 / t))
            for k in range(j):
                h1 = c1[k] * H[j, k] + c2[k] * H[j, k + 1]
                h2 = -c2[k] * H[j, k] + c1[k] * H[j, k + 1]
                H[j, k], H[j, k + 1] = h1, h2
            r = np.sqrt(H[j, j] * H[j, j] + H[j, j + 1] * H[j, j + 1])
            c1[j] = H[j, j] / r
            c2[j] = H[j, j + 1] / r
            H[j, j] = r
            H[j, j + 1] = 0.0
            s[j + 1] = -c2[j] * s[j]
            s[j] = s[j] * c1[j]
        y = np.zeros(R, dt)
        for l in range(j, -1, -1):
            y[l] = s[l] / H[l, l]
            for k in range(l - 1, -1, -1):
                s[k] -= H[l, k] * y[l]
        for l in range(j, -1, -1):
            x = x + y[l] * V[l]
        v0 = Mb - pre(A @ x)
    return x, cycle, (nv / nb if nb > 0 else 0.0)


def relres(rp, ci, val, b, x):
    return lib().om_relres_c(len(rp) - 1, _c(rp, np.int32), _c(ci, np.int32), _cplx(val), _cplx(b), _cplx(x))


def project(phi, mvec, lptr, lidx, logscale=True):
    """phi: (nq, n) complex, mvec: (nm, n) complex dense."""
    phi = np.ascontiguousarray(phi, np.complex128); mvec = np.ascontiguousarray(mvec, np.complex128)
    nq, n = phi.shape
    out = np.zeros(2 * int(lptr[-1]))
    lib().om_project(n, nq, mvec.shape[0], phi.view(np.float64), mvec.view(np.float64), _c(lptr, np.int32),
                     _c(lidx, np.int32), 1 if logscale else 0, out)
    return out.view(np.complex128)


def jacobian_nodal(mesh, dphi, aphi, lptr, lidx, proj, logscale=True):
    dphi = np.ascontiguousarray(dphi, np.complex128); aphi = np.ascontiguousarray(aphi, np.complex128)
    nq, nm = dphi.shape[0], aphi.shape[0]
    nqm = int(lptr[-1])
    J = np.zeros((2 * nqm, 2 * mesh.n))
    lib().om_jacobian_nodal(mesh.h, nq, nm, dphi.view(np.float64), aphi.view(np.float64), _c(lptr, np.int32),
                            _c(lidx, np.int32), _cplx(proj), 1 if logscale else 0, J)
    return J


def jacobian_el(mesh, dphi, aphi, lptr, lidx, proj, logscale=True, mvec_offset=False):
    """Element-basis Jacobian (2 nQM x 2 elen); mvec_offset reproduces the
    mvec overload's kappa view at offset n (jacobian.cc:397)."""
    dphi = np.ascontiguousarray(dphi, np.complex128); aphi = np.ascontiguousarray(aphi, np.complex128)
    nq, nm = dphi.shape[0], aphi.shape[0]
    nqm = int(lptr[-1])
    J = np.zeros((2 * nqm, 2 * mesh.elen))
    lib().om_jacobian_el(mesh.h, nq, nm, dphi.view(np.float64), aphi.view(np.float64), _c(lptr, np.int32),
                         _c(lidx, np.int32), _cplx(proj), 1 if logscale else 0, 1 if mvec_offset else 0, J)
    return J


def jacobian_cw_el(mesh, dphi, aphi, lptr, lidx, proj=None, logscale=True):
    """GenerateJacobian_cw_mesh with elbasis = true (jacobian.cc:727-790) +
    the log row scaling: -IntFG_el / proj and -IntGradFGradG_el / proj on
    real fields.  Returns (Jmua, Jkap), each nqm x elen."""
    dphi = np.asarray(dphi, np.float64).astype(np.complex128)
    aphi = np.asarray(aphi, np.float64).astype(np.complex128)
    nqm = int(lptr[-1])
    p = np.ones(nqm) if proj is None else np.asarray(proj, np.float64)
    J = jacobian_el(mesh, dphi, aphi, lptr, lidx, p.astype(np.complex128), logscale)
    return J[:nqm, :mesh.elen].copy(), J[:nqm, mesh.elen:].copy()


def jacobian_cw_nodal(mesh, dphi, aphi, lptr, lidx, proj=None, logscale=True):
    """GenerateJacobian_cw_mesh (jacobian.cc:727-790) + the log row scaling
    (:160-186) for real fields (nq x n, nm x n): the element-integral loops of
    om_jacobian_nodal on the fields with zero imaginary parts, whose real
    blocks are -IntFG/proj and -IntGradFGradG/proj.  Returns (Jmua, Jkap),
    each nqm x n."""
    dphi = np.asarray(dphi, np.float64).astype(np.complex128)
    aphi = np.asarray(aphi, np.float64).astype(np.complex128)
    nqm = int(lptr[-1])
    p = np.ones(nqm) if proj is None else np.asarray(proj, np.float64)
    J = jacobian_nodal(mesh, dphi, aphi, lptr, lidx, p.astype(np.complex128), logscale)
    return J[:nqm, :mesh.n].copy(), J[:nqm, mesh.n:].copy()


def qvec_point(mesh, pts, neumann=True):
    """QVec_Point for each row of pts -> (npts x n)."""
    pts = np.atleast_2d(np.asarray(pts, np.float64))
    out = np.zeros((pts.shape[0], mesh.n))
    for i, p in enumerate(pts):
        row = np.zeros(mesh.n)
        lib().om_qvec_point(mesh.h, np.ascontiguousarray(p), 1 if neumann else 0, row)
        out[i] = row
    return out


class Raster:
    """Raster_Pixel (raster.cc:18-100, raster_px.cc:10-54) restated:
    bounding box of the mesh, element of every grid point, the grid map B,
    and the solution voxels: grid points inside the mesh when the basis grid
    bdim equals the intermediate grid gdim, else basis voxels whose linear
    pixel support holds a grid point inside the mesh (SubsampleLinPixel of
    the inside mask), with G = Grid2LinPixMatrix mapping grid to basis."""

    def __init__(self, mesh, gdim, bb=None, bdim=None):
        self.mesh = mesh
        self.dim = mesh.dim
        self.gdim = np.ascontiguousarray(gdim, np.int32)
        self.bdim = self.gdim.copy() if bdim is None else np.ascontiguousarray(bdim, np.int32)
        if bb is None:
            lo, hi = np.zeros(self.dim), np.zeros(self.dim)
            lib().om_bbox(mesh.h, lo, hi)
        else:
            lo, hi = np.ascontiguousarray(bb[0], np.float64), np.ascontiguousarray(bb[1], np.float64)
        self.bbmin, self.bbmax = lo, hi
        self.glen = int(np.prod(self.gdim))
        self.blen = int(np.prod(self.bdim))
        self.elref = np.zeros(self.glen, np.int32)
        lib().om_raster_elref(mesh.h, self.gdim, lo, hi, self.elref)
        n = lib().om_raster_slen(self.dim, self.gdim, self.bdim, lo, hi, self.elref, None)
        s2b = np.zeros(max(n, 1), np.int64)
        lib().om_raster_slen(self.dim, self.gdim, self.bdim, lo, hi, self.elref, s2b.ctypes.data_as(C.c_void_p))
        self.sol2basis = s2b[:n].astype(np.int32)
        self.slen = n

    def _csr(self, fn, nrows, ncols, masked=True):
        rp = np.zeros(nrows + 1, np.int32)
        el = self.elref.ctypes.data_as(C.c_void_p) if masked else None
        fn(self.dim, self.gdim, self.bdim, el, rp, None, None)
        ci = np.zeros(max(rp[-1], 1), np.int32)
        val = np.zeros(max(rp[-1], 1))
        fn(self.dim, self.gdim, self.bdim, el, rp, ci.ctypes.data_as(C.c_void_p), val.ctypes.data_as(C.c_void_p))
        return rp, ci[:rp[-1]], val[:rp[-1]]

    def gmap(self):
        """G = Grid2LinPixMatrix (blen x glen) as (rowptr, colidx, val)."""
        return self._csr(lib().om_grid2linpix, self.blen, self.glen)

    def gimap(self):
        """GI = LinPix2GridMatrix (glen x blen) as (rowptr, colidx, val)."""
        return self._csr(lib().om_linpix2grid, self.glen, self.blen)

    def subsample(self, gf, masked=False):
        """SubsampleLinPixel of a grid image (elref mask optional)."""
        bf = np.zeros(self.blen)
        el = self.elref.ctypes.data_as(C.c_void_p) if masked else None
        lib().om_subsample_linpixel(self.dim, self.gdim, self.bdim, el, _c(gf, np.float64), bf)
        return bf

    def bmap(self):
        nn = self.dim + 1
        cols = np.zeros(self.glen * nn, np.int32)
        w = np.zeros(self.glen * nn)
        lib().om_raster_bmap(self.mesh.h, self.gdim, self.bbmin, self.bbmax, self.elref, cols, w)
        return cols.reshape(self.glen, nn), w.reshape(self.glen, nn)

    def gsize(self):
        return self.bbmax - self.bbmin

    def nodemap(self):
        nc = 1 << self.dim
        cols = np.zeros(self.mesh.n * nc, np.int32)
        w = np.zeros(self.mesh.n * nc)
        lib().om_raster_nodemap(self.mesh.h, self.gdim, self.bbmin, self.bbmax, self.elref, cols, w)
        return cols.reshape(self.mesh.n, nc), w.reshape(self.mesh.n, nc)


def image_gradient(gdim, gsize, im, mask):
    gdim = np.ascontiguousarray(gdim, np.int32)
    im = np.ascontiguousarray(im, np.complex128)
    out = np.zeros(len(gdim) * im.size, np.complex128)
    lib().om_image_gradient(len(gdim), gdim, np.ascontiguousarray(gsize, np.float64), im.view(np.float64),
                            np.ascontiguousarray(mask, np.int32), out.view(np.float64))
    return out.reshape(len(gdim), im.size)


def jacobian_grid(raster, dphi, aphi, lptr, lidx, proj, logscale=True):
    dphi = np.ascontiguousarray(dphi, np.complex128); aphi = np.ascontiguousarray(aphi, np.complex128)
    nq, nm = dphi.shape[0], aphi.shape[0]
    nqm = int(lptr[-1])
    J = np.zeros((2 * nqm, 2 * raster.slen))
    lib().om_jacobian_grid_b(raster.mesh.h, raster.gdim, raster.bdim, raster.bbmin, raster.bbmax, raster.elref, nq,
                             nm, dphi.view(np.float64), aphi.view(np.float64), _c(lptr, np.int32),
                             _c(lidx, np.int32), _cplx(proj), 1 if logscale else 0, J)
    return J


def jacobian_grid_cw(raster, dphi, aphi, lptr, lidx, proj=None, logscale=True):
    """GenerateJacobian_cw_grid (jacobian.cc:610-721) + log row scaling:
    real fields (nq x n, nm x n) -> (Jmua, Jkap), each nqm x slen."""
    dphi = np.ascontiguousarray(dphi, np.float64); aphi = np.ascontiguousarray(aphi, np.float64)
    nq, nm = dphi.shape[0], aphi.shape[0]
    nqm = int(lptr[-1])
    Jm = np.zeros((nqm, raster.slen)); Jk = np.zeros((nqm, raster.slen))
    pj = None if proj is None else np.ascontiguousarray(proj, np.float64)
    lib().om_jacobian_grid_cw_b(raster.mesh.h, raster.gdim, raster.bdim, raster.bbmin, raster.bbmax, raster.elref, nq,
                                nm, dphi, aphi, _c(lptr, np.int32), _c(lidx, np.int32),
                              None if pj is None else pj.ctypes.data_as(C.c_void_p), 1 if logscale else 0,
                              Jm.ctypes.data_as(C.c_void_p), Jk.ctypes.data_as(C.c_void_p))
    return Jm, Jk


def adjoint_sources(cproj, proj_real, data, sd, mvec, lptr, lidx):
    """The adjoint source of every source q in AddDataGradient
    (toastmodule.cc:1772-1808): cproj the nQM linear complex projections,
    proj_real the log projections [ln amp ; phase] (2 nQM), data / sd the
    measured data and its standard deviations (2 nQM), mvec (nM x n) complex
    dense.  Returns (nq x n) complex."""
    nqm = int(lptr[-1])
    nq = len(lptr) - 1
    out = np.zeros((nq, mvec.shape[1]), np.complex128)
    for q in range(nq):
        for k in range(lptr[q], lptr[q + 1]):
            b_mod = (data[k] - proj_real[k]) / sd[k]
            b_arg = (data[nqm + k] - proj_real[nqm + k]) / sd[nqm + k]
            rp, ip = cproj[k].real, cproj[k].imag
            dn = 1.0 / (rp * rp + ip * ip)
            qs = mvec[lidx[k]]
            term = -2.0 * b_mod / (1.0 * sd[k])
            out[q] += qs * complex(term * rp * dn, -term * ip * dn)
            term = -2.0 * b_arg / (1.0 * sd[nqm + k])
            out[q] += qs * complex(-term * ip * dn, -term * rp * dn)
    return out


def data_gradient(raster, dphi, aphi, grad=None):
    """AddDataGradient's raster part: dphi, aphi (nq x n) complex direct and
    adjoint fields; returns grad (2 slen) decremented as the reference does."""
    dphi = np.ascontiguousarray(dphi, np.complex128); aphi = np.ascontiguousarray(aphi, np.complex128)
    g = np.zeros(2 * raster.slen) if grad is None else np.array(grad, np.float64)
    lib().om_data_gradient_b(raster.mesh.h, raster.gdim, raster.bdim, raster.bbmin, raster.bbmax, raster.elref,
                             dphi.shape[0], dphi.view(np.float64), aphi.view(np.float64), g)
    return g


def forward(mesh, qm, mua, mus, ref, freq, qwidth=2.0, mwidth=2.0, solver="bicgstab", tol=1e-15,
            maxit=10000, diag=False, logscale=True):
    """fwdfem restated (src/supertoast/fwdfem.cc:56-260): returns (proj, phi, sysmat)."""
    Q, M, lp, li = qm
    cm, ck, c2 = coeffs(mua, mus, ref)
    omega = freq * 2.0 * np.pi * 1e-6
    rp, ci, val = mesh.sysmat(cm, ck, c2, omega)
    qv = mesh.qvec_gaussian(Q, qwidth)
    mv = mesh.qvec_gaussian(M, mwidth) * c2[None, :]
    nq = Q.shape[0]
    phi = np.zeros((nq, mesh.n), np.complex128)
    for i in range(nq):
        if solver == "cg":
            x, _, _ = pcg(rp, ci, val.real.copy(), qv[i], tol, maxit, diag)
            phi[i] = x
        elif solver == "gmres":
            phi[i] = gmres(rp, ci, val if freq else val.real.copy(), qv[i], tol, maxit, diag)[0]
        else:
            x, _, _ = bicgstab(rp, ci, val, qv[i].astype(np.complex128), tol, maxit, diag)
            phi[i] = x
    proj = project(phi, mv.astype(np.complex128), lp, li, logscale)
    return proj, phi, (rp, ci, val), qv, mv


def read_fem(path):
    """fwdfem output vector ('[v0 v1 ...]', fwdfem.cc:457-462)."""
    with open(path) as f:
        txt = f.read().replace("[", " ").replace("]", " ")
    return np.array(txt.split(), dtype=np.float64)


def cmpvec(a, b):
    """Largest per-element relative difference |a-b|/max(|a|,|b|) (bintools/cmpvec.cc)."""
    a = np.asarray(a, np.float64); b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(a), np.abs(b))
    den[den == 0] = 1.0
    return float(np.max(np.abs(a - b) / den))
