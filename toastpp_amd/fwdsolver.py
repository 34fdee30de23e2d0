"""FwdSolver: Python mirror of libstoast's TFwdSolver<T> (fwdsolver.h:77-508).

Method names follow the reference (Allocate, AssembleSystemMatrix, Reset,
CalcFields, ProjectAll, SetLinSolver, SetPrecon, ...); all work runs in
libtoastpp_amd on the GPU.
"""
import ctypes as C

import numpy as np

from . import _lib as L

# SpMM variant names, in the order of their C-ABI codes
SPMM_VARIANTS = ("lds", "direct", "dma", "rowblock", "rowblock_pipelined", "rowblock_mfma")


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Solution:
    """FEM coefficients of the diffusion operator (Solution, solution.h:12-16):
    cmua = c mua / n, ckappa = c / (3 n (mua + mus')), c2a = c / (2 n A(n))."""

    def __init__(self, mua, mus, ref):
        mua, mus, ref = _f64(mua), _f64(mus), _f64(ref)
        n = len(mua)
        if not (len(mus) == n and len(ref) == n):
            raise ValueError("mua, mus and ref must have the same length")
        self.cmua, self.ckappa, self.c2a = np.zeros(n), np.zeros(n), np.zeros(n)
        L.call("tmi_optical_to_coeffs", n, mua, mus, ref, self.cmua, self.ckappa, self.c2a)


class FwdSolver:
    """TFwdSolver<double> (complex=False) or TFwdSolver<std::complex<double>>."""

    def __init__(self, mesh, solver="BICGSTAB", tol=1e-10, complex=True):
        self.mesh = mesh
        self.complex = bool(complex)
        h = C.c_void_p()
        name = solver.upper()
        # COCG (complex systems only) has no reference counterpart: the
        # conjugate orthogonal CG for the complex symmetric diffusion matrix
        ls = {"DIRECT": L.LSOLVER_DIRECT, "CG": L.LSOLVER_CG, "BICGSTAB": L.LSOLVER_BICGSTAB,
              "GMRES": L.LSOLVER_GMRES, "COCG": L.LSOLVER_COCG}.get(name)
        if ls is None:
            raise ValueError("unsupported linear solver %r (DIRECT, CG, BICGSTAB, GMRES, COCG)" % solver)
        L.call("tmi_solver_create", mesh.handle_ptr(), 1 if self.complex else 0, ls, float(tol), C.byref(h))
        self._h = h
        self.stats = None
        self.unwrap_phase = False   # fwdsolver.cc:74

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and h.value and L._lib is not None:
            L._lib.tmi_solver_free(h)
            self._h = None

    # --- configuration (SetLinSolver / SetPrecon / SetLinSolverMaxit) ---
    def SetLinSolver(self, solver, tol=1e-10):
        L.call("tmi_solver_set_linsolver", self._h, solver.encode(), float(tol))

    def SetPrecon(self, precon):
        p = {"NONE": L.PRECON_NONE, "DIAG": L.PRECON_DIAG}.get(str(precon).upper(), precon)
        L.call("tmi_solver_set_precon", self._h, int(p))

    def SetLinSolverMaxit(self, maxit):
        L.call("tmi_solver_set_maxit", self._h, int(maxit))

    def SetMode(self, mode):
        """'fp64' (single right-hand side), 'fp64_block' or 'mixed' (FP32 block + FP64 refinement)."""
        m = {"fp64": L.MODE_FP64_SINGLE, "mixed": L.MODE_MIXED_BLOCK, "fp64_block": L.MODE_FP64_BLOCK}.get(mode, mode)
        L.call("tmi_solver_set_mode", self._h, int(m))

    def SetTile(self, tile):
        L.call("tmi_solver_set_tile", self._h, int(tile))

    def SetOrder(self, order):
        """'cluster' (default), 'morton' or 'natural' node order of the
        solver's copy of the system; call Allocate afterwards."""
        L.call("tmi_solver_set_order", self._h, {"natural": L.ORDER_NATURAL, "morton": L.ORDER_MORTON,
                                                 "cluster": L.ORDER_CLUSTER}[order])

    def SetRowsPerWave(self, rows):
        L.call("tmi_solver_set_rows_per_wave", self._h, int(rows))

    def SetSpmmVariant(self, variant):
        """'direct': per-nonzero gathers; 'lds': LDS-staged 64-column
        SpMM; 'dma': LDS-staged with the LDS DMA, double-buffered; 'rowblock':
        blocks of 8 rows gathering each shared neighbour row once;
        'rowblock_pipelined': the same with double-buffered gathers;
        'rowblock_mfma' (default): each block of 8 rows a dense product on
        the matrix cores over the same shared rows (also the 32- and
        16-column tiles)."""
        L.call("tmi_solver_set_spmm_variant", self._h, SPMM_VARIANTS.index(variant))

    # --- lifecycle ---
    def Allocate(self):
        L.call("tmi_solver_allocate", self._h)

    def _coeffs(self, sol, elbasis):
        want = self.mesh.ElementCount() if elbasis else self.mesh.NodeCount()
        c = [_f64(sol.cmua), _f64(sol.ckappa), _f64(sol.c2a)]
        if any(len(a) != want for a in c):
            raise ValueError("the solution holds %d values per parameter, the %s basis needs %d"
                             % (len(c[0]), "element" if elbasis else "nodal", want))
        return c

    def AssembleSystemMatrix(self, sol, omega=0.0, elbasis=False):
        """AssembleSystemMatrix(sol, omega, elbasis) (fwdsolver.cc:338-424):
        nodal coefficients, or one per element with elbasis=True."""
        L.call("tmi_solver_assemble_basis", self._h, *self._coeffs(sol, elbasis), float(omega), 1 if elbasis else 0)

    def Reset(self, sol, omega=0.0, elbasis=False):
        L.call("tmi_solver_reset_basis", self._h, *self._coeffs(sol, elbasis), float(omega), 1 if elbasis else 0)

    def nnz(self):
        v = C.c_int()
        L.call("tmi_solver_nnz", self._h, C.byref(v))
        return v.value

    def SystemMatrix(self):
        """(rowptr, colidx, values) of the FP64 system matrix F."""
        nz = self.nnz()
        n = self.mesh.NodeCount()
        rp = np.zeros(n + 1, np.int32)
        ci = np.zeros(nz, np.int32)
        val = np.zeros(2 * nz if self.complex else nz)
        L.call("tmi_solver_get_sysmat", self._h, L.ptr(rp), L.ptr(ci), L.ptr(val))
        return rp, ci, (val.view(np.complex128) if self.complex else val)

    def CalcFields(self, qvec, phi=None, allow_notconv=False):
        """qvec: (nQ x n) scipy CSR matrix or dense array (row q = source q).
        Returns phi (nQ x n); per-column records in self.stats."""
        from scipy import sparse

        q = sparse.csr_matrix(qvec)
        nq, n = q.shape
        if n != self.mesh.NodeCount():
            raise ValueError("qvec has %d columns, the mesh %d nodes" % (n, self.mesh.NodeCount()))
        dt = np.complex128 if self.complex else np.float64
        vals = np.ascontiguousarray(q.data, dtype=dt)
        if phi is None:
            phi = np.zeros((nq, n), dt)
        phi = np.ascontiguousarray(phi, dtype=dt)
        if phi.shape != (nq, n):
            raise ValueError("phi has shape %s, expected (%d, %d) (sources x nodes)" % (phi.shape, nq, n))
        stats = (L.ColStat * nq)()
        st = L.call("tmi_solver_calc_fields", self._h, nq, np.ascontiguousarray(q.indptr, np.int32),
                    np.ascontiguousarray(q.indices, np.int32), vals.view(np.float64), phi.view(np.float64),
                    C.cast(stats, C.c_void_p), allow_notconv=allow_notconv)
        self.stats = [(s.iters, s.sweeps, s.relres, bool(s.converged)) for s in stats]
        self.converged = st == L.TMI_OK
        return phi

    def CalcFieldsDevice(self, nq, d_rhs, d_phi):
        """GPU-resident CalcFields: d_rhs, d_phi are device pointers (int) to
        [n][nq] node-major arrays; returns the per-column records."""
        stats = (L.ColStat * nq)()
        st = L.call("tmi_solver_calc_fields_device", self._h, nq, C.c_void_p(d_rhs), C.c_void_p(d_phi),
                    C.cast(stats, C.c_void_p), allow_notconv=True)
        self.stats = [(s.iters, s.sweeps, s.relres, bool(s.converged)) for s in stats]
        self.converged = st == L.TMI_OK
        return self.stats

    def CalcFieldsFreqsDevice(self, nq, d_rhs, d_phi, omegas):
        """Several frequencies of this mesh in one block solve: column q of
        the [n][nq] device arrays is solved at omegas[q] (the unit of Reset's
        omega) with the optical coefficients of the last Reset; COCG, block
        modes.  No reference counterpart (the reference solves each frequency
        after its own AssembleSystemMatrix).  Returns the per-column records."""
        om = np.ascontiguousarray(omegas, dtype=np.float64)
        if om.shape != (nq,):
            raise ValueError("omegas has shape %s, expected (%d,)" % (om.shape, nq))
        stats = (L.ColStat * nq)()
        st = L.call("tmi_solver_calc_fields_freqs_device", self._h, nq, C.c_void_p(d_rhs), C.c_void_p(d_phi),
                    om.ctypes.data_as(C.c_void_p), C.cast(stats, C.c_void_p), allow_notconv=True)
        self.stats = [(s.iters, s.sweeps, s.relres, bool(s.converged)) for s in stats]
        self.converged = st == L.TMI_OK
        return self.stats

    def CalcFieldsFreqs(self, qvec, omegas):
        """CalcFields with column q solved at omegas[q] (see
        CalcFieldsFreqsDevice); qvec (nQ x n), returns phi (nQ x n)."""
        from scipy import sparse

        q = sparse.csr_matrix(qvec)
        nq, n = q.shape
        if n != self.mesh.NodeCount():
            raise ValueError("qvec has %d columns, the mesh %d nodes" % (n, self.mesh.NodeCount()))
        B = np.ascontiguousarray(q.T.toarray().astype(np.complex128))   # [n][nq]
        dB = L.DeviceArray.from_host(B)
        dX = L.DeviceArray(B.shape, np.complex128)
        self.CalcFieldsFreqsDevice(nq, dB.ptr, dX.ptr, omegas)
        return np.ascontiguousarray(dX.to_host().T)

    def Autotune(self, nq, d_rhs, d_phi, ld=None):
        """Pick tile width, row schedule and SpMM variant for nq device
        right-hand sides ([n][ld] node-major); returns them as a dict."""
        t, r, v, ms = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        L.call("tmi_solver_autotune", self._h, int(nq), C.c_void_p(d_rhs), C.c_void_p(d_phi), int(ld or nq),
               C.byref(t), C.byref(r), C.byref(v), C.byref(ms))
        return {"tile": t.value, "rows_per_wave": r.value, "spmm_variant": SPMM_VARIANTS[v.value],
                "solve_ms": ms.value}

    def SetCompaction(self, min_width):
        """narrowest tile of the column compaction (0: off, default 2)"""
        L.call("tmi_solver_set_compaction", self._h, int(min_width))

    def SetFallback(self, enable=True, mixed_maxit=0):
        """mixed mode: re-solve the columns the FP32 iteration leaves above the
        tolerance with the FP64 block iteration (default on); mixed_maxit caps
        the FP32 iteration (0: the solver's maxit)"""
        L.call("tmi_solver_set_fallback", self._h, 1 if enable else 0, int(mixed_maxit))

    def WorkBytes(self, nq):
        """device bytes of the workspace a solve of nq columns uses"""
        b = C.c_double()
        L.call("tmi_solver_work_bytes", self._h, int(nq), C.byref(b))
        return b.value

    def SetGraphs(self, enable=True):
        """replay the block iteration's check periods from captured HIP graphs
        (default) or launch every kernel"""
        L.call("tmi_solver_set_graphs", self._h, 1 if enable else 0)

    def LastCounts(self):
        """(kernel launch calls, graph launches, host synchronisations) of the
        last CalcFields"""
        k, g, s = C.c_longlong(), C.c_longlong(), C.c_longlong()
        L.call("tmi_solver_last_counts", self._h, C.byref(k), C.byref(g), C.byref(s))
        return k.value, g.value, s.value

    def LastTiming(self):
        ms, it = C.c_double(), C.c_double()
        L.call("tmi_solver_last_timing", self._h, C.byref(ms), C.byref(it))
        return ms.value, it.value

    def TimeKernel(self, which=0, reps=20):
        """(avg_ms, algorithmic_bytes) of the inner SpMM (0) or FP64 residual SpMM (1)."""
        ms, b = C.c_double(), C.c_double()
        L.call("tmi_solver_time_kernel", self._h, int(which), int(reps), C.byref(ms), C.byref(b))
        return ms.value, b.value

    def ProjectAllDevice(self, mvec, d_phi, ld, scale="log"):
        """ProjectAll on GPU-resident fields: d_phi is a device pointer (int)
        to [n][ld] node-major fields (the first nQ entries of each row)."""
        from scipy import sparse

        m = sparse.csr_matrix(mvec)
        dt = np.complex128 if self.complex else np.float64
        nq, nm, nqm = self.mesh.qm_dims()
        if m.shape != (nm, self.mesh.NodeCount()):
            raise ValueError("mvec has shape %s, expected (%d, %d) (detectors x nodes)"
                             % (m.shape, nm, self.mesh.NodeCount()))
        if int(ld) < nq:
            raise ValueError("field row stride %d below the source count %d" % (ld, nq))
        proj = np.zeros(nqm, dt)
        vals = np.ascontiguousarray(m.data, dtype=dt)
        L.call("tmi_solver_project_all_device", self._h, m.shape[0], np.ascontiguousarray(m.indptr, np.int32),
               np.ascontiguousarray(m.indices, np.int32), vals.view(np.float64), C.c_void_p(d_phi), ld,
               L.DATA_LOG if scale == "log" else L.DATA_LIN, proj.view(np.float64))
        return proj

    def SetPhaseUnwrap(self, unwrap):
        """fwdsolver.h:142-143."""
        self.unwrap_phase = bool(unwrap)

    def ProjectAllReal(self, mvec, phi, scale="log"):
        """ProjectAll_real (fwdsolver.cc:940-992): [Re ..., Im ...] of the
        projections (log: ln amplitude, phase).  With phase unwrapping on and
        log scaling, positive phases are shifted by -2 pi (the branch the
        reference compiles, fwdsolver.cc:985-989)."""
        p = self.ProjectAll(mvec, phi, scale)
        if not self.complex:
            return np.asarray(p, np.float64)
        out = np.concatenate([p.real, p.imag])
        if self.unwrap_phase and scale != "lin":
            ph = out[len(p):]
            ph[ph > 0.0] -= 2.0 * np.pi
        return out

    def ProjectAll(self, mvec, phi, scale="log"):
        """mvec: (nM x n) CSR/dense; phi: (nQ x n) -> nQM projections in link order."""
        from scipy import sparse

        m = sparse.csr_matrix(mvec)
        dt = np.complex128 if self.complex else np.float64
        nq, nm, nqm = self.mesh.qm_dims()
        n = self.mesh.NodeCount()
        if m.shape != (nm, n):
            raise ValueError("mvec has shape %s, expected (%d, %d) (detectors x nodes)" % (m.shape, nm, n))
        proj = np.zeros(nqm, dt)
        vals = np.ascontiguousarray(m.data, dtype=dt)
        phi = np.ascontiguousarray(phi, dtype=dt)
        if phi.shape != (nq, n):
            raise ValueError("phi has shape %s, expected (%d, %d) (sources x nodes)" % (phi.shape, nq, n))
        L.call("tmi_solver_project_all", self._h, m.shape[0], np.ascontiguousarray(m.indptr, np.int32),
               np.ascontiguousarray(m.indices, np.int32), vals.view(np.float64), phi.view(np.float64),
               L.DATA_LOG if scale == "log" else L.DATA_LIN, proj.view(np.float64))
        return proj
