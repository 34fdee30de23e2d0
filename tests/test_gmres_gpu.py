"""GMRES(20) on the GPU (gmres_imp.hpp:41-230 as IterativeSolve runs it)
against the oracle's restatement of the same loop, through the C-ABI:
the fields, the restart-cycle count of every column, and the fields of a
sparse LU solve.  Complex and real systems, no and diagonal
preconditioning, one tile and several tiles with a ragged last tile.
The oracle's GMRES is pinned by the reference's golden vectors of the
2D GMRES targets (test_oracle.py)."""
import numpy as np
import pytest
from scipy import sparse
from scipy.sparse import linalg as spla

from oracle import oracle as O
from tests.cases import CASES

pytestmark = pytest.mark.gpu

TMI = pytest.importorskip("toastpp_amd")


@pytest.fixture(scope="module")
def disk():
    case = CASES["2d_cplx_gmres"]
    om = O.Mesh(case["mesh"])
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    return case, om, gm


@pytest.mark.parametrize("cplx", [True, False])
@pytest.mark.parametrize("precon", [None, "DIAG"])
@pytest.mark.parametrize("tile", [0, 8])
def test_gmres_vs_oracle(disk, cplx, precon, tile):
    case, om, gm = disk
    n = om.n
    rng = np.random.default_rng(11)
    mua = 0.02 + 0.01 * rng.random(n)
    mus = 1.5 + rng.random(n)
    ref = np.full(n, 1.4)
    freq = 100.0 if cplx else 0.0
    omega = freq * 2 * np.pi * 1e-6
    tol = 1e-12
    fws = TMI.FwdSolver(gm, "GMRES", tol, complex=cplx)
    if precon:
        fws.SetPrecon(precon)
    fws.SetLinSolverMaxit(500)
    if tile:
        fws.SetTile(tile)
    fws.Allocate()
    fws.Reset(TMI.Solution(mua, mus, ref), omega)
    qv = sparse.csr_matrix(gm.Qvec(width=2.0).T)
    if not cplx:
        qv = qv.real
    nq = qv.shape[0] if tile == 0 else 19   # 19 columns on 8-wide tiles: 8 + 8 + 3
    qv = qv[:nq]
    print("GMRES cplx=%s precon=%s tile=%d nq=%d" % (cplx, precon, tile, nq), flush=True)
    phi = fws.CalcFields(qv)
    cm, ck, c2 = O.coeffs(mua, mus, ref)
    rp, ci, val = om.sysmat(cm, ck, c2, omega)
    if not cplx:
        val = np.ascontiguousarray(val.real)
    lu = spla.splu(sparse.csr_matrix((val, ci, rp), shape=(n, n)).tocsc())
    dense_q = qv.toarray()
    worst_o = worst_lu = 0.0
    for q in range(nq):
        xo, cyc, rr = O.gmres(rp, ci, val, dense_q[q], tol, 500, precon == "DIAG")
        it, steps, relres, conv = fws.stats[q]
        assert conv and relres <= tol
        # same loop: the same number of restart cycles (a column whose last
        # residual lands within rounding of the tolerance may differ by one)
        assert abs(it - cyc) <= 1, (q, it, cyc)
        worst_o = max(worst_o, np.linalg.norm(phi[q] - xo) / np.linalg.norm(xo))
        xl = lu.solve(dense_q[q])
        worst_lu = max(worst_lu, np.linalg.norm(phi[q] - xl) / np.linalg.norm(xl))
    print("  max rel L2 vs oracle GMRES %.3g, vs LU %.3g; cycles %s" %
          (worst_o, worst_lu, sorted({s[0] for s in fws.stats})))
    assert worst_o <= 1e-8 and worst_lu <= 1e-8


def test_gmres_cycle_limit_and_zero_rhs(disk):
    """maxit bounds the restart cycles (the last cycle takes no Arnoldi step,
    gmres_imp.hpp:133-137), and a zero right-hand side stops at once."""
    case, om, gm = disk
    n = om.n
    fws = TMI.FwdSolver(gm, "GMRES", 1e-15, complex=True)
    fws.SetLinSolverMaxit(3)
    fws.Allocate()
    fws.Reset(TMI.Solution(np.full(n, 0.025), np.full(n, 2.0), np.full(n, 1.4)), 100.0 * 2 * np.pi * 1e-6)
    qv = sparse.csr_matrix(gm.Qvec(width=2.0).T)[:2].tolil()
    qv[1, :] = 0
    phi = fws.CalcFields(qv.tocsr(), allow_notconv=True)
    it0, steps0, rr0, conv0 = fws.stats[0]
    assert it0 == 3 and steps0 == 40 and not conv0
    cm, ck, c2 = O.coeffs(np.full(n, 0.025), np.full(n, 2.0), np.full(n, 1.4))
    rp, ci, val = om.sysmat(cm, ck, c2, 100.0 * 2 * np.pi * 1e-6)
    xo, cyc, rr = O.gmres(rp, ci, val, qv.toarray()[0].astype(complex), 1e-15, 3)
    assert cyc == 3
    assert np.linalg.norm(phi[0] - xo) / np.linalg.norm(xo) <= 1e-10
    assert abs(rr0 - rr) <= 1e-6 * rr
    assert fws.stats[1][0] == 0 and fws.stats[1][3] and np.all(phi[1] == 0)


def test_gmres_device_entry_autotune_and_work_bytes(disk):
    """GPU-resident CalcFields with GMRES ([n][nq] node-major device
    buffers) gives the host entry point's fields bit for bit; the autotuner
    runs a GMRES solve without kernel timings; the workspace estimate covers
    the restart basis."""
    case, om, gm = disk
    from toastpp_amd import _lib as L
    n = om.n
    fws = TMI.FwdSolver(gm, "GMRES", 1e-12, complex=True)
    fws.SetPrecon("DIAG")
    fws.Allocate()
    fws.Reset(TMI.Solution(np.full(n, 0.025), np.full(n, 2.0), np.full(n, 1.4)), 100.0 * 2 * np.pi * 1e-6)
    qv = sparse.csr_matrix(gm.Qvec(width=2.0).T)[:12]
    phi_h = fws.CalcFields(qv)
    nq = qv.shape[0]
    rhs = L.DeviceArray.from_host(np.ascontiguousarray(qv.toarray().T.astype(np.complex128)))
    phi = L.DeviceArray.from_host(np.zeros((n, nq), np.complex128))
    stats = fws.CalcFieldsDevice(nq, rhs.ptr, phi.ptr)
    assert all(s[3] for s in stats)
    assert np.array_equal(phi.to_host().T, phi_h)
    phi2 = L.DeviceArray.from_host(np.zeros((n, nq), np.complex128))
    tuned = fws.Autotune(nq, rhs.ptr, phi2.ptr)
    print("GMRES autotune:", tuned)
    assert tuned["solve_ms"] > 0
    assert fws.WorkBytes(16) >= 16 * n * 16 * 21
