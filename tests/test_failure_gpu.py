"""Failure and fallback paths of CalcFields on the GPU, through the C-ABI.

* The mixed mode's FP64 fallback: with the FP32 iteration capped at a few
  iterations every column is handed to the FP64 block iteration, which
  finishes it; the fields agree with the oracle, the records count both
  parts.  With the fallback off the same solve reports every column
  unconverged.
* maxit exhaustion (SetLinSolverMaxit) in the FP64 modes returns
  TMI_ERR_NOT_CONVERGED with per-column records whose residual is the true
  FP64 residual of the returned fields (the reference's
  IterativeSolverResult, matrix.h:34-37).
* Non-finite coefficients are refused on the host before any GPU work.
* The autotuner picks its settings and the solve after it still matches the
  oracle; kernel timing refuses a workspace left in another precision.

Case: the reference's 3D golden mesh cyl2 (9,043 nodes, Tet4), 100 MHz.
"""
import numpy as np
import pytest
from scipy import sparse

from oracle import oracle as O
from tests.cases import CASES

pytestmark = pytest.mark.gpu

T = pytest.importorskip("toastpp_amd")
from toastpp_amd import _lib as L  # noqa: E402

CASE = CASES["3d_cplx_direct"]
OMEGA = 100.0 * 2 * np.pi * 1e-6


@pytest.fixture(scope="module")
def cyl2():
    gm = T.Mesh(CASE["mesh"])
    gm.ReadQM(CASE["qm"])
    n = gm.NodeCount()
    rng = np.random.default_rng(11)
    mua, mus, ref = 0.01 + 0.01 * rng.random(n), 1.0 + rng.random(n), np.full(n, 1.4)
    B = np.ascontiguousarray(gm.Qvec(width=2.0).toarray().astype(np.complex128))   # n x 80
    om = O.Mesh(CASE["mesh"])
    cm, ck, c2 = O.coeffs(mua, mus, ref)
    rp, ci, val = om.sysmat(cm, ck, c2, OMEGA)
    return dict(gm=gm, n=n, mua=mua, mus=mus, ref=ref, B=B, sys=(rp, ci, val))


def _solver(c, mode, tol=1e-10):
    f = T.FwdSolver(c["gm"], "BICGSTAB", tol, complex=True)
    f.SetMode(mode)
    f.SetPrecon("DIAG")
    f.Allocate()
    f.Reset(T.Solution(c["mua"], c["mus"], c["ref"]), OMEGA)
    return f


def _solve(f, B):
    dB = L.DeviceArray.from_host(B)
    dX = L.DeviceArray(B.shape, np.complex128)
    st = f.CalcFieldsDevice(B.shape[1], dB.ptr, dX.ptr)
    return st, dX.to_host()


def _oracle(c, cols, tol=1e-14):
    rp, ci, val = c["sys"]
    return [O.bicgstab(rp, ci, val, np.ascontiguousarray(c["B"][:, j]), tol, 50000, True)[0] for j in cols]


def test_fp64_fallback_finishes_capped_mixed_columns(cyl2):
    B = np.ascontiguousarray(cyl2["B"][:, :16])
    f = _solver(cyl2, "mixed")
    f.SetFallback(True, mixed_maxit=4)
    st, X = _solve(f, B)
    assert f.converged
    for it, sw, rr, conv in st:
        assert conv and rr <= 1e-10
        assert it > 4 and sw >= 2   # 4 FP32 iterations + the FP64 ones; the fallback counts one sweep
    for j, x in zip((0, 15), _oracle(cyl2, (0, 15))):
        assert np.linalg.norm(X[:, j] - x) / np.linalg.norm(x) <= 1e-8


def test_capped_mixed_without_fallback_reports_unconverged(cyl2):
    B = np.ascontiguousarray(cyl2["B"][:, :16])
    f = _solver(cyl2, "mixed")
    f.SetFallback(False, mixed_maxit=4)
    st, X = _solve(f, B)
    assert not f.converged
    rp, ci, val = cyl2["sys"]
    for j, (it, sw, rr, conv) in enumerate(st):
        assert not conv and it == 4 and rr > 1e-10
        # the record holds the FP64 true residual of the returned field
        true = O.relres(rp, ci, val, np.ascontiguousarray(B[:, j]), np.ascontiguousarray(X[:, j]))
        assert abs(rr - true) <= 1e-6 * true


@pytest.mark.parametrize("mode", ["fp64", "fp64_block"])
def test_maxit_exhaustion_reports_true_residual(cyl2, mode):
    B = np.ascontiguousarray(cyl2["B"][:, :8])
    f = _solver(cyl2, mode)
    f.SetLinSolverMaxit(5)
    st, X = _solve(f, B)
    assert not f.converged
    rp, ci, val = cyl2["sys"]
    for j, (it, sw, rr, conv) in enumerate(st):
        assert not conv and it == 5
        true = O.relres(rp, ci, val, np.ascontiguousarray(B[:, j]), np.ascontiguousarray(X[:, j]))
        assert rr > 1e-10 and abs(rr - true) <= 1e-6 * true
    # the host-buffer CalcFields raises NotConverged unless the caller allows it
    q = sparse.csr_matrix(B[:, :2].T)
    with pytest.raises(L.NotConverged):
        f.CalcFields(q)
    phi = f.CalcFields(q, allow_notconv=True)
    assert phi.shape == (2, cyl2["n"]) and not f.converged


def test_real_pcg_maxit_exhaustion(cyl2):
    gm = cyl2["gm"]
    f = T.FwdSolver(gm, "CG", 1e-10, complex=False)
    f.SetPrecon("DIAG")
    f.SetMode("fp64_block")
    f.Allocate()
    f.Reset(T.Solution(cyl2["mua"], cyl2["mus"], cyl2["ref"]), 0.0)
    f.SetLinSolverMaxit(3)
    B = np.ascontiguousarray(cyl2["B"][:, :4].real)
    dB = L.DeviceArray.from_host(B)
    dX = L.DeviceArray(B.shape, np.float64)
    st = f.CalcFieldsDevice(4, dB.ptr, dX.ptr)
    assert not f.converged and all(s[0] == 3 and not s[3] for s in st)


def test_nonfinite_coefficients_refused(cyl2):
    f = T.FwdSolver(cyl2["gm"], "BICGSTAB", 1e-10)
    f.Allocate()
    sol = T.Solution(cyl2["mua"], cyl2["mus"], cyl2["ref"])
    for attr in ("cmua", "ckappa", "c2a"):
        bad = T.Solution(cyl2["mua"], cyl2["mus"], cyl2["ref"])
        getattr(bad, attr)[7] = np.nan
        for call in (f.Reset, f.AssembleSystemMatrix):
            with pytest.raises(L.TmiError) as ei:
                call(bad, OMEGA)
            assert ei.value.code == L.TMI_ERR_INVALID_ARG
    with pytest.raises(L.TmiError):
        f.Reset(sol, float("inf"))
    f.Reset(sol, OMEGA)   # the solver is still usable


def test_autotune_then_fields_match_oracle(cyl2):
    B = np.ascontiguousarray(cyl2["B"][:, :64])
    f = _solver(cyl2, "mixed")
    dB = L.DeviceArray.from_host(B)
    dX = L.DeviceArray(B.shape, np.complex128)
    tuned = f.Autotune(64, dB.ptr, dX.ptr)
    assert tuned["tile"] in (8, 16, 32, 64) and tuned["spmm_variant"] in (
        "lds", "direct", "dma", "rowblock", "rowblock_pipelined", "rowblock_mfma")
    assert tuned["solve_ms"] > 0
    st = f.CalcFieldsDevice(64, dB.ptr, dX.ptr)
    assert all(s[3] for s in st)
    X = dX.to_host()
    for j, x in zip((0, 63), _oracle(cyl2, (0, 63))):
        assert np.linalg.norm(X[:, j] - x) / np.linalg.norm(x) <= 1e-8


def test_time_kernel_refuses_workspace_of_another_precision(cyl2):
    B = np.ascontiguousarray(cyl2["B"][:, :64])
    f = _solver(cyl2, "fp64_block")
    _solve(f, B)
    f.SetMode("mixed")   # the 64-column workspace still holds FP64 vectors
    with pytest.raises(L.TmiError) as ei:
        f.TimeKernel(0, reps=2)
    assert ei.value.code == L.TMI_ERR_STATE
    _solve(f, B)
    ms, nbytes = f.TimeKernel(0, reps=2)
    assert ms > 0 and nbytes > 0
