"""COCG (conjugate orthogonal CG: PCG with the unconjugated bilinear form,
for the complex symmetric K = A + i w B) on the GPU, through the C-ABI.
COCG has no counterpart in the reference; its fields are held to the
reference's own answers: the golden projections of the reference's complex
test targets (test/2D/fwdfem, test/3D/fwdfem) and the FP64 oracle BiCGSTAB
fields of the 3D golden case, in every precision mode and with every
64-column SpMM variant.  Also: one SpMM per iteration needs fewer SpMMs
than BiCGSTAB's two, graph replay equals direct launches, and COCG is
refused for real systems as CG is for complex ones."""
import numpy as np
import pytest
from scipy import sparse

from oracle import oracle as O
from tests.cases import CASES, compare

pytestmark = pytest.mark.gpu

TMI = pytest.importorskip("toastpp_amd")
from toastpp_amd import _lib as L  # noqa: E402

VARIANTS = ["lds", "direct", "dma", "rowblock", "rowblock_pipelined"]


def _solver(case, gm, mode, tol, precon="DIAG", variant=None, solver="COCG", graphs=True):
    n = gm.NodeCount()
    fws = TMI.FwdSolver(gm, solver, tol, complex=True)
    fws.SetMode(mode)
    if precon:
        fws.SetPrecon(precon)
    if variant:
        fws.SetSpmmVariant(variant)
    fws.SetGraphs(graphs)
    fws.Allocate()
    fws.Reset(TMI.Solution(np.full(n, case["mua"]), np.full(n, case["mus"]), np.full(n, case["n"])),
              case["freq"] * 2 * np.pi * 1e-6)
    return fws


@pytest.fixture(scope="module")
def golden_3d():
    case = CASES["3d_cplx_direct"]
    om = O.Mesh(case["mesh"])
    n = om.n
    cm, ck, c2 = O.coeffs(np.full(n, case["mua"]), np.full(n, case["mus"]), np.full(n, case["n"]))
    rp, ci, val = om.sysmat(cm, ck, c2, case["freq"] * 2 * np.pi * 1e-6)
    Q = om.read_qm(case["qm"])[0]
    qv = om.qvec_gaussian(Q[:4], 2.0)
    ref = [O.bicgstab(rp, ci, val, qv[i].astype(np.complex128), 1e-15, 20000, True)[0] for i in range(4)]
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    return case, gm, ref


# the complex targets whose golden files hold converged fields: the GMRES
# targets (test/2D/fwdfem/Makefile:128-148) hold the point where the
# reference's GMRES(20) stopped, whose phase lies ~2e-7 from the converged
# field -- a closer solve differs from them by more than their 1e-7 tolerance
COCG_GOLDEN = sorted(k for k in CASES if CASES[k]["freq"] != 0.0 and CASES[k]["solver"] != "gmres")


@pytest.mark.parametrize("name", COCG_GOLDEN)
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_cocg_golden_vectors(name, mode):
    """The reference's golden projections of the complex BiCGSTAB and direct
    targets, solved by COCG at tolerance 1e-15 (the reference suite's
    setting)."""
    case = CASES[name]
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    fws = _solver(case, gm, mode, 1e-15, precon="DIAG" if case["precon"] == "diag" else None)
    print("COCG golden %s mode=%s" % (name, mode), flush=True)
    phi = fws.CalcFields(sparse.csr_matrix(gm.Qvec(width=2.0).T), allow_notconv=True)
    worst = max(s[2] for s in fws.stats)
    assert worst <= 1e-12, "FP64 relative residual %.3g" % worst
    proj = fws.ProjectAll(sparse.csr_matrix(gm.Mvec(width=2.0, ref=case["n"]).T), phi)
    for fname, diff, tol in compare(case, proj):
        assert diff <= tol, "%s: %s rel diff %.3g > %g" % (name, fname, diff, tol)


@pytest.mark.parametrize("mode", ["fp64", "mixed", "fp64_block"])
@pytest.mark.parametrize("precon", [None, "DIAG"])
def test_cocg_fields_vs_oracle_3d(golden_3d, mode, precon):
    case, gm, ref = golden_3d
    fws = _solver(case, gm, mode, 1e-12, precon=precon)
    print("COCG 3D mode=%s precon=%s" % (mode, precon), flush=True)
    phi = fws.CalcFields(sparse.csr_matrix(gm.Qvec(width=2.0).T))
    assert all(s[3] for s in fws.stats) and max(s[2] for s in fws.stats) <= 1e-12
    for i in range(4):
        rel = np.linalg.norm(phi[i] - ref[i]) / np.linalg.norm(ref[i])
        assert rel <= 1e-8, "column %d rel L2 %.3g" % (i, rel)


@pytest.mark.parametrize("variant", VARIANTS)
def test_cocg_spmm_variants_vs_oracle_3d(golden_3d, variant):
    case, gm, ref = golden_3d
    fws = _solver(case, gm, "mixed", 1e-12, variant=variant)
    print("COCG variant %s" % variant, flush=True)
    phi = fws.CalcFields(sparse.csr_matrix(gm.Qvec(width=2.0).T))
    assert all(s[3] for s in fws.stats)
    for i in range(4):
        rel = np.linalg.norm(phi[i] - ref[i]) / np.linalg.norm(ref[i])
        assert rel <= 1e-8, "%s column %d rel L2 %.3g" % (variant, i, rel)


def test_cocg_fewer_spmms_than_bicgstab(golden_3d):
    """One SpMM per COCG iteration against two per BiCGSTAB iteration: on the
    3D golden case at tolerance 1e-10 COCG needs fewer SpMMs per column."""
    case, gm, ref = golden_3d
    qv = sparse.csr_matrix(gm.Qvec(width=2.0).T)
    its = {}
    for solver in ("COCG", "BICGSTAB"):
        fws = _solver(case, gm, "fp64_block", 1e-10, solver=solver)
        fws.CalcFields(qv)
        assert all(s[3] for s in fws.stats)
        its[solver] = np.array([s[0] for s in fws.stats], dtype=float)
    spmm_cocg, spmm_bicg = its["COCG"], 2 * its["BICGSTAB"]
    print("SpMMs per column: COCG %.1f, BiCGSTAB %.1f" % (spmm_cocg.mean(), spmm_bicg.mean()))
    assert spmm_cocg.mean() < spmm_bicg.mean()


@pytest.mark.parametrize("mode", ["mixed", "fp64_block"])
def test_cocg_graph_replay_identical(golden_3d, mode):
    case, gm, ref = golden_3d
    B = np.ascontiguousarray(gm.Qvec(width=2.0).toarray().astype(np.complex128))[:, :48]
    dB = L.DeviceArray.from_host(B)
    out = {}
    for graphs in (True, False):
        fws = _solver(case, gm, mode, 1e-10, graphs=graphs)
        dX = L.DeviceArray(B.shape, B.dtype)
        st1 = fws.CalcFieldsDevice(B.shape[1], dB.ptr, dX.ptr)
        X1 = dX.to_host()
        st2 = fws.CalcFieldsDevice(B.shape[1], dB.ptr, dX.ptr)
        assert np.array_equal(dX.to_host(), X1) and st1 == st2
        out[graphs] = (X1, st1, fws.LastCounts())
    assert np.array_equal(out[True][0], out[False][0]) and out[True][1] == out[False][1]
    assert out[True][2][1] > 0 and out[False][2][1] == 0


def test_cocg_and_cg_field_restrictions(golden_3d):
    """COCG is for complex systems and CG for real ones: the other pairing is
    refused at creation and by SetLinSolver, with TMI_ERR_UNSUPPORTED."""
    case, gm, ref = golden_3d
    with pytest.raises(L.TmiError, match="COCG"):
        TMI.FwdSolver(gm, "COCG", 1e-10, complex=False)
    with pytest.raises(L.TmiError, match="CG on a complex"):
        TMI.FwdSolver(gm, "CG", 1e-10, complex=True)
    f = TMI.FwdSolver(gm, "CG", 1e-10, complex=False)
    with pytest.raises(L.TmiError, match="COCG"):
        f.SetLinSolver("COCG", 1e-10)
    g = TMI.FwdSolver(gm, "BICGSTAB", 1e-10, complex=True)
    g.SetLinSolver("COCG", 1e-10)   # accepted for a complex system
