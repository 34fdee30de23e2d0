"""GPU parity of the forward path against the FP64 oracle and the reference's
golden vectors.  Everything goes through the C-ABI (toastpp_amd._lib)."""
import numpy as np
import pytest
from scipy import sparse

from oracle import oracle as O
from tests.cases import CASES, compare

pytestmark = pytest.mark.gpu

TMI = pytest.importorskip("toastpp_amd")

MESHES = {"2d": CASES["2d_real_cg"], "3d": CASES["3d_cplx_direct"]}


@pytest.fixture(scope="module", params=["2d", "3d"])
def pair(request):
    case = MESHES[request.param]
    om = O.Mesh(case["mesh"])
    gm = TMI.Mesh(case["mesh"])
    return case, om, gm


def test_setup_boundary_and_sizes(pair):
    case, om, gm = pair
    assert gm.NodeCount() == om.n and gm.ElementCount() == om.elen
    assert np.array_equal(gm.BoundaryFlags(), om.bnd())
    # signed element sizes: same operations as Tetrahedron4/Triangle3::ComputeSize
    assert np.array_equal(gm.ElementSize(), om.size())


def test_csr_pattern_bit_exact(pair):
    case, om, gm = pair
    rp, ci = gm.SparseStructure()
    orp, oci = om.csr()
    assert np.array_equal(rp, orp)
    assert np.array_equal(ci, oci)


@pytest.mark.parametrize("freq", [0.0, 100.0])
def test_system_matrix(pair, freq):
    case, om, gm = pair
    n = om.n
    rng = np.random.default_rng(7)
    mua = 0.01 + 0.02 * rng.random(n)
    mus = 1.0 + rng.random(n)
    ref = 1.3 + 0.2 * rng.random(n)
    cm, ck, c2 = O.coeffs(mua, mus, ref)
    omega = freq * 2 * np.pi * 1e-6
    _, _, oval = om.sysmat(cm, ck, c2, omega)
    fws = TMI.FwdSolver(gm, "BICGSTAB", 1e-10, complex=freq != 0.0)
    fws.Allocate()
    sol = TMI.Solution(mua, mus, ref)
    assert np.array_equal(sol.cmua, cm) and np.array_equal(sol.ckappa, ck) and np.array_equal(sol.c2a, c2)
    fws.AssembleSystemMatrix(sol, omega)
    rp, ci, val = fws.SystemMatrix()
    ref_val = oval if freq else oval.real
    # FP64 values: same element integrals, same per-entry summation order
    # (elements ascending, PFF then PDD then BNDPFF then i*omega*FF).
    mism = np.count_nonzero(val != ref_val)
    rel = np.max(np.abs(val - ref_val)) / np.max(np.abs(ref_val))
    assert rel <= 1e-15, "max rel diff %g (%d entries differ)" % (rel, mism)
    if om.dim == 3:
        # Tet4 path: only +-*/ and sqrt, all correctly rounded on the GPU -> bit-exact
        assert mism == 0, "%d entries differ (max rel %g)" % (mism, rel)
    print("sysmat entries differing from oracle: %d of %d (max rel %.3g)" % (mism, val.size, rel))


def test_qm_and_profiles(pair):
    case, om, gm = pair
    gm.ReadQM(case["qm"])
    Q, M, lp, li = om.read_qm(case["qm"])
    gQ, gM, glp, gli = gm.QMPositions()
    assert np.array_equal(gQ, Q) and np.array_equal(glp, lp) and np.array_equal(gli, li)
    assert np.max(np.abs(gM - M)) <= 1e-12
    oq = om.qvec_gaussian(Q, 2.0)
    gq = gm.Qvec(width=2.0).toarray().T.real
    assert np.max(np.abs(gq - oq)) <= 1e-14 * np.max(np.abs(oq))
    n = om.n
    c2 = O.coeffs(np.full(n, 0.01), np.ones(n), np.full(n, 1.4))[2]
    omv = om.qvec_gaussian(M, 2.0) * c2[None, :]
    gmv = gm.Mvec(width=2.0, ref=1.4).toarray().T.real
    assert np.max(np.abs(gmv - omv)) <= 1e-14 * np.max(np.abs(omv))


def _run_case(name, mode, tol=1e-15):
    case = CASES[name]
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    n = gm.NodeCount()
    mua, mus, ref = np.full(n, case["mua"]), np.full(n, case["mus"]), np.full(n, case["n"])
    cplx = case["freq"] != 0.0
    ls = "GMRES" if case["solver"] == "gmres" else "BICGSTAB" if cplx else "CG"
    fws = TMI.FwdSolver(gm, ls, tol, complex=cplx)
    fws.SetMode(mode)
    if case["precon"] == "diag":
        fws.SetPrecon("DIAG")
    fws.Allocate()
    fws.Reset(TMI.Solution(mua, mus, ref), case["freq"] * 2 * np.pi * 1e-6)
    qv = sparse.csr_matrix(gm.Qvec(width=2.0).T)
    mv = sparse.csr_matrix(gm.Mvec(width=2.0, ref=case["n"]).T)
    if not cplx:
        qv = qv.real
        mv = mv.real
    phi = fws.CalcFields(qv, allow_notconv=True)
    proj = fws.ProjectAll(mv, phi)
    if not cplx:
        proj = proj.astype(np.complex128)
    return case, fws, phi, proj


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_golden_vectors(name, mode):
    # the reference suite solves to tol 1e-15 / maxit 10000 and accepts what
    # the solver reaches; the refinement loop here stops at the FP64 floor
    case, fws, phi, proj = _run_case(name, mode)
    worst = max(s[2] for s in fws.stats)
    assert worst <= 1e-13, "FP64 relative residual %.3g" % worst
    for fname, diff, tol in compare(case, proj):
        assert diff <= tol, "%s: %s rel diff %.3g > %g" % (name, fname, diff, tol)


@pytest.mark.parametrize("mode", ["fp64", "mixed", "fp64_block"])
def test_fields_vs_oracle_3d(mode):
    """Fields of the 3D golden case against the FP64 oracle BiCGSTAB (1e-15)."""
    case, fws, phi, proj = _run_case("3d_cplx_direct", mode, tol=1e-12)
    assert all(s[3] for s in fws.stats)
    om = O.Mesh(case["mesh"])
    n = om.n
    cm, ck, c2 = O.coeffs(np.full(n, case["mua"]), np.full(n, case["mus"]), np.full(n, case["n"]))
    rp, ci, val = om.sysmat(cm, ck, c2, case["freq"] * 2 * np.pi * 1e-6)
    Q = om.read_qm(case["qm"])[0]
    qv = om.qvec_gaussian(Q[:4], 2.0)
    for i in range(4):
        x, it, err = O.bicgstab(rp, ci, val, qv[i].astype(np.complex128), 1e-15, 20000, True)
        rel = np.linalg.norm(phi[i] - x) / np.linalg.norm(x)
        assert rel <= 1e-8, "column %d rel L2 %.3g" % (i, rel)


def test_point_sources_match_oracle(pair):
    case, om, gm = pair
    gm.ReadQM(case["qm"])
    Q = om.read_qm(case["qm"])[0]
    for neumann in (True, False):
        gq = gm.Qvec(type="Neumann" if neumann else "Isotropic", shape="Point").toarray().T.real
        oq = O.qvec_point(om, Q, neumann=neumann)
        assert np.array_equal(gq, oq)
    assert gm.Mvec(shape="Point").nnz == 0   # the reference's CalcMvec has no point profile


def test_project_all_real_phase_unwrap():
    """ProjectAll_real with SetPhaseUnwrap (fwdsolver.cc:951-989): the
    log-amplitude half is unchanged, every phase ends up in (-2 pi, 0] and
    differs from the wrapped phase by exactly 0 or -2 pi."""
    case, fws, phi, proj = _run_case("2d_cplx_bicgstab", "mixed", tol=1e-10)
    gm = fws.mesh
    mv = sparse.csr_matrix(gm.Mvec(width=2.0, ref=case["n"]).T)
    wrapped = fws.ProjectAllReal(mv, phi)
    fws.SetPhaseUnwrap(True)
    unwrapped = fws.ProjectAllReal(mv, phi)
    nqm = len(proj)
    assert np.array_equal(unwrapped[:nqm], wrapped[:nqm])
    shift = unwrapped[nqm:] - wrapped[nqm:]
    assert np.all((shift == 0.0) | (shift == -2.0 * np.pi))
    assert np.all(unwrapped[nqm:] <= 0.0) and np.all(unwrapped[nqm:] > -2.0 * np.pi)
    assert np.any(shift != 0.0) or np.all(wrapped[nqm:] <= 0.0)
    assert np.array_equal(fws.ProjectAllReal(mv, phi, scale="lin"), np.concatenate(
        [fws.ProjectAll(mv, phi, "lin").real, fws.ProjectAll(mv, phi, "lin").imag]))


def test_module_level_functions(pair):
    """toast's module-level Sysmat_CW / Jacobian (mesh.py:344-349) on Mesh
    and Raster objects give the methods' results."""
    case, om, gm = pair
    n = om.n
    rng = np.random.default_rng(3)
    mua, mus, ref = 0.01 + 0.01 * rng.random(n), 1.0 + rng.random(n), np.full(n, 1.4)
    A = TMI.Sysmat_CW(gm, mua, mus, ref, 0)
    B = gm.Sysmat_CW(mua, mus, ref)
    assert (A != B).nnz == 0
    gm.ReadQM(case["qm"])
    nq = gm.QMPositions()[0].shape[0]
    nm = gm.QMPositions()[1].shape[0]
    dphi = rng.standard_normal((n, nq)) + 1j * rng.standard_normal((n, nq))
    aphi = rng.standard_normal((n, nm)) + 1j * rng.standard_normal((n, nm))
    nqm = int(gm.QMPositions()[2][-1])
    proj = rng.standard_normal(nqm) + 1j * rng.standard_normal(nqm) + 3.0
    grid = np.array([16] * om.dim)
    r = TMI.Raster(gm, grid)
    assert np.array_equal(TMI.Jacobian(gm, r, dphi, aphi, proj), gm.Jacobian(r, dphi, aphi, proj))
