"""Element basis on the GPU against the FP64 oracle, through the C-ABI.

* AssembleSystemMatrix(sol, omega, elbasis = true) (fwdsolver.cc:338-424,
  ASSEMBLE_*_EL, mesh.cc:1985-1993): one coefficient per element; the 3D
  matrix equals the oracle's bit for bit, the 2D one to the last bit of the
  device hypot in the Tri3 boundary mass.
* Fields of the element-basis system against the oracle's.
* GenerateJacobian(RASTER_ELBASIS) (jacobian.cc:441-497 with IntFG_el /
  IntGradFGradG_el :879-939) and GenerateJacobian_cw on the element basis,
  on arbitrary fields and on solved ones.
* The mvec overloads (jacobian.cc:86-101, :160-177) for the nodal, element
  and raster bases: the projections computed from the measurement vectors;
  the element basis laid out as the reference's overload lays it out (its
  kappa view at offset nlen, :397).

The oracle's element-basis assembly and Jacobian are pinned in
tests/test_oracle.py (closed-form local matrices, finite differences in the
element parameters).
"""
import numpy as np
import pytest
from scipy import sparse

from oracle import oracle as O
from tests.cases import CASES

pytestmark = pytest.mark.gpu

T = pytest.importorskip("toastpp_amd")
from toastpp_amd import _lib as L  # noqa: E402
from toastpp_amd.jacobian import jacobian, jacobian_cw  # noqa: E402

NAMES = ["2d_real_cg", "3d_cplx_direct"]


def _random_fields(rng, n, k):
    return (rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))) * np.exp(-rng.random((n, 1)) * 5)


def _el_params(rng, E):
    return 0.01 + 0.02 * rng.random(E), 1.0 + rng.random(E), 1.3 + 0.2 * rng.random(E)


@pytest.fixture(scope="module", params=NAMES)
def pair(request):
    case = CASES[request.param]
    om = O.Mesh(case["mesh"])
    gm = T.Mesh(case["mesh"])
    # a few sources and detectors of the case's QM file, ragged links
    Q, M, lp, li = om.read_qm(case["qm"])
    Qs, Ms = Q[:3], M[:4]
    links = [[0, 1, 3], [2], [0, 1, 2, 3]]
    gm.SetQM(Qs, Ms, links)
    lp2 = np.array([0, 3, 4, 8], np.int32)
    li2 = np.array([0, 1, 3, 2, 0, 1, 2, 3], np.int32)
    return request.param, case, om, gm, lp2, li2


@pytest.mark.parametrize("freq", [0.0, 100.0])
def test_element_basis_system_matrix(pair, freq):
    name, case, om, gm, _, _ = pair
    E = om.elen
    rng = np.random.default_rng(21)
    mua, mus, ref = _el_params(rng, E)
    cm, ck, c2 = O.coeffs(mua, mus, ref)
    omega = freq * 2 * np.pi * 1e-6
    _, _, oval = om.sysmat_el(cm, ck, c2, omega)
    f = T.FwdSolver(gm, "BICGSTAB", 1e-10, complex=freq != 0.0)
    f.Allocate()
    sol = T.Solution(mua, mus, ref)
    f.AssembleSystemMatrix(sol, omega, elbasis=True)
    rp, ci, val = f.SystemMatrix()
    ref_val = oval if freq else oval.real
    if om.dim == 3:
        assert np.array_equal(val, ref_val)
    else:
        assert np.max(np.abs(val - ref_val)) <= 1e-15 * np.max(np.abs(ref_val))
    # a nodal-length solution is refused for the element basis, and vice versa
    with pytest.raises(ValueError):
        f.AssembleSystemMatrix(T.Solution(*(a[:om.n] for a in (mua, mus, ref))), omega, elbasis=True)
    with pytest.raises(ValueError):
        f.AssembleSystemMatrix(sol, omega, elbasis=False)


def test_element_basis_fields_vs_oracle(pair):
    name, case, om, gm, _, _ = pair
    E = om.elen
    rng = np.random.default_rng(22)
    mua, mus, ref = _el_params(rng, E)
    cm, ck, c2 = O.coeffs(mua, mus, ref)
    omega = 100.0 * 2 * np.pi * 1e-6
    rp, ci, val = om.sysmat_el(cm, ck, c2, omega)
    f = T.FwdSolver(gm, "BICGSTAB", 1e-12)
    f.SetMode("mixed")
    f.SetPrecon("DIAG")
    f.Allocate()
    f.Reset(T.Solution(mua, mus, ref), omega, elbasis=True)
    qv = sparse.csr_matrix(gm.Qvec(width=2.0).T)
    phi = f.CalcFields(qv)
    for i in range(qv.shape[0]):
        x = O.bicgstab(rp, ci, val, qv[i].toarray()[0].astype(np.complex128), 1e-14, 20000, True)[0]
        assert np.linalg.norm(phi[i] - x) / np.linalg.norm(x) <= 1e-8


@pytest.mark.parametrize("scale", ["log", "lin"])
def test_element_jacobian_random_fields(pair, scale):
    name, case, om, gm, lp, li = pair
    n, E = om.n, om.elen
    nq, nm, nqm = gm.qm_dims()
    rng = np.random.default_rng(23)
    dphi, aphi = _random_fields(rng, n, nq), _random_fields(rng, n, nm)
    proj = rng.standard_normal(nqm) + 1j * rng.standard_normal(nqm)
    Jg = jacobian(gm, "element", dphi, aphi, proj, scale=scale)
    Jo = O.jacobian_el(om, dphi.T, aphi.T, lp, li, proj, logscale=scale == "log")
    assert Jg.shape == (2 * nqm, 2 * E)
    # same products, sums and division in the same order: identical
    assert np.max(np.abs(Jg - Jo)) <= 1e-13 * np.max(np.abs(Jo))
    # the device entry point on resident fields with row strides wider than the counts
    dd = L.DeviceArray.from_host(np.ascontiguousarray(np.hstack([dphi, np.zeros((n, 2))])))
    da = L.DeviceArray.from_host(np.ascontiguousarray(np.hstack([aphi, np.zeros((n, 1))])))
    dp = L.DeviceArray.from_host(proj)
    dJ = L.DeviceArray(Jg.shape, np.float64)
    L.call("tmi_jacobian_element_device", gm.handle_ptr(), nq, nm, L.C.c_void_p(dd.ptr), nq + 2,
           L.C.c_void_p(da.ptr), nm + 1, L.C.c_void_p(dp.ptr), L.DATA_LOG if scale == "log" else L.DATA_LIN,
           L.C.c_void_p(dJ.ptr))
    assert np.array_equal(dJ.to_host(), Jg)


def test_element_jacobian_cw(pair):
    name, case, om, gm, lp, li = pair
    n, E = om.n, om.elen
    nq, nm, nqm = gm.qm_dims()
    rng = np.random.default_rng(24)
    dphi = np.abs(rng.standard_normal((n, nq))) + 0.1
    aphi = np.abs(rng.standard_normal((n, nm))) + 0.1
    proj = np.abs(rng.standard_normal(nqm)) + 0.5
    Jm, Jk = jacobian_cw(gm, "element", dphi, aphi, proj, mua=True, kappa=True)
    Om, Ok = O.jacobian_cw_el(om, dphi.T, aphi.T, lp, li, proj, logscale=True)
    assert Jm.shape == (nqm, E) and Jk.shape == (nqm, E)
    assert np.max(np.abs(Jm - Om)) <= 1e-13 * np.max(np.abs(Om))
    assert np.max(np.abs(Jk - Ok)) <= 1e-13 * np.max(np.abs(Ok))


def test_mvec_overloads_all_bases(pair):
    """GenerateJacobian(raster, mesh, mvec, dphi, aphi, ...): projections from
    mvec (ProjectSingle, DATA_LIN), then the Jacobian as with given
    projections -- for the nodal and raster bases the same as the
    given-projection overload fed the projections of the same fields."""
    name, case, om, gm, lp, li = pair
    n, E = om.n, om.elen
    nq, nm, nqm = gm.qm_dims()
    rng = np.random.default_rng(25)
    dphi, aphi = _random_fields(rng, n, nq), _random_fields(rng, n, nm)
    mvec = sparse.csr_matrix(gm.Mvec(width=2.0, ref=1.4).T)    # nM x n
    proj = O.project(dphi.T, mvec.toarray().astype(np.complex128), lp, li, logscale=False)
    for basis in (None, "raster"):
        rast = None if basis is None else T.Raster(gm, np.array([24] * om.dim))
        Ja = jacobian(gm, rast, dphi, aphi, mvec=mvec)
        Jb = jacobian(gm, rast, dphi, aphi, proj)
        assert np.max(np.abs(Ja - Jb)) <= 1e-12 * np.max(np.abs(Jb)), basis
    Je = jacobian(gm, "element", dphi, aphi, mvec=mvec)
    Jo = O.jacobian_el(om, dphi.T, aphi.T, lp, li, proj, logscale=True, mvec_offset=True)
    assert np.max(np.abs(Je - Jo)) <= 1e-12 * np.max(np.abs(Jo))
    assert E > n and np.array_equal(Je[:, n:E], Je[:, E:2 * E - n])   # the reference's overlapped layout
    # CW: rows scaled by 1 / ProjectAll(mvec, dphi)
    dr, ar = np.abs(dphi), np.abs(aphi)
    mr = sparse.csr_matrix(mvec.real)
    pr = O.project(dr.T, mr.toarray().astype(np.complex128), lp, li, logscale=False).real
    for basis in (None, "element"):
        Ma, Ka = jacobian_cw(gm, basis, dr, ar, mua=True, kappa=True, mvec=mr)
        Mb, Kb = jacobian_cw(gm, basis, dr, ar, pr, mua=True, kappa=True)
        assert np.max(np.abs(Ma - Mb)) <= 1e-12 * np.max(np.abs(Mb)), basis
        assert np.max(np.abs(Ka - Kb)) <= 1e-12 * np.max(np.abs(Kb)), basis


def test_element_jacobian_of_solved_fields_matches_oracle():
    """End to end on the 3D golden mesh: element-basis parameters, fields and
    adjoint fields solved on the GPU, projections, element Jacobian -- against
    the oracle's Jacobian of the oracle's own fields (1e-6, the Jacobian
    tolerance of the forward path)."""
    case = CASES["3d_cplx_direct"]
    om = O.Mesh(case["mesh"])
    gm = T.Mesh(case["mesh"])
    Q, M, lp, li = om.read_qm(case["qm"])
    Qs, Ms = Q[:2], M[:3]
    gm.SetQM(Qs, Ms)
    lp2 = np.array([0, 3, 6], np.int32)
    li2 = np.array([0, 1, 2, 0, 1, 2], np.int32)
    E = om.elen
    rng = np.random.default_rng(26)
    mua, mus, ref = _el_params(rng, E)
    cm, ck, c2 = O.coeffs(mua, mus, ref)
    omega = 100.0 * 2 * np.pi * 1e-6
    f = T.FwdSolver(gm, "BICGSTAB", 1e-13)
    f.SetMode("mixed")
    f.SetPrecon("DIAG")
    f.Allocate()
    f.Reset(T.Solution(mua, mus, ref), omega, elbasis=True)
    qv = sparse.csr_matrix(gm.Qvec(width=2.0).T)
    mv = sparse.csr_matrix(gm.Mvec(width=2.0, ref=1.4).T)
    dphi = f.CalcFields(qv).T
    aphi = f.CalcFields(mv).T
    Jg = jacobian(gm, "element", dphi, aphi, mvec=mv)
    rp, ci, val = om.sysmat_el(cm, ck, c2, omega)
    od = np.array([O.bicgstab(rp, ci, val, qv[i].toarray()[0].astype(complex), 1e-15, 20000, True)[0]
                   for i in range(2)])
    oa = np.array([O.bicgstab(rp, ci, val, mv[i].toarray()[0].astype(complex), 1e-15, 20000, True)[0]
                   for i in range(3)])
    proj = O.project(od, mv.toarray().astype(complex), lp2, li2, logscale=False)
    Jo = O.jacobian_el(om, od, oa, lp2, li2, proj, logscale=True, mvec_offset=True)
    nqm = 6
    for r in (slice(0, nqm), slice(nqm, 2 * nqm)):
        for c in (slice(0, E), slice(E, 2 * E)):
            assert np.linalg.norm(Jg[r, c] - Jo[r, c]) / np.linalg.norm(Jo[r, c]) <= 1e-6
