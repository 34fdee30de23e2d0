"""GPU nodal Jacobian (GenerateJacobian_mesh, jacobian.cc:441-497) against the
FP64 oracle on the same fields, through the C-ABI.  The oracle's Jacobian is
pinned by finite differences of the projections (test_oracle.py)."""
import numpy as np
import pytest
from scipy import sparse

from oracle import oracle as O
from tests.cases import CASES

pytestmark = pytest.mark.gpu

TMI = pytest.importorskip("toastpp_amd")
from toastpp_amd.jacobian import jacobian  # noqa: E402

TOL = 1e-6   # Jacobian parity bound (relative to the largest entry of each block)


def _blockwise_rel(Jg, Jo, nqm, n):
    worst = 0.0
    for r in (slice(0, nqm), slice(nqm, 2 * nqm)):
        for c in (slice(0, n), slice(n, 2 * n)):
            ref = Jo[r, c]
            worst = max(worst, np.max(np.abs(Jg[r, c] - ref)) / np.max(np.abs(ref)))
    return worst


def _blockwise_fro(Jg, Jo, nqm, n):
    worst = 0.0
    for r in (slice(0, nqm), slice(nqm, 2 * nqm)):
        for c in (slice(0, n), slice(n, 2 * n)):
            worst = max(worst, np.linalg.norm(Jg[r, c] - Jo[r, c]) / np.linalg.norm(Jo[r, c]))
    return worst


def _random_fields(rng, n, k):
    return (rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))) * np.exp(-rng.random((n, 1)) * 5)


@pytest.mark.parametrize("name", ["2d_real_cg", "3d_cplx_direct"])
@pytest.mark.parametrize("scale", ["log", "lin"])
def test_nodal_jacobian_random_fields(name, scale):
    """Arbitrary complex fields: every element integral, index and sign."""
    case = CASES[name]
    om = O.Mesh(case["mesh"])
    Q, Mp, lp, li = om.read_qm(case["qm"])
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    n, nq, nm, nqm = om.n, Q.shape[0], Mp.shape[0], int(lp[-1])
    assert gm.qm_dims() == (nq, nm, nqm)
    rng = np.random.default_rng(11)
    dphi = _random_fields(rng, n, nq)
    aphi = _random_fields(rng, n, nm)
    proj = rng.standard_normal(nqm) + 1j * rng.standard_normal(nqm)
    if scale == "log":
        Jg = gm.Jacobian(None, dphi, aphi, proj)
    else:
        Jg = jacobian(gm, None, dphi, aphi, proj, scale="lin")
    Jo = O.jacobian_nodal(om, dphi.T, aphi.T, lp, li, proj, logscale=scale == "log")
    assert Jg.shape == (2 * nqm, 2 * n)
    rel = _blockwise_rel(Jg, Jo, nqm, n)
    print("%s %s: worst block-relative diff %.3g" % (name, scale, rel))
    assert rel <= 1e-12   # FP64 both sides; only the summation order differs


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_nodal_jacobian_optical_2d(mode):
    """JacobianOptical on the 2D golden mesh: GPU fields, projections and
    Jacobian against the oracle's BiCGSTAB fields and Jacobian.  The log
    Jacobian divides by the projections, which span ~20 decades between near
    and far source/detector pairs, so the far rows carry the fields' error
    relative to their own size: the solves run near the FP64 floor (relative
    residual 3e-13; at 1e-13 a column of the single-column FP64 path can end
    with its true residual just above the tolerance, depending on the
    summation order, and is reported unconverged), and the bound is on each
    block's relative Frobenius norm, with the worst single entry reported and
    held to 1e-5."""
    tol = 3e-13
    case = CASES["2d_cplx_bicgstab"]
    om = O.Mesh(case["mesh"])
    Q, Mp, lp, li = om.read_qm(case["qm"])
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    n = om.n
    rng = np.random.default_rng(5)
    mua = 0.01 + 0.01 * rng.random(n)
    mus = 1.0 + 0.5 * rng.random(n)
    ref = np.full(n, 1.4)
    freq = 100.0
    qvec = gm.Qvec(width=2.0)
    mvec = gm.Mvec(width=2.0, ref=1.4)
    Jg = gm.JacobianOptical(None, qvec, mvec, mua, mus, ref, freq, solver="bicgstab", tol=tol,
                            mode=mode)
    # oracle: same system, tight BiCGSTAB per source and detector
    cm, ck, c2 = O.coeffs(mua, mus, ref)
    rp, ci, val = om.sysmat(cm, ck, c2, freq * 2 * np.pi * 1e-6)
    qv = om.qvec_gaussian(Q, 2.0)
    mv = om.qvec_gaussian(Mp, 2.0) * c2[None, :]
    d = np.array([O.bicgstab(rp, ci, val, q.astype(complex), 1e-15, 20000, True)[0] for q in qv])
    a = np.array([O.bicgstab(rp, ci, val, q.astype(complex), 1e-15, 20000, True)[0] for q in mv])
    proj = O.project(d, mv.astype(complex), lp, li, logscale=False)
    Jo = O.jacobian_nodal(om, d, a, lp, li, proj, logscale=True)
    nqm = int(lp[-1])
    rel = _blockwise_rel(Jg, Jo, nqm, n)
    fro = _blockwise_fro(Jg, Jo, nqm, n)
    print("JacobianOptical 2D %s: block-relative Frobenius %.3g, worst entry %.3g" % (mode, fro, rel))
    assert fro <= TOL
    assert rel <= 10 * TOL


def test_nodal_jacobian_rejects_bad_shapes():
    case = CASES["2d_real_cg"]
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    n = gm.NodeCount()
    nq, nm, nqm = gm.qm_dims()
    with pytest.raises(ValueError):
        gm.Jacobian(None, np.zeros((n, nq + 1), complex), np.zeros((n, nm), complex), np.ones(nqm, complex))


@pytest.mark.parametrize("name", ["2d_real_cg", "3d_cplx_direct"])
@pytest.mark.parametrize("scale", ["log", "lin"])
def test_nodal_jacobian_cw_random_fields(name, scale):
    """CW nodal Jacobian (GenerateJacobian_cw_mesh): real fields, mua and
    kappa blocks against the oracle's element-integral loops."""
    from toastpp_amd.jacobian import jacobian_cw
    case = CASES[name]
    om = O.Mesh(case["mesh"])
    Q, Mp, lp, li = om.read_qm(case["qm"])
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    n, nq, nm, nqm = om.n, Q.shape[0], Mp.shape[0], int(lp[-1])
    rng = np.random.default_rng(21)
    dphi = rng.standard_normal((n, nq)) * np.exp(-rng.random((n, 1)) * 5)
    aphi = rng.standard_normal((n, nm)) * np.exp(-rng.random((n, 1)) * 5)
    proj = rng.random(nqm) + 0.5
    Jm, Jk = jacobian_cw(gm, None, dphi, aphi, proj if scale == "log" else None, scale=scale, kappa=True)
    Om, Ok = O.jacobian_cw_nodal(om, dphi.T, aphi.T, lp, li, proj, logscale=scale == "log")
    assert Jm.shape == Jk.shape == (nqm, n)
    for g, o, blk in ((Jm, Om, "mua"), (Jk, Ok, "kappa")):
        rel = np.max(np.abs(g - o)) / np.max(np.abs(o))
        print("%s %s CW %s: max diff / max |J| = %.3g" % (name, scale, blk, rel))
        assert rel <= 1e-12   # FP64 both sides; only the summation order differs
