"""GPU raster with a basis grid coarser than the intermediate grid
(Raster_Pixel(bdim, gdim), raster_px.cc:10-54; toast.Raster(mesh, grd,
intgrd)) against the oracle, through the C-ABI: the maps G
(Grid2LinPixMatrix) and GI (LinPix2GridMatrix), the solution voxels, the
maps between every pair of bases, and the grid Jacobians and the data
gradient taken through G (Map_GridToSol).  The oracle's maps are pinned in
test_oracle_raster.py."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import CASES

pytestmark = pytest.mark.gpu
# fields near the FP64 floor: at 1e-13 a column of the single-column FP64
# path can end with its true residual just above the tolerance, depending on
# the summation order, and is reported unconverged
FP64_FLOOR_TOL = 3e-13

TMI = pytest.importorskip("toastpp_amd")

# (basis grid, intermediate grid): unaligned and aligned (every basis point a grid point)
GRIDS = {"2d": ([24, 20], [64, 64]), "2d_aligned": ([33, 17], [65, 65]), "3d": ([12, 9, 10], [24, 22, 20])}


@pytest.fixture(scope="module", params=sorted(GRIDS))
def pair(request):
    case = CASES["3d_cplx_direct" if request.param == "3d" else "2d_real_cg"]
    om = O.Mesh(case["mesh"])
    gm = TMI.Mesh(case["mesh"])
    bd, gd = GRIDS[request.param]
    return om, gm, O.Raster(om, gd, bdim=bd), TMI.Raster(gm, np.array(bd), np.array(gd))


def _dense(rp, ci, val, shape):
    A = np.zeros(shape)
    for r in range(shape[0]):
        A[r, ci[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    return A


def test_basis_structure(pair):
    om, gm, orr, gr = pair
    print("basis %s grid %s: slen %d" % (list(orr.bdim), list(orr.gdim), orr.slen), flush=True)
    assert gr.GLen() == orr.glen and gr.BLen() == orr.blen and gr.SLen() == orr.slen
    assert np.array_equal(gr.Elref(), orr.elref)
    assert np.array_equal(gr.Sol2Basis(), orr.sol2basis)
    Gg = gr.GetMatrix("G->B")
    rp, ci, val = orr.gmap()
    assert np.array_equal(Gg.indptr, rp) and np.array_equal(Gg.indices, ci)
    assert np.array_equal(Gg.data, val)   # same operations: bit-identical weights
    GIg = gr.GetMatrix("B->G")
    rp, ci, val = orr.gimap()
    assert np.array_equal(GIg.indptr, rp) and np.array_equal(GIg.indices, ci)
    assert np.array_equal(GIg.data, val)


def test_basis_maps(pair):
    om, gm, orr, gr = pair
    rng = np.random.default_rng(4)
    G = _dense(*orr.gmap(), (orr.blen, orr.glen))
    GI = _dense(*orr.gimap(), (orr.glen, orr.blen))
    g = rng.standard_normal((orr.glen, 2)) + 1j * rng.standard_normal((orr.glen, 2))
    b = gr.GridToBasis(g)
    assert np.max(np.abs(b - G @ g)) <= 1e-13 * np.abs(G @ g).max()
    bv = rng.standard_normal(orr.blen)
    gv = gr.BasisToGrid(bv)
    assert not np.iscomplexobj(gv)
    assert np.max(np.abs(gv - GI @ bv)) <= 1e-13 * np.abs(GI @ bv).max()
    # M -> S through the chain equals D G B applied to the field
    m = rng.standard_normal(om.n)
    Bm = gr.GetMatrix("M->G")
    ref = (G @ (Bm @ m))[orr.sol2basis]
    s = gr.Map("M->S", m)
    assert np.max(np.abs(s - ref)) <= 1e-13 * np.abs(ref).max()
    assert np.allclose(gr.GetMatrix("M->B") @ m, G @ (Bm @ m), rtol=0, atol=1e-13 * np.abs(ref).max())
    # S -> M: scatter, GI, NodeMapMatrix
    sv = rng.standard_normal(orr.slen)
    full = np.zeros(orr.blen)
    full[orr.sol2basis] = sv
    ref_m = gr.GetMatrix("G->M") @ (GI @ full)
    assert np.allclose(gr.Map("S->M", sv), ref_m, rtol=0, atol=1e-13 * np.abs(ref_m).max())


def _links(gm, dim, rng):
    nq, nm = 3, 4
    gm.SetQM(rng.standard_normal((nq, dim)), rng.standard_normal((nm, dim)), links=[[0, 2, 3], [1], [0, 1, 2, 3]])
    return nq, nm, np.array([0, 3, 4, 8], np.int32), np.array([0, 2, 3, 1, 0, 1, 2, 3], np.int32)


@pytest.mark.parametrize("scale", ["log", "lin"])
def test_basis_jacobian_vs_oracle(pair, scale):
    om, gm, orr, gr = pair
    n, dim = om.n, om.dim
    rng = np.random.default_rng(7)
    nq, nm, lp, li = _links(gm, dim, rng)
    nqm = len(li)
    dphi = rng.standard_normal((n, nq)) + 1j * rng.standard_normal((n, nq))
    aphi = rng.standard_normal((n, nm)) + 1j * rng.standard_normal((n, nm))
    proj = rng.standard_normal(nqm) + 1j * rng.standard_normal(nqm)
    from toastpp_amd.jacobian import jacobian
    Jg = jacobian(gm, gr, dphi, aphi, proj, scale=scale)
    Jo = O.jacobian_grid(orr, dphi.T, aphi.T, lp, li, proj, logscale=scale == "log")
    assert Jg.shape == Jo.shape == (2 * nqm, 2 * orr.slen)
    rel = np.max(np.abs(Jg - Jo)) / np.max(np.abs(Jo))
    print("basis Jacobian %s %s: max diff / max |J| = %.3g" % (list(orr.bdim), scale, rel))
    assert rel <= 1e-13


@pytest.mark.parametrize("scale", ["log", "lin"])
def test_basis_jacobian_cw_vs_oracle(pair, scale):
    om, gm, orr, gr = pair
    n, dim = om.n, om.dim
    rng = np.random.default_rng(17)
    nq, nm, lp, li = _links(gm, dim, rng)
    dphi = rng.standard_normal((n, nq))
    aphi = rng.standard_normal((n, nm))
    proj = rng.random(len(li)) + 0.5
    from toastpp_amd.jacobian import jacobian_cw
    Jm, Jk = jacobian_cw(gm, gr, dphi, aphi, proj, scale=scale, mua=True, kappa=True)
    Om, Ok = O.jacobian_grid_cw(orr, dphi.T, aphi.T, lp, li, proj, logscale=scale == "log")
    assert Jm.shape == Om.shape == (len(li), orr.slen)
    for g, o, name in ((Jm, Om, "mua"), (Jk, Ok, "kappa")):
        rel = np.max(np.abs(g - o)) / np.max(np.abs(o))
        print("basis CW Jacobian %s %s %s: max diff / max |J| = %.3g" % (list(orr.bdim), scale, name, rel))
        assert rel <= 1e-13
    Jk2 = jacobian_cw(gm, gr, dphi, aphi, proj, scale=scale, mua=False, kappa=True)[1]
    assert np.array_equal(Jk2, Jk)


def test_basis_data_gradient_vs_oracle(pair):
    om, gm, orr, gr = pair
    n = om.n
    rng = np.random.default_rng(23)
    nq = 3
    d = rng.standard_normal((n, nq)) + 1j * rng.standard_normal((n, nq))
    a = rng.standard_normal((n, nq)) + 1j * rng.standard_normal((n, nq))
    g0 = rng.standard_normal(2 * orr.slen)
    gg = gr.DataGradient(d, a, g0)
    go = O.data_gradient(orr, d.T, a.T, g0)
    rel = np.max(np.abs(gg - go)) / np.max(np.abs(go - g0))
    print("basis data gradient %s: max diff / max |increment| = %.3g" % (list(orr.bdim), rel))
    assert rel <= 1e-13


def test_basis_jacobian_optical_2d():
    """JacobianOptical with a 32 x 32 basis over a 64 x 64 intermediate grid:
    GPU fields and Jacobian against the oracle's fields and basis Jacobian
    (1e-6 block Frobenius)."""
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
    gr = TMI.Raster(gm, np.array([32, 32]), np.array([64, 64]))
    Jg = gm.JacobianOptical(gr, gm.Qvec(width=2.0), gm.Mvec(width=2.0, ref=1.4), mua, mus, ref, freq,
                            solver="bicgstab", tol=FP64_FLOOR_TOL, mode="fp64")
    cm, ck, c2 = O.coeffs(mua, mus, ref)
    rp, ci, val = om.sysmat(cm, ck, c2, freq * 2 * np.pi * 1e-6)
    qv = om.qvec_gaussian(Q, 2.0)
    mv = om.qvec_gaussian(Mp, 2.0) * c2[None, :]
    d = np.array([O.bicgstab(rp, ci, val, q.astype(complex), 1e-15, 20000, True)[0] for q in qv])
    a = np.array([O.bicgstab(rp, ci, val, q.astype(complex), 1e-15, 20000, True)[0] for q in mv])
    proj = O.project(d, mv.astype(complex), lp, li, logscale=False)
    Jo = O.jacobian_grid(O.Raster(om, [64, 64], bdim=[32, 32]), d, a, lp, li, proj, logscale=True)
    nqm, slen = int(lp[-1]), gr.SLen()
    assert Jg.shape == Jo.shape == (2 * nqm, 2 * slen)
    worst = 0.0
    for r in (slice(0, nqm), slice(nqm, 2 * nqm)):
        for c in (slice(0, slen), slice(slen, 2 * slen)):
            worst = max(worst, np.linalg.norm(Jg[r, c] - Jo[r, c]) / np.linalg.norm(Jo[r, c]))
    print("basis JacobianOptical 2D: block-relative Frobenius %.3g" % worst)
    assert worst <= 1e-6
