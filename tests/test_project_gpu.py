"""ProjectAll (fwdsolver.cc:1415-1447) on the GPU against the oracle's serial
dot (om_project, vector_imp.hpp:1045-1052 order), through the C-ABI: long
mvec rows split over several segments, more than 64 sources (two source
blocks per wave), a partial link pattern, host and device field layouts,
complex and real systems, and run-to-run reproducibility of the
fixed-order segment sums."""
import numpy as np
import pytest
from scipy import sparse

from oracle import oracle as O
from tests.cases import CASES

pytestmark = pytest.mark.gpu

TMI = pytest.importorskip("toastpp_amd")
from toastpp_amd import _lib as L  # noqa: E402


@pytest.mark.parametrize("cplx", [True, False])
@pytest.mark.parametrize("nq", [70, 5])
def test_project_all_vs_oracle(cplx, nq):
    case = CASES["3d_cplx_direct"]
    gm = TMI.Mesh(case["mesh"])
    n = gm.NodeCount()
    rng = np.random.default_rng(5 + nq)
    om = O.Mesh(case["mesh"])
    bnd = np.flatnonzero(om.bnd())
    xyz = np.asarray(om.nodes())
    Q = xyz[rng.choice(bnd, nq, replace=False)]
    M = xyz[rng.choice(bnd, 6, replace=False)]
    # every source linked to a random subset of the detectors (at least one)
    links = []
    for q in range(nq):
        sel = np.flatnonzero(rng.random(6) < 0.6)
        links.append(sel if len(sel) else np.array([q % 6]))
    gm.SetQM(Q, M, links=[list(map(int, l)) for l in links])
    nq_, nm, nqm = gm.qm_dims()
    assert (nq_, nm) == (nq, 6)
    lptr = np.zeros(nq + 1, np.int32)
    lptr[1:] = np.cumsum([len(l) for l in links])
    lidx = np.concatenate(links).astype(np.int32)
    # wide profiles: rows over the whole boundary (1377 nodes: several
    # segments per row for either source-block width)
    mv = sparse.csr_matrix(gm.Mvec(width=40.0, ref=1.4).T)
    assert mv.shape == (6, n) and np.diff(mv.indptr).max() > 1024
    dt = np.complex128 if cplx else np.float64
    phi = rng.standard_normal((nq, n)) + (1j * rng.standard_normal((nq, n)) if cplx else 0.0)
    phi = phi.astype(dt)
    mvd = mv.toarray() if cplx else mv.toarray().real
    fws = TMI.FwdSolver(gm, "BICGSTAB" if cplx else "CG", 1e-10, complex=cplx)
    fws.Allocate()
    if not cplx:
        mv = sparse.csr_matrix(mvd)
    p_host = fws.ProjectAll(mv, phi, scale="lin")
    # node-major device fields with a padded row stride
    ld = nq + 3
    Xn = np.zeros((n, ld), dt)
    Xn[:, :nq] = phi.T
    dX = L.DeviceArray.from_host(np.ascontiguousarray(Xn))
    p_dev = fws.ProjectAllDevice(mv, dX.ptr, ld, scale="lin")
    p_dev2 = fws.ProjectAllDevice(mv, dX.ptr, ld, scale="lin")
    assert np.array_equal(p_dev, p_dev2)   # fixed-order sums: reproducible
    po = O.project(phi.astype(np.complex128), mvd.astype(np.complex128), lptr, lidx, logscale=False)
    if not cplx:
        po = po.real
    # each link against the oracle's serial sum, relative to the sum of the
    # magnitudes of its terms (the rounding scale of any summation order)
    absphi, absmv = np.abs(phi), np.abs(mvd)
    for q in range(nq):
        for t, k in enumerate(range(lptr[q], lptr[q + 1])):
            m = lidx[k]
            scale = float(absphi[q] @ absmv[m])
            for p in (p_host, p_dev):
                assert abs(p[k] - po[k]) <= 1e-14 * scale, (q, m, p[k], po[k], scale)
    # the log scaling of the same projections
    pl = fws.ProjectAllDevice(mv, dX.ptr, ld, scale="log")
    if cplx:
        assert np.allclose(pl, np.log(p_dev), rtol=0, atol=1e-13)
    else:
        ok = p_dev > 0
        assert np.allclose(pl[ok], np.log(p_dev[ok]), rtol=0, atol=1e-13)
