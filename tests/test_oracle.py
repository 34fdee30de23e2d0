"""The FP64 CPU oracle, pinned before it is trusted as the checker.

Pins (none of them restates the oracle's own formulas):
  * the reference's golden vectors (test/2D/fwdfem, test/3D/fwdfem), for
    PCG, BiCGSTAB, GMRES(20) and the direct targets;
  * closed forms of linear-element integrals (barycentric monomials);
  * structural identities of the assembled operators (row sums, volume,
    perimeter/surface, symmetry);
  * the CSR pattern against a brute-force set of element node pairs;
  * the nodal Jacobian against finite differences of the projections
    (the check of test/2D/matlab/jacobian_test.m).
"""
import itertools
import math

import numpy as np
import pytest
from scipy import sparse

from oracle import oracle as O
from tests.cases import CASES, compare
from toastpp_amd import meshgen as G


@pytest.fixture(scope="module")
def disk():
    nodes, el = G.disk(10.0, 6)
    return O.Mesh(nodes=nodes, elems=el)


@pytest.fixture(scope="module")
def cyl():
    nodes, el = G.cylinder(10.0, 8.0, 4, 3)
    return O.Mesh(nodes=nodes, elems=el)


# ------------------------------------------------------------ golden vectors
@pytest.mark.parametrize("name", ["2d_real_cg", "2d_cplx_bicgstab", "2d_cplx_gmres", "2d_cplx_gmres_diag",
                                  "3d_real_direct", "3d_cplx_direct"])
def test_oracle_golden_vectors(name):
    case = CASES[name]
    m = O.Mesh(case["mesh"])
    qm = m.read_qm(case["qm"])
    n = m.n
    proj, *_ = O.forward(m, qm, np.full(n, case["mua"]), np.full(n, case["mus"]), np.full(n, case["n"]),
                         case["freq"], solver=case["solver"], tol=1e-15, maxit=10000,
                         diag=case["precon"] == "diag")
    for fname, diff, tol in compare(case, proj):
        assert diff <= tol, "%s %s: %.3g > %g" % (name, fname, diff, tol)


def test_golden_mesh_boundary_matches_file_flags():
    """MarkBoundary finds exactly the nodes the mesh files flag 'B'."""
    for path in (CASES["2d_real_cg"]["mesh"], CASES["3d_cplx_direct"]["mesh"]):
        m = O.Mesh(path)
        flags = []
        with open(path) as f:
            for line in f:
                if line[:1] in "NB" and line[1:2] == "[":
                    flags.append(line[0] == "B")
        assert np.array_equal(m.bnd().astype(bool), np.array(flags[:m.n]))


# ------------------------------------------------------ element closed forms
def _bary_int(size, powers, dim):
    """integral of prod l_i^a_i over a simplex of measure `size`."""
    s = sum(powers)
    num = math.prod(math.factorial(a) for a in powers)
    return size * math.factorial(dim) * num / math.factorial(s + dim)


@pytest.mark.parametrize("which", ["disk", "cyl"])
def test_intfff_closed_form(which, disk, cyl):
    m = disk if which == "disk" else cyl
    size = m.size()
    rng = np.random.default_rng(0)
    for e in rng.choice(m.elen, 5, replace=False):
        for i, j, k in itertools.product(range(m.nn), repeat=3):
            p = [0] * m.nn
            for t in (i, j, k):
                p[t] += 1
            assert O.lib().om_el_intfff(m.h, int(e), i, j, k) == pytest.approx(_bary_int(size[e], p, m.dim), rel=1e-14)


@pytest.mark.parametrize("which", ["disk", "cyl"])
def test_mass_and_stiffness_identities(which, disk, cyl):
    m = disk if which == "disk" else cyl
    size = m.size()
    assert np.all(size > 0)
    rp, ci, ff = m.massmat()
    # sum of all mass entries = measure of the domain; row sums = int phi_i
    assert ff.sum() == pytest.approx(size.sum(), rel=1e-13)
    el = m.elements()
    intf = np.zeros(m.n)
    np.add.at(intf, el.ravel(), np.repeat(size / m.nn, m.nn))
    rows = np.add.reduceat(ff, rp[:-1])
    assert np.allclose(rows, intf, rtol=1e-12)
    # pure diffusion operator (cmua = c2a = 0, ckappa = 1): zero row sums, symmetric
    n = m.n
    _, _, K = m.sysmat(np.zeros(n), np.ones(n), np.zeros(n), 0.0)
    rsum = np.add.reduceat(K.real, rp[:-1])
    assert np.max(np.abs(rsum)) <= 1e-12 * np.max(np.abs(K.real))
    from scipy import sparse
    A = sparse.csr_matrix((K.real, ci, rp), shape=(n, n))
    assert abs(A - A.T).max() <= 1e-14 * abs(A).max()


def test_boundary_term_measures(disk, cyl):
    # Tri3 (BndIntPFF = intbff * nodal mean, tri3.h:217-218): sum = perimeter
    n = disk.n
    rp, ci, K = disk.sysmat(np.zeros(n), np.zeros(n), np.ones(n), 0.0)
    x = disk.nodes()
    bnd = disk.bnd().astype(bool)
    ang = np.arctan2(x[bnd, 1], x[bnd, 0])
    xb = x[bnd][np.argsort(ang)]
    perim = np.sum(np.linalg.norm(np.roll(xb, -1, axis=0) - xb, axis=1))
    assert K.real.sum() == pytest.approx(perim, rel=1e-12)
    # Tet4 (exact face integrals): sum = surface area of the polyhedral cylinder
    n = cyl.n
    rp, ci, K = cyl.sysmat(np.zeros(n), np.zeros(n), np.ones(n), 0.0)
    R, H, kr = 10.0, 8.0, 4
    side = 6 * kr
    poly_area = 0.5 * side * R * R * math.sin(2 * math.pi / side)
    poly_perim = side * 2 * R * math.sin(math.pi / side)
    assert K.real.sum() == pytest.approx(2 * poly_area + poly_perim * H, rel=1e-12)


def test_omega_term_is_mass_matrix(cyl):
    n = cyl.n
    om = 0.0123
    rp, ci, K = cyl.sysmat(np.zeros(n), np.zeros(n), np.zeros(n), om)
    _, _, ff = cyl.massmat()
    assert np.allclose(K.imag, om * ff, rtol=1e-15, atol=0)


def test_csr_brute_force(cyl, disk):
    for m in (cyl, disk):
        rp, ci = m.csr()
        el = m.elements()
        pairs = sorted({(a, b) for e in el for a in e for b in e})
        assert len(pairs) == len(ci)
        rows = np.repeat(np.arange(m.n), np.diff(rp))
        assert list(zip(rows.tolist(), ci.tolist())) == pairs


def test_a_keijzer_known_value():
    # n = 1 -> no mismatch: A = 1 (R0 = 0, theta_c = 90 deg)
    assert O.a_keijzer(1.0 + 1e-15) == pytest.approx(1.0, abs=1e-6)
    # monotone increase with n
    assert O.a_keijzer(1.4) > O.a_keijzer(1.33) > 1.0


def test_pull_to_boundary_lands_on_boundary(disk):
    x = disk.nodes()
    bnd = disk.bnd().astype(bool)
    xb = x[bnd]
    for ang in np.linspace(0.1, 6.0, 7):
        p = np.array([11.0 * math.cos(ang), 11.0 * math.sin(ang)])
        out = np.zeros(2)
        assert O.lib().om_pull_to_boundary(disk.h, p, out) == 1
        # the pulled point lies on the line of a boundary edge next to the
        # nearest boundary node: distance to that line ~ 0
        nmin = np.argmin(np.linalg.norm(xb - p, axis=1))
        best = np.inf
        order = np.argsort(np.arctan2(xb[:, 1], xb[:, 0]))
        ring = xb[order]
        k = int(np.where(order == nmin)[0][0])
        for a, b in ((ring[k - 1], ring[k]), (ring[k], ring[(k + 1) % len(ring)])):
            t = b - a
            nrm = np.array([-t[1], t[0]]) / np.linalg.norm(t)
            best = min(best, abs(np.dot(out - a, nrm)))
        assert best < 1e-9
        assert 9.5 < np.linalg.norm(out) < 10.5


# --------------------------------------------------------------- Jacobian FD
def test_nodal_jacobian_finite_differences():
    nodes, el = G.cylinder(6.0, 6.0, 3, 3)
    m = O.Mesh(nodes=nodes, elems=el)
    n = m.n
    Q = np.array([[6.0, 0.0, 1.0], [0.0, 6.0, -1.0]])
    M = np.array([[-6.0, 0.0, 0.0], [0.0, -6.0, 1.0], [4.2, -4.2, 0.0]])
    lp = np.array([0, 3, 6], np.int32)
    li = np.array([0, 1, 2, 0, 1, 2], np.int32)
    rng = np.random.default_rng(3)
    cm = 0.01 + 0.01 * rng.random(n)
    ck = 0.1 + 0.05 * rng.random(n)
    c2 = np.full(n, 0.2)
    omega = 2 * np.pi * 100e-6
    qv = m.qvec_gaussian(Q, 2.0)
    mv = m.qvec_gaussian(M, 2.0) * c2[None, :]

    def fields(cm_, ck_):
        rp, ci, val = m.sysmat(cm_, ck_, c2, omega)
        d = np.array([O.bicgstab(rp, ci, val, q.astype(complex), 1e-15, 5000)[0] for q in qv])
        a = np.array([O.bicgstab(rp, ci, val, q.astype(complex), 1e-15, 5000)[0] for q in mv])
        return d, a

    d, a = fields(cm, ck)
    proj = O.project(d, mv.astype(complex), lp, li, logscale=False)
    J = O.jacobian_nodal(m, d, a, lp, li, proj, logscale=True)
    nqm = len(li)
    for k in rng.choice(n, 4, replace=False):
        for blk, base in ((0, cm), (1, ck)):
            h = 1e-6 * base[k]
            cp, cn = cm.copy(), ck.copy()
            cq, cr = cm.copy(), ck.copy()
            if blk == 0:
                cp[k] += h; cq[k] -= h
            else:
                cn[k] += h; cr[k] -= h
            dp, _ = fields(cp, cn)
            dm, _ = fields(cq, cr)
            yp = O.project(dp, mv.astype(complex), lp, li, logscale=True)
            ym = O.project(dm, mv.astype(complex), lp, li, logscale=True)
            fd = (yp - ym) / (2 * h)
            col = k + blk * n
            jc = J[:nqm, col] + 1j * J[nqm:, col]
            scale = np.max(np.abs(jc))
            assert np.max(np.abs(jc - fd)) <= 1e-5 * scale, (k, blk, jc, fd)


def test_csr_is_per_mesh():
    """Each mesh keeps its own pattern: a mesh created after another was
    freed (possibly at the same address) gets its own CSR."""
    for _ in range(3):
        a = O.Mesh(nodes=G.disk(10.0, 4)[0], elems=G.disk(10.0, 4)[1])
        rpa, _ = a.csr()
        del a
        nodes, el = G.cylinder(6.0, 6.0, 3, 2)
        b = O.Mesh(nodes=nodes, elems=el)
        rp, ci = b.csr()
        pairs = sorted({(x, y) for e in el for x in e for y in e})
        rows = np.repeat(np.arange(b.n), np.diff(rp))
        assert list(zip(rows.tolist(), ci.tolist())) == pairs


def test_qvec_point_closed_forms(disk, cyl):
    rng = np.random.default_rng(9)
    for m in (disk, cyl):
        x = m.nodes()
        el = m.elements()
        # at a node: the unit vector of that node (isotropic source)
        k = int(rng.integers(m.n))
        q = O.qvec_point(m, x[k], neumann=False)[0]
        assert q[k] == pytest.approx(1.0, abs=1e-12) and np.abs(np.delete(q, k)).max() <= 1e-12
        # inside an element: partition of unity and linear reproduction
        e = int(rng.integers(m.elen))
        w = rng.random(m.nn); w /= w.sum()
        p = w @ x[el[e]]
        q = O.qvec_point(m, p, neumann=False)[0]
        assert q.sum() == pytest.approx(1.0, abs=1e-12)
        assert np.allclose(q @ x, p, atol=1e-12)
        # Neumann sources keep boundary nodes only
        qn = O.qvec_point(m, p, neumann=True)[0]
        bnd = m.bnd().astype(bool)
        assert np.all(qn[~bnd] == 0) and np.allclose(qn[bnd], q[bnd])


# --------------------------------------------------------- element basis
def _el_system(m, rp, ci, cm, ck, c2, omega):
    return m.sysmat_el(cm, ck, c2, omega)


@pytest.mark.parametrize("case", ["disk", "cyl"])
def test_element_basis_assembly_closed_forms(case):
    """AssembleSystemMatrix(elbasis = true): constant element coefficients
    give the nodal assembly with the same constants (IntPFF / IntPDD /
    BndIntPFF of a constant P are P IntFF / P IntDD / P BndIntFF); a unit
    coefficient on one element gives that element's local matrices: mass
    V/20 (1 + delta_ij) (Tet4) or A/12 (1 + delta_ij) (Tri3), a stiffness
    with zero row sums and u^T K u = |a|^2 V for u = a.x, and a boundary
    mass whose entries sum to the element's boundary measure."""
    nodes, el = (G.disk(10.0, 5) if case == "disk" else G.cylinder(6.0, 6.0, 3, 3))
    m = O.Mesh(nodes=nodes, elems=el)
    n, E = m.n, m.elen
    omega = 2 * np.pi * 100e-6
    rp, ci, kn = m.sysmat(np.full(n, 0.02), np.full(n, 0.3), np.full(n, 0.15), omega)
    _, _, ke = m.sysmat_el(np.full(E, 0.02), np.full(E, 0.3), np.full(E, 0.15), omega)
    assert np.max(np.abs(ke - kn)) <= 1e-14 * np.max(np.abs(kn))
    rows = np.repeat(np.arange(n), np.diff(rp))
    dim = nodes.shape[1]
    size = m.size()
    bnd_el = np.flatnonzero(m.bndside().reshape(E, -1).any(axis=1))
    for e in (0, E // 2, int(bnd_el[0])):
        nd = el[e]
        unit = np.zeros(E); unit[e] = 1.0
        _, _, kf = m.sysmat_el(unit, 0 * unit, 0 * unit, 0.0)
        _, _, kd = m.sysmat_el(0 * unit, unit, 0 * unit, 0.0)
        _, _, kb = m.sysmat_el(0 * unit, 0 * unit, unit, 0.0)
        inel = np.isin(rows, nd) & np.isin(ci, nd)
        assert np.all(kf[~inel] == 0) and np.all(kd[~inel] == 0) and np.all(kb[~inel] == 0)
        c = 1.0 / 20.0 if dim == 3 else 1.0 / 12.0
        for k in np.flatnonzero(inel):
            assert kf[k].real == pytest.approx(c * size[e] * (2.0 if rows[k] == ci[k] else 1.0), rel=1e-14)
        Kd = sparse.csr_matrix((kd.real, ci, rp), shape=(n, n))
        assert np.max(np.abs(Kd @ np.ones(n))) <= 1e-13 * np.max(np.abs(kd))
        a = np.array([0.3, -1.2, 0.7][:dim])
        u = nodes @ a
        assert u @ (Kd @ u) == pytest.approx(a @ a * size[e], rel=1e-12)
        # boundary measure of the element's boundary sides (Heron areas / edge lengths)
        side_nodes = [[0, 1, 2], [0, 3, 1], [0, 2, 3], [1, 3, 2]] if dim == 3 else [[0, 1], [1, 2], [2, 0]]
        meas = 0.0
        for s_, on in enumerate(m.bndside()[e]):
            if not on:
                continue
            P = nodes[nd[side_nodes[s_]]]
            meas += (np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0])) / 2 if dim == 3
                     else np.linalg.norm(P[1] - P[0]))
        assert kb.real.sum() == pytest.approx(meas, rel=1e-12, abs=1e-300)


def test_element_basis_jacobian_finite_differences():
    """GenerateJacobian(RASTER_ELBASIS) against central finite differences
    of the log projections in per-element cmua / ckappa (the per-element
    check of test/2D/matlab/jacobian_test.m), with the fields of the
    element-basis system matrix."""
    nodes, el = G.cylinder(6.0, 6.0, 3, 3)
    m = O.Mesh(nodes=nodes, elems=el)
    E = m.elen
    Q = np.array([[6.0, 0.0, 1.0], [0.0, 6.0, -1.0]])
    M = np.array([[-6.0, 0.0, 0.0], [0.0, -6.0, 1.0], [4.2, -4.2, 0.0]])
    lp = np.array([0, 3, 6], np.int32)
    li = np.array([0, 1, 2, 0, 1, 2], np.int32)
    rng = np.random.default_rng(4)
    cm = 0.01 + 0.01 * rng.random(E)
    ck = 0.1 + 0.05 * rng.random(E)
    c2 = np.full(E, 0.2)
    omega = 2 * np.pi * 100e-6
    qv = m.qvec_gaussian(Q, 2.0)
    mv = m.qvec_gaussian(M, 2.0) * 0.2

    def fields(cm_, ck_):
        rp, ci, val = m.sysmat_el(cm_, ck_, c2, omega)
        d = np.array([O.bicgstab(rp, ci, val, q.astype(complex), 1e-15, 5000)[0] for q in qv])
        a = np.array([O.bicgstab(rp, ci, val, q.astype(complex), 1e-15, 5000)[0] for q in mv])
        return d, a

    d, a = fields(cm, ck)
    proj = O.project(d, mv.astype(complex), lp, li, logscale=False)
    J = O.jacobian_el(m, d, a, lp, li, proj, logscale=True)
    nqm = len(li)
    assert J.shape == (2 * nqm, 2 * E)
    for e in rng.choice(E, 4, replace=False):
        for blk, base in ((0, cm), (1, ck)):
            h = 1e-6 * base[e]
            cp, cn, cq, cr = cm.copy(), ck.copy(), cm.copy(), ck.copy()
            if blk == 0:
                cp[e] += h; cq[e] -= h
            else:
                cn[e] += h; cr[e] -= h
            yp = O.project(fields(cp, cn)[0], mv.astype(complex), lp, li, logscale=True)
            ym = O.project(fields(cq, cr)[0], mv.astype(complex), lp, li, logscale=True)
            fd = (yp - ym) / (2 * h)
            col = e + blk * E
            jc = J[:nqm, col] + 1j * J[nqm:, col]
            assert np.max(np.abs(jc - fd)) <= 1e-5 * np.max(np.abs(jc)), (e, blk)
    # the element sums of the mua block equal the node sums of the nodal
    # Jacobian on the same fields (sum_i IntFFF(i,j,k) = IntFF(j,k)); for
    # Tet4 so do the kappa blocks (sum_i IntFDD(i,j,k) = IntDD(j,k))
    Jn = O.jacobian_nodal(m, d, a, lp, li, proj, logscale=True)
    for blk in (0, 1):
        se = J[:, blk * E:(blk + 1) * E].sum(1)
        sn = Jn[:, blk * m.n:(blk + 1) * m.n].sum(1)
        assert np.max(np.abs(se - sn)) <= 1e-12 * np.max(np.abs(sn))
    # the mvec overload's kappa view at offset n: mua columns n.. hold kappa
    Jo = O.jacobian_el(m, d, a, lp, li, proj, logscale=True, mvec_offset=True)
    assert np.array_equal(Jo[:, :m.n], J[:, :m.n]) and np.array_equal(Jo[:, E:], J[:, E:])
    assert np.array_equal(Jo[:, m.n:E], J[:, E:2 * E - m.n])
