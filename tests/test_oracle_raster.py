"""The oracle's raster basis and grid Jacobian (Raster_Pixel with basis grid
== intermediate grid, GenerateJacobian_grid), pinned before it is trusted:

  * element of every grid point against a brute-force barycentric search;
  * grid map B: partition of unity and exact reproduction of linear fields;
  * ImageGradient: closed form on linear fields (spacing gsize/n, the
    reference's choice, jacobian.cc:961) and its mask rules on a hand case;
  * grid Jacobian: its voxel sums are a quadrature of the same integrals the
    nodal Jacobian's node sums give exactly (sum_i phi_i = 1), so the two
    must agree to quadrature error -- this fixes sign, Re/Im split, link
    order, 1/proj and the voxel-size scale.
"""
import numpy as np
import pytest

from oracle import oracle as O
from toastpp_amd import meshgen as G


def _barycentric_elref(nodes, el, pts, tol=1e-8):
    """min element index whose barycentric coordinates at p are >= -tol."""
    dim = nodes.shape[1]
    out = np.full(len(pts), -1)
    for e in range(len(el)):
        X = nodes[el[e]]
        T = (X[1:] - X[0]).T
        lam = np.linalg.solve(T, (pts - X[0]).T).T
        full = np.c_[1.0 - lam.sum(1), lam]
        inside = np.all(full >= -tol * 10, axis=1)
        sel = inside & (out < 0)
        out[sel] = e
    return out


@pytest.mark.parametrize("case", ["disk", "cyl"])
def test_elref_brute_force(case):
    if case == "disk":
        nodes, el = G.disk(10.0, 4)
        gdim = [13, 11]
    else:
        nodes, el = G.cylinder(6.0, 6.0, 2, 3)
        gdim = [7, 6, 5]
    m = O.Mesh(nodes=nodes, elems=el)
    R = O.Raster(m, gdim)
    axes = [R.bbmin[d] + np.arange(gdim[d]) * (R.bbmax[d] - R.bbmin[d]) / (gdim[d] - 1) for d in range(len(gdim))]
    mesh_ = np.meshgrid(*axes, indexing="ij")
    pts = np.stack([g.transpose().ravel() for g in mesh_], axis=1)   # x fastest
    ref = _barycentric_elref(nodes, el, pts)
    # points well inside or well outside agree exactly; points on shared
    # faces belong to the lowest-numbered element in both
    assert np.array_equal(R.elref, ref)
    assert R.slen == int(np.count_nonzero(ref >= 0))
    assert np.all(np.diff(R.sol2basis) > 0)


@pytest.mark.parametrize("case", ["disk", "cyl"])
def test_grid_map_linear_reproduction(case):
    nodes, el = (G.disk(10.0, 5) if case == "disk" else G.cylinder(6.0, 6.0, 3, 3))
    m = O.Mesh(nodes=nodes, elems=el)
    gdim = [21, 17] if case == "disk" else [9, 9, 8]
    R = O.Raster(m, gdim)
    cols, w = R.bmap()
    inside = R.elref >= 0
    assert np.allclose(w[inside].sum(1), 1.0, atol=1e-12)
    assert np.all(w[~inside] == 0) and np.all(cols[~inside] == -1)
    coef = np.array([0.7, -1.3, 0.4][:nodes.shape[1]])
    f = nodes @ coef + 2.5
    axes = [R.bbmin[d] + np.arange(gdim[d]) * (R.bbmax[d] - R.bbmin[d]) / (gdim[d] - 1) for d in range(len(gdim))]
    pts = np.stack([g.transpose().ravel() for g in np.meshgrid(*axes, indexing="ij")], axis=1)
    mapped = (w * np.where(cols >= 0, f[np.maximum(cols, 0)], 0.0)).sum(1)
    assert np.allclose(mapped[inside], pts[inside] @ coef + 2.5, rtol=1e-12, atol=1e-12)


def test_image_gradient_closed_form_and_mask_rules():
    gdim = [8, 6, 5]
    gsize = np.array([7.0, 10.0, 4.0])
    x = [np.arange(g) * s / (g - 1) for g, s in zip(gdim, gsize)]
    X = np.meshgrid(*x, indexing="ij")
    a = np.array([1.5 + 0.5j, -2.0, 0.25j])
    im = sum(a[d] * X[d] for d in range(3)).transpose().ravel()
    mask = np.zeros(im.size, np.int32)
    g = O.image_gradient(gdim, gsize, im, mask)
    for d in range(3):
        # grid spacing gsize/(n-1) over a finite-difference step gsize/n
        assert np.allclose(g[d], a[d] * gdim[d] / (gdim[d] - 1), rtol=1e-13, atol=1e-13)
    # mask rules along x on a 5 x 2 grid: centre masked with both neighbours
    # inside still takes the central difference (the reference tests the
    # neighbours first); a point with one masked neighbour uses the other
    gdim = [5, 2]
    im = np.arange(10, dtype=float) ** 2 + 0j
    mask = np.array([0, -1, 0, 0, -1,   0, 0, 0, -1, -1], np.int32)
    g = O.image_gradient(gdim, np.array([5.0, 2.0]), im, mask)[0].real
    ih = 5 / 5.0
    exp = np.array([0.0,                          # x=0: right neighbour masked, no left -> 0
                    (im[2] - im[0]).real * ih / 2,   # x=1 masked, both neighbours in mask -> central
                    (im[3] - im[2]).real * ih,      # x=2: left masked -> forward
                    (im[3] - im[2]).real * ih,      # x=3: right masked -> backward
                    0.0,                          # x=4 masked, left in mask but no right -> 0
                    (im[6] - im[5]).real * ih,
                    (im[7] - im[5]).real * ih / 2,
                    (im[7] - im[6]).real * ih,
                    0.0, 0.0])
    assert np.allclose(g, exp, rtol=1e-14)


def _fields(m, Q, M, mua, mus, ref, freq):
    cm, ck, c2 = O.coeffs(mua, mus, ref)
    rp, ci, val = m.sysmat(cm, ck, c2, freq * 2 * np.pi * 1e-6)
    qv = m.qvec_gaussian(Q, 2.0)
    mv = m.qvec_gaussian(M, 2.0) * c2[None, :]
    d = np.array([O.bicgstab(rp, ci, val, q.astype(complex), 1e-14, 20000, True)[0] for q in qv])
    a = np.array([O.bicgstab(rp, ci, val, q.astype(complex), 1e-14, 20000, True)[0] for q in mv])
    return d, a, mv


def test_grid_jacobian_quadrature_of_nodal_jacobian():
    nodes, el = G.disk(25.0, 16)
    m = O.Mesh(nodes=nodes, elems=el)
    n = m.n
    Q, M = G.ring(4, 25.0), G.ring(4, 25.0, offset=0.5)
    lp = np.array([0, 2, 4, 6, 8], np.int32)
    li = np.array([1, 2, 2, 3, 3, 0, 0, 1], np.int32)
    mua, mus, ref = np.full(n, 0.01), np.full(n, 1.0), np.full(n, 1.4)
    d, a, mv = _fields(m, Q, M, mua, mus, ref, 100.0)
    proj = O.project(d, mv.astype(complex), lp, li, logscale=False)
    Jn = O.jacobian_nodal(m, d, a, lp, li, proj, logscale=True)
    R = O.Raster(m, [96, 96])
    Jg = O.jacobian_grid(R, d, a, lp, li, proj, logscale=True)
    nqm = len(li)
    assert Jg.shape == (2 * nqm, 2 * R.slen)
    sn_mua, sg_mua = Jn[:, :n].sum(1), Jg[:, :R.slen].sum(1)
    sn_kap, sg_kap = Jn[:, n:].sum(1), Jg[:, R.slen:].sum(1)
    # the mua integrand is sampled exactly at the voxels: quadrature error only
    # (0.13 % at 96 x 96, 0.09 % at 192 x 192)
    assert np.max(np.abs(sg_mua / sn_mua - 1)) < 0.005
    # the kappa integrand uses finite differences with step gsize/n
    # (4.6 % at 96 x 96, 2.2 % at 192 x 192)
    assert np.max(np.abs(sg_kap / sn_kap - 1)) < 0.07
    # the wrong sign, a swapped Re/Im block or a missing voxel scale fails
    assert not np.allclose(-sg_mua, sn_mua, rtol=0.5)


def test_grid_jacobian_cw_oracle():
    """CW grid Jacobian (GenerateJacobian_cw_grid): its mua block equals the
    real part of the (pinned) complex grid Jacobian on the same real fields;
    its kappa block equals it at voxels whose neighbours are all inside the
    mesh and differs next to the boundary, where the reference takes the
    adjoint gradients without the mask; voxel sums match the CW nodal
    Jacobian's node sums (quadrature)."""
    nodes, el = G.disk(25.0, 16)
    m = O.Mesh(nodes=nodes, elems=el)
    n = m.n
    Q, M = G.ring(4, 25.0), G.ring(4, 25.0, offset=0.5)
    lp = np.array([0, 2, 4, 6, 8], np.int32)
    li = np.array([1, 2, 2, 3, 3, 0, 0, 1], np.int32)
    mua, mus, ref = np.full(n, 0.01), np.full(n, 1.0), np.full(n, 1.4)
    d, a, mv = _fields(m, Q, M, mua, mus, ref, 0.0)
    d, a, mv = d.real, a.real, mv.real
    proj = O.project(d.astype(complex), mv.astype(complex), lp, li, logscale=False).real
    R = O.Raster(m, [96, 96])
    nqm, slen = len(li), R.slen
    Om, Ok = O.jacobian_grid_cw(R, d, a, lp, li, proj, logscale=True)
    Jc = O.jacobian_grid(R, d.astype(complex), a.astype(complex), lp, li, proj.astype(complex), logscale=True)
    scale = np.max(np.abs(Om))
    assert np.max(np.abs(Om - Jc[:nqm, :slen])) <= 1e-14 * scale
    # interior voxels: both gradients see only inside points
    g = R.sol2basis
    nx, ny = R.gdim
    ix, iy = g % nx, g // nx
    inside = R.elref >= 0
    interior = np.ones(slen, bool)
    for dx, dy in ((1, 0), (-1, 0), (0, 1), (0, -1)):
        jx, jy = ix + dx, iy + dy
        ok = (jx >= 0) & (jx < nx) & (jy >= 0) & (jy < ny)
        interior &= ok & inside[np.clip(jy, 0, ny - 1) * nx + np.clip(jx, 0, nx - 1)]
    ks = np.max(np.abs(Ok))
    assert np.max(np.abs(Ok[:, interior] - Jc[:nqm, slen:][:, interior])) <= 1e-13 * ks
    assert np.max(np.abs(Ok[:, ~interior] - Jc[:nqm, slen:][:, ~interior])) > 1e-6 * ks
    # quadrature against the CW nodal Jacobian
    Nm, Nk = O.jacobian_cw_nodal(m, d, a, lp, li, proj, logscale=True)
    assert np.max(np.abs(Om.sum(1) / Nm.sum(1) - 1)) < 0.005
    assert np.max(np.abs(Ok.sum(1) / Nk.sum(1) - 1)) < 0.07


def test_data_gradient_oracle_matches_jacobian_identity():
    """AddDataGradient against the (pinned) grid Jacobian: with adjoint
    fields a_q = sum_m w_qm aphi_m (the adjoint source is linear in the
    measurement vectors), grad = sum_k (t_mod,k J[k] + t_arg,k J[k + nQM]) / sz
    with t = -2 (data - proj) / sd^2, J the log-data grid Jacobian and sz the
    voxel size the Jacobian carries and the gradient does not."""
    nodes, el = G.disk(25.0, 16)
    m = O.Mesh(nodes=nodes, elems=el)
    n = m.n
    Q, M = G.ring(4, 25.0), G.ring(4, 25.0, offset=0.5)
    lp = np.array([0, 2, 4, 6, 8], np.int32)
    li = np.array([1, 2, 2, 3, 3, 0, 0, 1], np.int32)
    mua, mus, ref = np.full(n, 0.01), np.full(n, 1.0), np.full(n, 1.4)
    d, a, mv = _fields(m, Q, M, mua, mus, ref, 100.0)
    nqm = len(li)
    cproj = O.project(d, mv.astype(complex), lp, li, logscale=False)
    lproj = np.log(cproj)
    preal = np.concatenate([lproj.real, lproj.imag])
    rng = np.random.default_rng(3)
    data = preal + 0.05 * rng.standard_normal(2 * nqm)
    sd = 0.5 + rng.random(2 * nqm)
    wq = O.adjoint_sources(cproj, preal, data, sd, mv.astype(complex), lp, li)
    # adjoint fields by linearity: the weights of wq applied to aphi
    W = np.zeros((4, 4), complex)
    for q in range(4):
        for k in range(lp[q], lp[q + 1]):
            rp, ip = cproj[k].real, cproj[k].imag
            dn = 1.0 / (rp * rp + ip * ip)
            ta = -2.0 * (data[k] - preal[k]) / sd[k] / sd[k]
            tp = -2.0 * (data[nqm + k] - preal[nqm + k]) / sd[nqm + k] / sd[nqm + k]
            W[q, li[k]] += complex(ta * rp * dn, -ta * ip * dn) + complex(-tp * ip * dn, -tp * rp * dn)
    assert np.allclose(wq, W @ mv, rtol=1e-13, atol=1e-13 * np.abs(wq).max())
    aq = W @ a
    R = O.Raster(m, [64, 64])
    g = O.data_gradient(R, d, aq)
    J = O.jacobian_grid(R, d, a, lp, li, cproj, logscale=True)
    sz = np.prod([(R.bbmax[i] - R.bbmin[i]) / (R.gdim[i] - 1) for i in range(2)])
    t_mod = -2.0 * (data[:nqm] - preal[:nqm]) / sd[:nqm] ** 2
    t_arg = -2.0 * (data[nqm:] - preal[nqm:]) / sd[nqm:] ** 2
    ref_g = (t_mod @ J[:nqm] + t_arg @ J[nqm:]) / sz
    assert g.shape == (2 * R.slen,)
    scale = np.abs(ref_g).max()
    assert np.max(np.abs(g - ref_g)) <= 1e-10 * scale
    # a sign or a block swap fails
    assert np.max(np.abs(g + ref_g)) > 1e-3 * scale


def _case_mesh(case):
    if case == "2d":
        nodes, el = G.disk(10.0, 6)
        return O.Mesh(nodes=nodes, elems=el), [17, 15]
    nodes, el = G.cylinder(8.0, 10.0, 4, 5)
    return O.Mesh(nodes=nodes, elems=el), [11, 10, 9]


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_nodemap_partition_and_linear_reproduction(case):
    """NodeMapMatrix: every node inside the box gets weights summing to 1;
    where all 2^dim surrounding grid points are inside the mesh the weights
    are the bi/trilinear ones, which reproduce a linear field exactly."""
    m, gd = _case_mesh(case)
    R = O.Raster(m, gd)
    cols, w = R.nodemap()
    nc = 1 << m.dim
    has = (cols >= 0).any(1)
    assert has.all()                                   # the mesh lies in its own bounding box
    assert np.allclose(w.sum(1), 1.0, rtol=0, atol=1e-13)
    full = (cols >= 0).sum(1) == nc
    assert full.mean() > 0.5
    # grid point coordinates, x fastest
    dW = (R.bbmax - R.bbmin) / (R.gdim - 1)
    idx = np.arange(R.glen)
    pos = []
    rest = idx.copy()
    for d in range(m.dim):
        pos.append(R.bbmin[d] + (rest % R.gdim[d]) * dW[d])
        rest //= R.gdim[d]
    a = np.array([0.7, -1.3, 0.4][:m.dim])
    gval = 2.0 + sum(a[d] * pos[d] for d in range(m.dim))
    x = m.nodes()
    f = np.where(cols >= 0, w * gval[np.maximum(cols, 0)], 0.0).sum(1)
    exact = 2.0 + x @ a
    assert np.max(np.abs(f[full] - exact[full])) <= 1e-11 * np.max(np.abs(exact))
    # partial support still reproduces constants (renormalised weights)
    fc = np.where(cols >= 0, w * 3.0, 0.0).sum(1)
    assert np.allclose(fc, 3.0, rtol=1e-13)


# ------------------------------------------------ basis grid coarser than the grid

def _tent_dense(R, masked=True):
    """Brute force of the linear pixel tents: w[b, g] = prod_d max(0, 1 -
    |u_g - u_b| (bdim_d - 1)) in unit-box coordinates, zero at grid points
    outside the mesh (elref < 0) when masked."""
    W = np.ones((R.blen, R.glen))
    for d in range(R.dim):
        sb = int(np.prod(R.bdim[:d])) if d else 1
        sg = int(np.prod(R.gdim[:d])) if d else 1
        ub = (np.arange(R.blen) // sb % R.bdim[d]) / (R.bdim[d] - 1.0)
        ug = (np.arange(R.glen) // sg % R.gdim[d]) / (R.gdim[d] - 1.0)
        W *= np.maximum(0.0, 1.0 - np.abs(ug[None, :] - ub[:, None]) * (R.bdim[d] - 1.0))
    if masked:
        W[:, R.elref < 0] = 0.0
    return W


def _dense(rp, ci, val, shape):
    A = np.zeros(shape)
    for r in range(shape[0]):
        A[r, ci[rp[r]:rp[r + 1]]] = val[rp[r]:rp[r + 1]]
    return A


def _coarse_case(case, aligned=False):
    """aligned: every basis point is a grid point ((g-1) a multiple of (b-1))"""
    if case == "2d":
        nodes, el = G.disk(10.0, 6)
        return O.Mesh(nodes=nodes, elems=el), [17, 15], ([9, 8] if aligned else [9, 6])
    nodes, el = G.cylinder(8.0, 10.0, 4, 5)
    return O.Mesh(nodes=nodes, elems=el), [13, 11, 9], ([5, 6, 5] if aligned else [5, 6, 4])


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_linpix_maps_brute_force(case):
    """Grid2LinPixMatrix / LinPix2GridMatrix / SubsampleLinPixel against the
    brute-force tents: G rows are the masked tents normalised to sum 1, GI
    the unnormalised masked tents transposed, and the solution voxels are
    exactly the basis voxels whose tent holds a grid point inside the mesh."""
    m, gd, bd = _coarse_case(case)
    R = O.Raster(m, gd, bdim=bd)
    W = _tent_dense(R)
    Gd = _dense(*R.gmap(), (R.blen, R.glen))
    rs = W.sum(1)
    ref = np.divide(W, rs[:, None], out=np.zeros_like(W), where=rs[:, None] > 0)
    assert np.max(np.abs(Gd - ref)) <= 1e-14
    # the structures differ only where a grid point sits on a neighbouring
    # basis point (d = 1 exactly, a weight of rounding size either side)
    diff = (Gd != 0) != (ref != 0)
    assert np.all(np.maximum(np.abs(Gd), np.abs(ref))[diff] <= 1e-14)
    GI = _dense(*R.gimap(), (R.glen, R.blen))
    assert np.max(np.abs(GI - W.T)) <= 1e-14
    # support: SubsampleLinPixel of the inside mask without elref
    bs = R.subsample((R.elref >= 0).astype(float), masked=False)
    assert np.array_equal(R.sol2basis, np.nonzero(bs > 0)[0])
    assert np.array_equal(R.sol2basis, np.nonzero(np.diff(R.gmap()[0]) > 0)[0])
    assert set(np.nonzero(rs > 1e-12)[0]) <= set(R.sol2basis)
    # masked subsampling is G applied to the image
    img = np.random.default_rng(1).standard_normal(R.glen)
    assert np.allclose(R.subsample(img, masked=True), Gd @ img, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("case", ["2d", "3d"])
def test_linpix_maps_linear_reproduction(case):
    """GI interpolates a linear field given at the basis points exactly at
    every grid point inside the mesh (the tents of a cell sum to 1); G
    averages a linear grid field to its value at interior basis points whose
    whole tent lies inside the mesh (weights symmetric about the basis
    point, which holds when basis points are grid points)."""
    m, gd, bd = _coarse_case(case, aligned=True)
    R = O.Raster(m, gd, bdim=bd)
    c = np.array([0.7, -1.3, 0.4])[:R.dim]

    def lin(dims, npts):
        out = np.full(npts, 0.25)
        for d in range(R.dim):
            s = int(np.prod(dims[:d])) if d else 1
            out += c[d] * (np.arange(npts) // s % dims[d]) / (dims[d] - 1.0)
        return out

    fb, fg = lin(R.bdim, R.blen), lin(R.gdim, R.glen)
    GI = _dense(*R.gimap(), (R.glen, R.blen))
    inside = R.elref >= 0
    assert np.allclose(GI[inside].sum(1), 1.0, atol=1e-13)
    assert np.allclose(GI[inside] @ fb, fg[inside], atol=1e-13)
    assert np.all(GI[~inside] == 0)
    Gd = _dense(*R.gmap(), (R.blen, R.glen))
    W = _tent_dense(R, masked=False)
    full = np.array([np.all(inside[W[b] > 0]) for b in range(R.blen)])
    interior = np.ones(R.blen, bool)
    for d in range(R.dim):
        s = int(np.prod(R.bdim[:d])) if d else 1
        i = np.arange(R.blen) // s % R.bdim[d]
        interior &= (i > 0) & (i < R.bdim[d] - 1)
    sel = full & interior
    assert sel.sum() >= 2
    assert np.allclose(Gd[sel] @ fg, fb[sel], atol=1e-13)


def test_grid_jacobian_basis_map_identity():
    """Jacobians and the data gradient in a coarser basis against the
    (pinned) grid-basis ones: every column of the basis version is the G row
    of its voxel applied to the grid version, with the voxel size of the
    basis (prod gsize/(bdim-1)) in place of the grid's."""
    nodes, el = G.disk(25.0, 12)
    m = O.Mesh(nodes=nodes, elems=el)
    n = m.n
    Q, M = G.ring(3, 25.0), G.ring(3, 25.0, offset=0.5)
    lp = np.array([0, 2, 4, 6], np.int32)
    li = np.array([1, 2, 2, 0, 0, 1], np.int32)
    mua, mus, ref = np.full(n, 0.01), np.full(n, 1.0), np.full(n, 1.4)
    d, a, mv = _fields(m, Q, M, mua, mus, ref, 100.0)
    proj = O.project(d, mv.astype(complex), lp, li, logscale=False)
    Rg = O.Raster(m, [41, 37])
    Rb = O.Raster(m, [41, 37], bdim=[15, 12])
    assert Rb.slen < Rg.slen
    Gd = _dense(*Rb.gmap(), (Rb.blen, Rb.glen))
    # grid columns -> basis solution columns
    P = Gd[Rb.sol2basis][:, Rg.sol2basis]
    szg = np.prod((Rg.bbmax - Rg.bbmin) / (Rg.gdim - 1.0))
    szb = np.prod((Rb.bbmax - Rb.bbmin) / (Rb.bdim - 1.0))
    Jg = O.jacobian_grid(Rg, d, a, lp, li, proj, logscale=True)
    Jb = O.jacobian_grid(Rb, d, a, lp, li, proj, logscale=True)
    ng, nb = Rg.slen, Rb.slen
    for blk in range(2):
        ref_b = (Jg[:, blk * ng:(blk + 1) * ng] / szg) @ P.T * szb
        got = Jb[:, blk * nb:(blk + 1) * nb]
        assert np.max(np.abs(got - ref_b)) <= 1e-12 * np.abs(ref_b).max()
    # CW on the real parts of the fields
    dr, ar = d.real.copy(), a.real.copy()
    pr = O.project(dr.astype(complex), mv.astype(complex), lp, li, logscale=False).real
    Jgm, Jgk = O.jacobian_grid_cw(Rg, dr, ar, lp, li, pr, logscale=True)
    Jbm, Jbk = O.jacobian_grid_cw(Rb, dr, ar, lp, li, pr, logscale=True)
    for Jgx, Jbx in ((Jgm, Jbm), (Jgk, Jbk)):
        ref_b = (Jgx / szg) @ P.T * szb
        assert np.max(np.abs(Jbx - ref_b)) <= 1e-12 * np.abs(ref_b).max()
    # data gradient (no voxel size)
    gg = O.data_gradient(Rg, d, a[:3])
    gb = O.data_gradient(Rb, d, a[:3])
    for blk in range(2):
        ref_b = P @ gg[blk * ng:(blk + 1) * ng]
        assert np.max(np.abs(gb[blk * nb:(blk + 1) * nb] - ref_b)) <= 1e-12 * np.abs(ref_b).max()


# ---------------------------------------------------------------------------
# Closed forms of every grid-Jacobian entry on linear fields
#
# With nodal fields that are linear functions of position, the grid map B
# reproduces them exactly at the grid points inside the mesh (pinned above),
# and ImageGradient's central and one-sided differences are both exact up to
# the reference's step choice: grid spacing h = gsize/(n-1) over the step
# dx = gsize/n (jacobian.cc:961,1000,1039), i.e. a factor n/(n-1).  A point
# with no usable neighbour along an axis gets 0 there.  So every solution
# voxel's entries follow from the field coefficients alone:
#   mua:   -f(x) g(x) / proj * sz
#   kappa: -sum_d Df_d Dg_d / proj * sz,   Df_d = a_d n_d/(n_d-1) [neighbour]
# (PMDF_mua / PMDF_kap / PMDF_log, jacobian.cc:57-79; sz = prod gsize/(n-1),
# :368-372; rows link-ordered, Re in row k and Im in row k + nQM).

def _grid_points(R):
    gdim = R.gdim
    h = (R.bbmax - R.bbmin) / (gdim - 1)
    idx = np.arange(R.glen)
    coords, stride = [], 1
    for d in range(R.dim):
        i_d = (idx // stride) % gdim[d]
        coords.append(R.bbmin[d] + i_d * h[d])
        stride *= gdim[d]
    return np.stack(coords, axis=1), h


def _neighbour_flags(R):
    """per grid point and axis: (left exists and inside, right exists and
    inside, left exists, right exists)"""
    gdim, inside = R.gdim, R.elref >= 0
    idx = np.arange(R.glen)
    out, stride = [], 1
    for d in range(R.dim):
        i_d = (idx // stride) % gdim[d]
        has_l, has_r = i_d > 0, i_d < gdim[d] - 1
        il = np.where(has_l, idx - stride, idx)
        ir = np.where(has_r, idx + stride, idx)
        out.append((has_l & inside[il], has_r & inside[ir], has_l, has_r, il, ir))
        stride *= gdim[d]
    return out


def _linear_case(case):
    rng = np.random.default_rng(5)
    if case == "disk":
        nodes, el = G.disk(10.0, 6)
        gdim = [23, 19]
    else:
        nodes, el = G.cylinder(6.0, 8.0, 3, 4)
        gdim = [11, 9, 13]
    dim = nodes.shape[1]
    nq, nm = 3, 2
    A = rng.normal(size=(nq, dim + 1)) + 1j * rng.normal(size=(nq, dim + 1))
    Bc = rng.normal(size=(nm, dim + 1)) + 1j * rng.normal(size=(nm, dim + 1))
    lp = np.array([0, 2, 3, 5], np.int32)
    li = np.array([0, 1, 1, 0, 1], np.int32)
    proj = rng.normal(size=5) + 1j * rng.normal(size=5)
    return nodes, el, gdim, A, Bc, lp, li, proj


def _lin(C_, X):
    return C_[:, :1] + C_[:, 1:] @ X.T   # rows: one field per coefficient row


@pytest.mark.parametrize("case", ["disk", "cyl"])
def test_grid_jacobian_closed_form_on_linear_fields(case):
    nodes, el, gdim, A, Bc, lp, li, proj = _linear_case(case)
    m = O.Mesh(nodes=nodes, elems=el)
    R = O.Raster(m, gdim)
    dim = nodes.shape[1]
    dphi, aphi = _lin(A, nodes), _lin(Bc, nodes)
    J = O.jacobian_grid(R, dphi, aphi, lp, li, proj, logscale=True)
    X, h = _grid_points(R)
    sol = np.flatnonzero(R.elref >= 0)
    assert np.array_equal(sol, R.sol2basis)
    Xs = X[sol]
    r = np.asarray(gdim, float) / (np.asarray(gdim, float) - 1.0)
    has_nb = np.stack([(fl[0] | fl[1])[sol] for fl in _neighbour_flags(R)], axis=1)   # slen x dim
    sz = np.prod(h)
    nqm = len(li)
    exp = np.zeros_like(J)
    k = 0
    for q in range(len(lp) - 1):
        fq = _lin(A[q:q + 1], Xs)[0]
        Df = A[q, 1:][None, :] * r[None, :] * has_nb
        for t in range(lp[q], lp[q + 1]):
            mi = li[t]
            gm = _lin(Bc[mi:mi + 1], Xs)[0]
            Dg = Bc[mi, 1:][None, :] * r[None, :] * has_nb
            pm = -fq * gm / proj[k] * sz
            pk = -(Df * Dg).sum(1) / proj[k] * sz
            exp[k, :R.slen], exp[k + nqm, :R.slen] = pm.real, pm.imag
            exp[k, R.slen:], exp[k + nqm, R.slen:] = pk.real, pk.imag
            k += 1
    for blk in (slice(0, R.slen), slice(R.slen, 2 * R.slen)):
        scale = np.max(np.abs(exp[:, blk]))
        assert np.max(np.abs(J[:, blk] - exp[:, blk])) <= 1e-12 * scale
    # some voxels sit next to the boundary, with a one-sided difference or none
    assert np.any(~has_nb) or any(np.any(~(fl[0] & fl[1])[sol]) for fl in _neighbour_flags(R))
    # the checks fail a dropped n/(n-1), a lost sign or swapped Re/Im rows
    assert np.max(np.abs(exp[:, R.slen:])) > 0
    bad = exp.copy()
    bad[:, R.slen:] /= r[0] ** 2
    assert np.max(np.abs(J - bad)) > 1e-6 * np.max(np.abs(exp))


@pytest.mark.parametrize("case", ["disk", "cyl"])
def test_grid_jacobian_cw_closed_form_on_linear_fields(case):
    """CW (GenerateJacobian_cw_grid, jacobian.cc:610-721): the direct-field
    gradient is masked, the adjoint gradient is not (ImageGradient without
    a mask, :656), so next to the boundary it differences against the zero
    the grid map gives outside the mesh.  Closed form: g-tilde = g inside,
    0 outside; central where both grid neighbours exist, one-sided at the
    grid's own edges."""
    nodes, el, gdim, A, Bc, lp, li, proj = _linear_case(case)
    A, Bc, proj = A.real.copy(), Bc.real.copy(), np.abs(proj.real) + 0.5
    m = O.Mesh(nodes=nodes, elems=el)
    R = O.Raster(m, gdim)
    dphi, aphi = _lin(A, nodes), _lin(Bc, nodes)
    Jm, Jk = O.jacobian_grid_cw(R, dphi, aphi, lp, li, proj, logscale=True)
    X, h = _grid_points(R)
    inside = R.elref >= 0
    sol = np.flatnonzero(inside)
    r = np.asarray(gdim, float) / (np.asarray(gdim, float) - 1.0)
    flags = _neighbour_flags(R)
    dx = (R.bbmax - R.bbmin) / np.asarray(gdim, float)
    sz = np.prod(h)
    k = 0
    for q in range(len(lp) - 1):
        f = _lin(A[q:q + 1], X)[0]
        Df = np.stack([A[q, 1 + d] * r[d] * (flags[d][0] | flags[d][1]) for d in range(R.dim)], axis=1)[sol]
        for t in range(lp[q], lp[q + 1]):
            mi = li[t]
            gt = np.where(inside, _lin(Bc[mi:mi + 1], X)[0], 0.0)
            Dg = []
            for d in range(R.dim):
                _, _, hl, hr, il, ir = flags[d]
                c = np.where(hl & hr, (gt[ir] - gt[il]) / (2 * dx[d]),
                             np.where(hr, (gt[ir] - gt) / dx[d], np.where(hl, (gt - gt[il]) / dx[d], 0.0)))
                Dg.append(c)
            Dg = np.stack(Dg, axis=1)[sol]
            em = -f[sol] * gt[sol] / proj[k] * sz
            ek = -(Df * Dg).sum(1) / proj[k] * sz
            assert np.max(np.abs(Jm[k] - em)) <= 1e-12 * np.max(np.abs(em))
            assert np.max(np.abs(Jk[k] - ek)) <= 1e-12 * np.max(np.abs(ek))
            k += 1


def test_image_gradient_branches_on_quadratic_image():
    """Each branch of ImageGradient (jacobian.cc:962-985 and its y/z
    twins) checked by value: on f = sum_d c_d x_d^2 + l.x + e x y the
    central difference gives f'(x) h/dx, the forward one (f'(x) + c h) h/dx,
    the backward one (f'(x) - c h) h/dx, with h the grid spacing
    gsize/(n-1) and dx = gsize/n the reference's step.  The branch follows
    the mask: both neighbours inside -> central (even at a masked centre);
    else a masked centre -> 0; else the one inside neighbour; else 0."""
    gdim = np.array([7, 5, 4])
    gsize = np.array([6.0, 8.0, 3.0])
    h = gsize / (gdim - 1)
    dx = gsize / gdim
    c = np.array([0.7, -1.1, 0.4])
    lcoef = np.array([0.3, 0.2, -0.5])
    e = 0.6
    glen = int(np.prod(gdim))
    idx = np.arange(glen)
    I = np.stack([idx % 7, (idx // 7) % 5, idx // 35], axis=1)
    X = I * h
    im = (c * X ** 2).sum(1) + X @ lcoef + e * X[:, 0] * X[:, 1]
    rng = np.random.default_rng(9)
    mask = np.where(rng.random(glen) < 0.3, -1, 0).astype(np.int32)
    g = O.image_gradient(gdim, gsize, im + 0j, mask)
    fprime = [2 * c[0] * X[:, 0] + lcoef[0] + e * X[:, 1],
              2 * c[1] * X[:, 1] + lcoef[1] + e * X[:, 0],
              2 * c[2] * X[:, 2] + lcoef[2]]
    strides = [1, 7, 35]
    counts = {"central": 0, "forward": 0, "backward": 0, "zero": 0}
    for d in range(3):
        for p in range(glen):
            i = I[p, d]
            bl = i > 0 and mask[p - strides[d]] >= 0
            br = i < gdim[d] - 1 and mask[p + strides[d]] >= 0
            if bl and br:
                exp, kind = fprime[d][p] * h[d] / dx[d], "central"
            elif mask[p] < 0:
                exp, kind = 0.0, "zero"
            elif bl:
                exp, kind = (fprime[d][p] - c[d] * h[d]) * h[d] / dx[d], "backward"
            elif br:
                exp, kind = (fprime[d][p] + c[d] * h[d]) * h[d] / dx[d], "forward"
            else:
                exp, kind = 0.0, "zero"
            counts[kind] += 1
            assert abs(g[d][p].real - exp) <= 1e-12 * (1 + abs(exp)), (d, p, kind)
            assert g[d][p].imag == 0.0
    assert min(counts.values()) > 10, counts
