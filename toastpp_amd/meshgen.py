"""Synthetic meshes, parameter maps and source/detector layouts of the
benchmark workloads (host-side input generation, numpy).

  disk(R, K)                 concentric-ring Tri3 disk: 1+3K(K+1) nodes, 6K^2 triangles
  cylinder(R, H, K, layers)  the disk extruded; each prism split into 3 Tet4
  ellipsoid(a, b, c, N)      cube-to-ball mapped (N+1)^3 grid, 6 Tet4 per cell

All elements are positively oriented (the reference's signed element size
requires it: tri3.cc:184-189, tet4.cc:1011-1014).  Node numbering is
structured (ring/layer or lexicographic), which keeps the system matrix
banded: neighbour rows of the SpMM stay close in memory.
"""
import numpy as np


def disk(R=25.0, K=32):
    rings = [np.zeros((1, 2))]
    ring_ids = [np.array([0])]
    nid = 1
    for k in range(1, K + 1):
        m = 6 * k
        th = 2 * np.pi * np.arange(m) / m
        rings.append(np.c_[R * k / K * np.cos(th), R * k / K * np.sin(th)])
        ring_ids.append(np.arange(nid, nid + m))
        nid += m
    nodes = np.vstack(rings)
    tris = []
    for k in range(1, K + 1):
        inner, outer = ring_ids[k - 1], ring_ids[k]
        ni, no = len(inner), len(outer)
        if k == 1:
            for j in range(no):
                tris.append((inner[0], outer[j], outer[(j + 1) % no]))
            continue
        # zipper the two rings by angle
        ai = 2 * np.pi * np.arange(ni + 1) / ni
        ao = 2 * np.pi * np.arange(no + 1) / no
        i = j = 0
        while i < ni or j < no:
            if j < no and (i >= ni or ao[j + 1] <= ai[i + 1] + 1e-12):
                tris.append((inner[i % ni], outer[j], outer[(j + 1) % no]))
                j += 1
            else:
                tris.append((inner[i % ni], outer[j % no], inner[(i + 1) % ni]))
                i += 1
    el = np.array(tris, np.int32)
    return nodes, _orient(nodes, el)


def _orient(nodes, el):
    el = el.copy()
    if nodes.shape[1] == 2:
        p0, p1, p2 = nodes[el[:, 0]], nodes[el[:, 1]], nodes[el[:, 2]]
        a = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])
    else:
        p0, p1, p2, p3 = (nodes[el[:, k]] for k in range(4))
        a = np.einsum("ij,ij->i", np.cross(p1 - p0, p2 - p0), p3 - p0)
    neg = a < 0
    el[neg, 1], el[neg, 2] = el[neg, 2].copy(), el[neg, 1].copy()
    if np.any(a == 0):
        raise ValueError("degenerate element in generated mesh")
    return el


def cylinder(R=25.0, H=50.0, K=20, layers=24):
    n2, t2 = disk(R, K)
    nl = layers + 1
    z = np.linspace(-H / 2, H / 2, nl)
    nodes = np.vstack([np.c_[n2, np.full(len(n2), zz)] for zz in z])
    N2 = len(n2)
    tets = []
    tri = np.sort(t2, axis=1)   # a < b < c: diagonals from the larger bottom to the smaller top index
    for L in range(layers):
        a, b, c = tri[:, 0] + L * N2, tri[:, 1] + L * N2, tri[:, 2] + L * N2
        A, B, Cc = a + N2, b + N2, c + N2
        tets += [np.c_[a, b, c, A], np.c_[b, c, A, B], np.c_[c, A, B, Cc]]
    el = np.vstack(tets).astype(np.int32)
    return nodes, _orient(nodes, el)


def ellipsoid(a=80.0, b=65.0, c=90.0, N=62):
    t = np.linspace(-1.0, 1.0, N + 1)
    X, Y, Z = np.meshgrid(t, t, t, indexing="ij")
    x, y, z = X.ravel(), Y.ravel(), Z.ravel()
    # cube -> ball (volume-preserving-ish analytic map), then scale to the ellipsoid
    xs = x * np.sqrt(1 - y * y / 2 - z * z / 2 + y * y * z * z / 3)
    ys = y * np.sqrt(1 - z * z / 2 - x * x / 2 + z * z * x * x / 3)
    zs = z * np.sqrt(1 - x * x / 2 - y * y / 2 + x * x * y * y / 3)
    nodes = np.c_[a * xs, b * ys, c * zs]
    idx = np.arange((N + 1) ** 3).reshape(N + 1, N + 1, N + 1)
    v = [idx[i:N + i, j:N + j, k:N + k].ravel() for i in (0, 1) for j in (0, 1) for k in (0, 1)]
    # v[ijk] with bits (i,j,k); Kuhn split of each cube into 6 tetrahedra along the 000-111 diagonal
    v000, v001, v010, v011, v100, v101, v110, v111 = v
    paths = [(v100, v110), (v100, v101), (v010, v110), (v010, v011), (v001, v101), (v001, v011)]
    el = np.vstack([np.c_[v000, p1, p2, v111] for p1, p2 in paths]).astype(np.int32)
    return nodes, _orient(nodes, el)


def blobs(nodes, base, peak, sigma, nblob, rng, rmax=15.0):
    """base + sum of Gaussian blobs with centres uniform in |r| <= rmax."""
    dim = nodes.shape[1]
    out = np.full(len(nodes), base, np.float64)
    for _ in range(nblob):
        while True:
            cpos = rng.uniform(-rmax, rmax, dim)
            if np.linalg.norm(cpos) <= rmax:
                break
        d2 = np.sum((nodes - cpos) ** 2, axis=1)
        out += peak * np.exp(-0.5 * d2 / sigma ** 2)
    return out


def params(nodes, seed):
    rng = np.random.default_rng(seed)
    mua = blobs(nodes, 0.025, 0.025, 3.0, 2, rng)
    mus = blobs(nodes, 2.0, 2.0, 3.0, 1, rng)
    ref = np.full(len(nodes), 1.4)
    return mua, mus, ref


def ring(nq, r, z=None, offset=0.0):
    th = 2 * np.pi * (np.arange(nq) + offset) / nq
    if z is None:
        return np.c_[r * np.cos(th), r * np.sin(th)]
    return np.c_[r * np.cos(th), r * np.sin(th), np.full(nq, z)]


def workload(name):
    """(nodes, elements, Q, M, mua, mus, ref, freq_MHz, description) of a named benchmark workload."""
    if name == "disk2d":   # plain method, small: 2D disk R=25 mm, K=32, CW, 16 Q / 16 M
        nodes, el = disk(25.0, 32)
        Q, M = ring(16, 25.0), ring(16, 25.0, offset=0.5)
        mua, mus, ref = params(nodes, 0)
        return nodes, el, Q, M, mua, mus, ref, 0.0, "2D disk R25 K32 Tri3, CW, 16Qx16M"
    if name == "cyl3d":    # 3D cylinder R25 H50, ~31.5k nodes, 100 MHz, 32 Q / 32 M in two rings
        nodes, el = cylinder(25.0, 50.0, 20, 24)
        Q = np.vstack([ring(16, 25.0, z) for z in (-12.5, 12.5)])
        M = np.vstack([ring(16, 25.0, z, 0.5) for z in (-12.5, 12.5)])
        mua, mus, ref = params(nodes, 1)
        return nodes, el, Q, M, mua, mus, ref, 100.0, "3D cylinder R25 H50 Tet4, 100 MHz, 32Qx32M"
    if name == "ellipsoid3d":   # 80x65x90 mm ellipsoid, ~250k nodes, 64 sources, 100 MHz
        nodes, el = ellipsoid(80.0, 65.0, 90.0, 62)
        Q = _ellipsoid_ring_points(80.0, 65.0, 90.0, [-45.0, -15.0, 15.0, 45.0], 16, 0.0)
        M = _ellipsoid_ring_points(80.0, 65.0, 90.0, [-45.0, -15.0, 15.0, 45.0], 16, 0.5)
        mua, mus, ref = params(nodes, 2)
        return nodes, el, Q, M, mua, mus, ref, 100.0, "3D ellipsoid 80x65x90 Tet4, 100 MHz, 64Qx64M"
    if name == "cyl3d_jac":   # Jacobian workload: cylinder ~103k nodes, 32x32
        nodes, el = cylinder(25.0, 50.0, 30, 36)
        Q = np.vstack([ring(16, 25.0, z) for z in (-12.5, 12.5)])
        M = np.vstack([ring(16, 25.0, z, 0.5) for z in (-12.5, 12.5)])
        mua, mus, ref = params(nodes, 3)
        return nodes, el, Q, M, mua, mus, ref, 100.0, "3D cylinder R25 H50 K30x36 Tet4, 100 MHz, 32Qx32M"
    raise KeyError(name)


def serving_pool():
    """Meshes of the serving workload: 2D disks of 3k-20k nodes and 3D
    cylinders / ellipsoids of 30k-250k nodes, each with 64 boundary sources
    (name, nodes, elements, sources, mua, mus, ref)."""
    out = []
    for name, K in (("disk3k", 32), ("disk19k", 80)):
        nodes, el = disk(25.0, K)
        out.append((name, nodes, el, ring(64, 25.0)))
    nodes, el = cylinder(25.0, 50.0, 20, 24)
    out.append(("cyl31k", nodes, el, np.vstack([ring(32, 25.0, z) for z in (-12.5, 12.5)])))
    for name, N in (("ellipsoid80k", 42), ("ellipsoid250k", 62)):
        nodes, el = ellipsoid(80.0, 65.0, 90.0, N)
        out.append((name, nodes, el, _ellipsoid_ring_points(80.0, 65.0, 90.0, [-45.0, -15.0, 15.0, 45.0], 16, 0.0)))
    nodes, el = cylinder(25.0, 50.0, 30, 36)
    out.append(("cyl103k", nodes, el, np.vstack([ring(32, 25.0, z) for z in (-12.5, 12.5)])))
    res = []
    for i, (name, nodes, el, Q) in enumerate(out):
        mua, mus, ref = params(nodes, 10 + i)
        res.append((name, nodes, el, Q, mua, mus, ref))
    return res


def ellipsoid_batch():
    """Three ellipsoids (20k, 80k, 250k nodes) of the multi-mesh batch, 64 sources each."""
    res = []
    for name, N in (("ellipsoid20k", 26), ("ellipsoid80k", 42), ("ellipsoid250k", 62)):
        nodes, el = ellipsoid(80.0, 65.0, 90.0, N)
        Q = _ellipsoid_ring_points(80.0, 65.0, 90.0, [-45.0, -15.0, 15.0, 45.0], 16, 0.0)
        mua, mus, ref = params(nodes, 2)
        res.append((name, nodes, el, Q, mua, mus, ref))
    return res


def _ellipsoid_ring_points(a, b, c, zs, n, offset):
    pts = []
    for z in zs:
        s = np.sqrt(max(0.0, 1 - (z / c) ** 2))
        th = 2 * np.pi * (np.arange(n) + offset) / n
        pts.append(np.c_[a * s * np.cos(th), b * s * np.sin(th), np.full(n, z)])
    return np.vstack(pts)
