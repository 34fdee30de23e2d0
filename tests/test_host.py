"""Host-side checks that need no GPU: the C-ABI library loads and exports
every entry point the header declares, the host mesh reader agrees with the
oracle's, the workload generators produce valid meshes, and bench.py refuses
multi-GPU runs."""
import os
import re
import subprocess
import sys

import numpy as np
import pytest

from tests.conftest import ROOT

HEADER = os.path.join(ROOT, "include", "toastpp_amd.h")


def declared_symbols():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"\b(tmi_[a-z0-9_]+)\s*\(", txt)))


def test_header_symbols_exported():
    from toastpp_amd import _lib
    L = _lib.lib()
    names = declared_symbols()
    assert len(names) > 40
    missing = [s for s in names if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding covers every declared entry point
    assert set(names) <= set(_lib.exported_symbols())


def test_host_mesh_reader_matches_oracle():
    import ctypes as C
    from oracle import oracle as O
    from toastpp_amd import _lib
    for sub, name in (("ref2d", "circle25_32.msh"), ("ref3d", "cyl2.msh")):
        path = os.path.join(ROOT, "tests", "golden", sub, name).encode()
        d, n, e, nn = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _lib.call("tmi_read_mesh_file", path, C.byref(d), C.byref(n), C.byref(e), C.byref(nn), None, None, None)
        x = np.zeros((n.value, d.value))
        el = np.zeros((e.value, nn.value), np.int32)
        reg = np.zeros(n.value, np.int32)
        _lib.call("tmi_read_mesh_file", path, C.byref(d), C.byref(n), C.byref(e), C.byref(nn), _lib.ptr(x),
                  _lib.ptr(el), _lib.ptr(reg))
        om = O.Mesh(path.decode())
        assert np.array_equal(x, om.nodes())
        assert np.array_equal(el, om.elements())
        assert np.array_equal(reg, om.regions())


def test_host_mesh_reader_errors(tmp_path):
    import ctypes as C
    from toastpp_amd import _lib
    bad = tmp_path / "bad.msh"
    bad.write_text("not a mesh\n")
    d = C.c_int()
    with pytest.raises(_lib.TmiError) as ei:
        _lib.call("tmi_read_mesh_file", str(bad).encode(), C.byref(d), None, None, None, None, None, None)
    assert ei.value.code == _lib.TMI_ERR_IO
    tri6 = tmp_path / "tri6.msh"
    tri6.write_text("MeshData 5.0\n\nNodeList 3 1\nN[0 0]\nN[1 0]\nN[0 1]\n\nElementList 1\np 1 2 3 1 2 3\n")
    with pytest.raises(_lib.TmiError) as ei:
        _lib.call("tmi_read_mesh_file", str(tri6).encode(), C.byref(d), None, None, None, None, None, None)
    assert ei.value.code == _lib.TMI_ERR_UNSUPPORTED


def test_no_gpu_call_fails_cleanly():
    """Without a GPU, device calls return an error code instead of exiting."""
    import toastpp_amd as T
    from toastpp_amd import _lib
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("GPU present")
    except ImportError:
        pass
    with pytest.raises(_lib.TmiError) as ei:
        T.Mesh(os.path.join(ROOT, "tests", "golden", "ref2d", "circle25_32.msh"))
    assert ei.value.code == _lib.TMI_ERR_DEVICE


@pytest.mark.parametrize("name,nodes,elems", [("disk2d", 3169, 6144), ("cyl3d", 31525, 172800),
                                              ("ellipsoid3d", 250047, 1429968), ("cyl3d_jac", 103267, 583200)])
def test_workloads(name, nodes, elems):
    from toastpp_amd import meshgen as G
    x, el, Q, M, mua, mus, ref, f, desc = G.workload(name)
    assert x.shape[0] == nodes and el.shape[0] == elems
    if x.shape[1] == 2:
        p0, p1, p2 = x[el[:, 0]], x[el[:, 1]], x[el[:, 2]]
        vol = (p1[:, 0] - p0[:, 0]) * (p2[:, 1] - p0[:, 1]) - (p2[:, 0] - p0[:, 0]) * (p1[:, 1] - p0[:, 1])
    else:
        p0, p1, p2, p3 = (x[el[:, k]] for k in range(4))
        vol = np.einsum("ij,ij->i", np.cross(p1 - p0, p2 - p0), p3 - p0)
    assert np.all(vol > 0)
    assert np.all(mua >= 0.025) and np.all(mus >= 2.0) and np.all(ref == 1.4)


def test_bench_refuses_multi_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_raster_and_data_file_entry_points_reject_bad_arguments():
    """Argument checks of the raster and data-file entry points run before
    any GPU work and come back as status codes with a message."""
    import ctypes as C
    from toastpp_amd import _lib
    L = _lib.lib()
    out = C.c_void_p()
    g = np.array([8, 8], np.int32)
    assert L.tmi_raster_create_basis(None, g, g, None, C.byref(out)) == _lib.TMI_ERR_INVALID_ARG
    assert b"NULL" in L.tmi_last_error()
    n = C.c_int()
    assert L.tmi_raster_basis_len(None, C.byref(n), None) == _lib.TMI_ERR_INVALID_ARG
    assert L.tmi_read_nim(b"/nonexistent/x.nim", 0, C.byref(n), None, None) == _lib.TMI_ERR_IO
    assert L.tmi_read_vector(b"/nonexistent/x.fem", C.byref(n), None) == _lib.TMI_ERR_IO
    v = np.zeros(3)
    assert L.tmi_write_nim(b"/tmp/x.nim", b"m", 3, 1, v, 7) == _lib.TMI_ERR_INVALID_ARG


def test_optical_to_coeffs_index_matched_and_nonfinite():
    """A_Keijzer(1) = 1 (param.cc:24-32: theta_c = pi/2, R0 = 0), so n = 1
    gives c2a = c0 / 2; n < 1 and non-finite parameters are refused."""
    import toastpp_amd as T
    from toastpp_amd import _lib
    s = T.Solution(np.full(3, 0.01), np.full(3, 1.0), np.ones(3))
    assert np.allclose(s.c2a, 0.3 / 2.0, rtol=1e-15, atol=0)
    assert np.allclose(s.cmua, 0.3 * 0.01, rtol=1e-15, atol=0)
    assert np.allclose(s.ckappa, 0.3 / (3.0 * 1.01), rtol=1e-15, atol=0)
    for mua, mus, ref in ((0.01, 1.0, 0.9), (np.nan, 1.0, 1.4), (0.01, np.inf, 1.4), (0.01, 1.0, np.nan)):
        with pytest.raises(_lib.TmiError) as ei:
            T.Solution(np.full(2, mua), np.full(2, mus), np.full(2, ref))
        assert ei.value.code == _lib.TMI_ERR_INVALID_ARG


class _FakeMesh:
    """Stands in for a device mesh: only the sizes the host checks read."""

    def __init__(self, n, nq, nm):
        self.n, self.dims = n, (nq, nm, nq * nm)

    def NodeCount(self):
        return self.n

    def qm_dims(self):
        return self.dims


def test_fwdsolver_rejects_mismatched_shapes_on_the_host():
    """CalcFields / ProjectAll check qvec, mvec and phi against the mesh
    before any C call (the C entry points copy nq x n entries)."""
    from toastpp_amd.fwdsolver import FwdSolver
    f = object.__new__(FwdSolver)   # no device handle: the checks run first
    f.mesh, f.complex, f._h = _FakeMesh(10, 2, 3), True, None
    with pytest.raises(ValueError):
        f.CalcFields(np.ones((2, 9)))                       # qvec narrower than the mesh
    with pytest.raises(ValueError):
        f.CalcFields(np.ones((2, 10)), phi=np.zeros((2, 9)))   # short phi
    with pytest.raises(ValueError):
        f.ProjectAll(np.ones((3, 10)), np.zeros((1, 10)))      # phi of too few sources
    with pytest.raises(ValueError):
        f.ProjectAll(np.ones((2, 10)), np.zeros((2, 10)))      # mvec of too few detectors
    with pytest.raises(ValueError):
        f.ProjectAllDevice(np.ones((3, 10)), 0, ld=1)          # row stride below the source count


def _node_order(kind, x, rp, ci):
    from toastpp_amd import _lib
    n = len(rp) - 1
    perm = np.full(n, -1, np.int32)
    _lib.call("tmi_node_order", kind, x.shape[1], n, np.ascontiguousarray(x, np.float64),
              np.ascontiguousarray(rp, np.int32), np.ascontiguousarray(ci, np.int32), perm)
    return perm


def _block_columns(perm, rp, ci, rows=8):
    """Distinct columns of consecutive `rows`-row blocks of the relabelled
    system, summed over the blocks (the row-block SpMM's gathers)."""
    tot = 0
    for b in range(0, len(perm), rows):
        tot += len(np.unique(np.concatenate([ci[rp[v]:rp[v + 1]] for v in perm[b:b + rows]])))
    return tot


def test_node_orders_are_permutations_and_clustering_shares_columns():
    """Host node orders (no GPU): each kind is a permutation; natural is the
    identity; Morton sorts the nodes by the interleaved bits of their
    quantised coordinates; the clustered order puts rows that share columns
    into the same 8-row blocks: on the reference's 3D golden mesh its blocks
    hold fewer distinct columns than blocks of 8 consecutive Morton rows."""
    from oracle import oracle as O
    from toastpp_amd import _lib
    om = O.Mesh(os.path.join(ROOT, "tests", "golden", "ref3d", "cyl2.msh"))
    rp, ci = om.csr()
    x = om.nodes()
    n = len(x)
    perms = {k: _node_order(k, x, rp, ci) for k in (_lib.ORDER_NATURAL, _lib.ORDER_MORTON, _lib.ORDER_CLUSTER)}
    for p in perms.values():
        assert np.array_equal(np.sort(p), np.arange(n))
    assert np.array_equal(perms[_lib.ORDER_NATURAL], np.arange(n))
    # Morton: the same sort done here with 21-bit quantisation over the
    # largest extent (cubic cells)
    lo, ext = x.min(0), (x.max(0) - x.min(0)).max()
    q = np.clip((x - lo) * (2097151.0 / ext), 0, 2097151).astype(np.uint64)
    key = np.zeros(n, np.uint64)
    for bit in range(21):
        for d in range(3):
            key |= ((q[:, d] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(3 * bit + d)
    assert np.array_equal(perms[_lib.ORDER_MORTON], np.argsort(key, kind="stable"))
    morton = _block_columns(perms[_lib.ORDER_MORTON], rp, ci)
    clustered = _block_columns(perms[_lib.ORDER_CLUSTER], rp, ci)
    assert clustered < 0.95 * morton, (clustered, morton)   # 54,194 against 59,799 distinct block columns


def test_cluster_order_matching_rule():
    """The pairing rule on a hand-built case: rows 0..7 on a line (x = row),
    each row's columns its two neighbours and itself; row 2 also holds column
    7 and row 7 column 2.  Visiting in x order, row 0 pairs with row 1 (they
    share columns 0, 1); row 2 shares 2 columns with row 3 ({2, 3}) and 2 with
    row 7 ({2, 7}): the tie goes to the group visited first, row 3; rows 4 and
    5, then 6 and 7.  Pairs of pairs: (0,1) with (2,3), (4,5) with (6,7)."""
    cols = [[0, 1], [0, 1, 2], [1, 2, 3, 7], [2, 3, 4], [3, 4, 5], [4, 5, 6], [5, 6, 7], [2, 6, 7]]
    rp = np.cumsum([0] + [len(c) for c in cols]).astype(np.int32)
    ci = np.concatenate(cols).astype(np.int32)
    x = np.c_[np.arange(8.0), np.zeros(8), np.zeros(8)]
    from toastpp_amd import _lib
    perm = _node_order(_lib.ORDER_CLUSTER, x, rp, ci)
    assert perm.tolist() == [0, 1, 2, 3, 4, 5, 6, 7]
    # a group that cannot fill up goes last: 9 rows, the ninth isolated
    cols9 = cols + [[8]]
    rp9 = np.cumsum([0] + [len(c) for c in cols9]).astype(np.int32)
    ci9 = np.concatenate(cols9).astype(np.int32)
    x9 = np.c_[np.r_[np.arange(8.0), -1.0], np.zeros(9), np.zeros(9)]
    perm9 = _node_order(_lib.ORDER_CLUSTER, x9, rp9, ci9)
    assert perm9[-1] == 8 and sorted(perm9[:8].tolist()) == list(range(8))


def test_node_order_rejects_bad_input():
    from toastpp_amd import _lib
    x = np.zeros((2, 3))
    rp = np.array([0, 1, 2], np.int32)
    with pytest.raises(_lib.TmiError) as ei:
        _node_order(7, x, rp, np.array([0, 1], np.int32))
    assert ei.value.code == _lib.TMI_ERR_INVALID_ARG
    with pytest.raises(_lib.TmiError) as ei:
        _node_order(_lib.ORDER_CLUSTER, x, rp, np.array([0, 5], np.int32))
    assert ei.value.code == _lib.TMI_ERR_INVALID_ARG
