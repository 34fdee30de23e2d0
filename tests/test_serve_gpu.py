"""Serving loop (tmi_server_*): every result is the same alone and in the
mix; the HBM budget bounds what runs at once; failures come back per
request.  Through the C-ABI on the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TMI = pytest.importorskip("toastpp_amd")
from toastpp_amd import _lib as L  # noqa: E402
from toastpp_amd import meshgen as G  # noqa: E402
from toastpp_amd.serve import Server  # noqa: E402


def _prepared(kind, freq):
    if kind == "disk":
        nodes, el = G.disk(25.0, 24)
        Q = G.ring(32, 25.0)
    else:
        nodes, el = G.cylinder(25.0, 50.0, 10, 12)
        Q = np.vstack([G.ring(16, 25.0, z) for z in (-12.5, 12.5)])
    mesh = TMI.Mesh(data=(nodes, el))
    mesh.SetQM(Q, Q[:1])
    mua, mus, ref = G.params(nodes, 1)
    cplx = freq != 0
    f = TMI.FwdSolver(mesh, "BICGSTAB" if cplx else "CG", 1e-10, complex=cplx)
    f.SetPrecon("DIAG")
    f.Allocate()
    f.Reset(TMI.Solution(mua, mus, ref), freq * 2 * np.pi * 1e-6)
    B = mesh.Qvec(width=2.0).toarray()
    B = np.ascontiguousarray(B if cplx else B.real)
    return mesh, f, L.DeviceArray.from_host(B), B.shape


@pytest.fixture(scope="module")
def pool():
    return [_prepared("disk", 100.0), _prepared("cyl", 0.0), _prepared("cyl", 200.0), _prepared("disk", 0.0)]


def _requests(pool, seed):
    rng = np.random.default_rng(seed)
    return [(int(rng.integers(len(pool))), int(rng.integers(4, 33))) for _ in range(10)]


def _run(pool, reqs, slots, concurrent, budget=64.0):
    srv = Server(slots=slots, hbm_budget_gib=budget)
    outs, ids = [], []
    for (m, nq) in reqs:
        mesh, f, dB, shape = pool[m]
        dX = L.DeviceArray((shape[0], shape[1]), np.complex128 if f.complex else np.float64)
        L.call("tmi_memcpy_h2d", dX._p, np.zeros(dX.shape, dX.dtype).ctypes.data_as(L.C.c_void_p), dX.nbytes)
        # the first nq columns of the source block, in place (row stride = all sources)
        rid = srv.submit(f, nq, dB.ptr, dX.ptr, shape[1])
        outs.append((dX, nq))
        ids.append(rid)
        if not concurrent:
            srv.wait(rid)
    res = []
    for rid, (dX, nq) in zip(ids, outs):
        if concurrent:
            st, times, status = srv.wait(rid)
        res.append(dX.to_host()[:, :nq])
    srv.close()
    return res


def test_results_identical_alone_and_in_the_mix(pool):
    reqs = _requests(pool, 3)
    alone = _run(pool, reqs, slots=1, concurrent=False)
    mixed = _run(pool, reqs, slots=4, concurrent=True)
    for a, b in zip(alone, mixed):
        assert np.array_equal(a, b)   # bit-identical: fixed-order reductions, no cross-request state


def test_hbm_budget_serializes(pool):
    mesh, f, dB, shape = pool[2]
    need = f.WorkBytes(16)
    srv = Server(slots=4, hbm_budget_gib=1.5 * need / 2.0 ** 30)   # room for one request at a time
    bufs, ids = [], []
    for _ in range(4):
        dX = L.DeviceArray(shape, np.complex128)
        L.call("tmi_memcpy_h2d", dX._p, np.zeros(shape, np.complex128).ctypes.data_as(L.C.c_void_p), dX.nbytes)
        bufs.append(dX)
        ids.append(srv.submit(f, 16, dB.ptr, dX.ptr, shape[1]))
    spans = [srv.wait(r)[1] for r in ids]
    srv.close()
    spans.sort(key=lambda t: t[1])
    for a, b in zip(spans, spans[1:]):
        assert b[1] >= a[2] - 1e-6, "requests overlapped despite the HBM budget"


def test_request_larger_than_budget_is_refused(pool):
    mesh, f, dB, shape = pool[0]
    srv = Server(slots=2, hbm_budget_gib=1e-6)
    with pytest.raises(L.TmiError):
        srv.submit(f, 8, dB.ptr, dB.ptr, shape[1])
    srv.close()


def test_retained_workspaces_counted_within_budget(pool):
    """Slots keep their workspaces between requests; the server counts them
    (held) beside the running requests' bounds (reserved) and stays within
    the budget: with room for one 32-column workspace only, the second
    large request makes the idle slot give its workspace back first.  The
    results equal those of a server with room for everything."""
    mesh, f, dB, shape = pool[2]
    big = f.WorkBytes(32)
    budget = 1.05 * big
    srv = Server(slots=2, hbm_budget_gib=budget / 2.0 ** 30)
    outs = []
    peak = 0.0
    for nq in (32, 32, 4, 32):
        dX = L.DeviceArray(shape, np.complex128)
        L.call("tmi_memcpy_h2d", dX._p, np.zeros(shape, np.complex128).ctypes.data_as(L.C.c_void_p), dX.nbytes)
        rid = srv.submit(f, nq, dB.ptr, dX.ptr, shape[1])
        while True:
            running, queued, held = srv.counts()
            peak = max(peak, held)
            try:
                st, _, status = srv.wait(rid, timeout=0.002)
                break
            except L.TmiError as e:
                if e.code != L.TMI_ERR_TIMEOUT:
                    raise
        assert status == L.TMI_OK and all(s[3] for s in st)
        peak = max(peak, srv.counts()[2])
        outs.append((dX.to_host()[:, :nq], nq))
    srv.close()
    assert peak <= budget, "counted device bytes %.0f above the budget %.0f" % (peak, budget)
    ref = _run(pool, [(2, nq) for _, nq in outs], slots=1, concurrent=False, budget=64.0)
    for (a, _), b in zip(outs, ref):
        assert np.array_equal(a, b)
