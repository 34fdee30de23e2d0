"""HIP-graph replay of the block Krylov iteration: each check period (8
iterations and the reliable update when it falls due) is captured once per
workspace and operator and replayed as one graph launch.  The captured
kernels, their arguments and their order are those of the direct launches,
so the fields and the per-column records are identical with graphs on and
off -- in the mixed and FP64 modes, for BiCGSTAB and PCG, across column
compaction -- and repeated solves replay the graphs already captured.
Through the C-ABI on the GPU (3D golden mesh of the reference's suite)."""
import numpy as np
import pytest

from tests.cases import CASES

pytestmark = pytest.mark.gpu

T = pytest.importorskip("toastpp_amd")
from toastpp_amd import _lib as L  # noqa: E402

CASE = CASES["3d_cplx_direct"]


@pytest.fixture(scope="module")
def cyl2():
    gm = T.Mesh(CASE["mesh"])
    gm.ReadQM(CASE["qm"])
    n = gm.NodeCount()
    rng = np.random.default_rng(31)
    mua, mus, ref = 0.01 + 0.01 * rng.random(n), 1.0 + rng.random(n), np.full(n, 1.4)
    B = np.ascontiguousarray(gm.Qvec(width=2.0).toarray().astype(np.complex128))[:, :48]
    return gm, T.Solution(mua, mus, ref), B


def _solve(gm, sol, B, solver, mode, freq, graphs):
    cplx = freq != 0.0
    f = T.FwdSolver(gm, solver, 1e-10, complex=cplx)
    f.SetMode(mode)
    f.SetPrecon("DIAG")
    f.SetGraphs(graphs)
    f.Allocate()
    f.Reset(sol, freq * 2 * np.pi * 1e-6)
    Bc = np.ascontiguousarray(B if cplx else B.real)
    dB = L.DeviceArray.from_host(Bc)
    dX = L.DeviceArray(Bc.shape, Bc.dtype)
    st1 = f.CalcFieldsDevice(Bc.shape[1], dB.ptr, dX.ptr)
    c1 = f.LastCounts()
    X1 = dX.to_host()
    st2 = f.CalcFieldsDevice(Bc.shape[1], dB.ptr, dX.ptr)   # replays the captured graphs
    c2 = f.LastCounts()
    assert np.array_equal(dX.to_host(), X1) and st1 == st2
    return X1, st1, c1, c2


@pytest.mark.parametrize("solver,mode,freq", [("BICGSTAB", "mixed", 100.0), ("BICGSTAB", "fp64_block", 100.0),
                                              ("CG", "mixed", 0.0), ("CG", "fp64_block", 0.0)])
def test_graph_replay_identical_to_direct_launches(cyl2, solver, mode, freq):
    gm, sol, B = cyl2
    Xg, sg, cg1, cg2 = _solve(gm, sol, B, solver, mode, freq, True)
    Xd, sd, cd1, cd2 = _solve(gm, sol, B, solver, mode, freq, False)
    assert np.array_equal(Xg, Xd)
    assert sg == sd and all(s[3] for s in sg)
    launches_g, graphs_g, syncs_g = cg2
    launches_d, graphs_d, syncs_d = cd2
    assert graphs_d == 0 and graphs_g > 0
    # every period's 64-104 kernel launches became one graph launch
    assert launches_g + 8 * graphs_g <= launches_d
    assert syncs_g == syncs_d
