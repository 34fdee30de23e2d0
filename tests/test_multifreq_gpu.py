"""Several frequencies of one mesh in one block solve
(tmi_solver_calc_fields_freqs_device; no reference counterpart -- the
reference assembles and solves each frequency on its own,
fwdsolver.cc:405-424, :653-810).  Column q is solved at omega_q with the
optical coefficients of the last Reset, through the C-ABI on the GPU, and
held to: the oracle's FP64 BiCGSTAB fields of the system assembled at that
frequency, and the same solver's single-frequency solves after Reset at
each omega.  Frequencies {0, 50, 100, 200} MHz as the serving workload
draws them; 64-column tiles (row-block and per-row SpMMs) and the narrow
tiles of compaction; mixed and FP64 block modes, with and without the
diagonal preconditioner; graph replay identical to direct launches."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import CASES

pytestmark = pytest.mark.gpu

T = pytest.importorskip("toastpp_amd")
from toastpp_amd import _lib as L  # noqa: E402

FREQS = np.array([0.0, 50.0, 100.0, 200.0])   # MHz
W = FREQS * 2 * np.pi * 1e-6                   # the unit of Reset's omega


@pytest.fixture(scope="module")
def cyl2():
    case = CASES["3d_cplx_direct"]
    gm = T.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    n = gm.NodeCount()
    rng = np.random.default_rng(12)
    mua, mus, ref = 0.01 + 0.01 * rng.random(n), 1.0 + rng.random(n), np.full(n, 1.4)
    Qd = np.ascontiguousarray(gm.Qvec(width=2.0).toarray().astype(np.complex128))   # [n][80]
    # 48 columns: 12 sources at each of the 4 frequencies, interleaved
    src = np.arange(48) % 12 * 6
    om_col = W[np.arange(48) // 12]
    B = np.ascontiguousarray(Qd[:, src])
    om = O.Mesh(case["mesh"])
    return case, gm, om, (mua, mus, ref), B, src, om_col


def _solver(gm, sol, mode, precon, variant=None, graphs=True, tol=1e-12):
    f = T.FwdSolver(gm, "COCG", tol, complex=True)
    f.SetMode(mode)
    if precon:
        f.SetPrecon(precon)
    if variant:
        f.SetSpmmVariant(variant)
    f.SetGraphs(graphs)
    f.Allocate()
    f.Reset(T.Solution(*sol), 0.7)   # any omega: the multi-frequency solve takes each column's own
    return f


def _freq_solve(f, B, om_col):
    dB = L.DeviceArray.from_host(B)
    dX = L.DeviceArray(B.shape, np.complex128)
    st = f.CalcFieldsFreqsDevice(B.shape[1], dB.ptr, dX.ptr, om_col)
    return dX.to_host(), st


@pytest.mark.parametrize("mode", ["mixed", "fp64_block"])
@pytest.mark.parametrize("precon", [None, "DIAG"])
def test_multifreq_matches_oracle_and_single_frequency_solves(cyl2, mode, precon):
    case, gm, om, sol, B, src, om_col = cyl2
    print("multifreq mode=%s precon=%s" % (mode, precon), flush=True)
    f = _solver(gm, sol, mode, precon)
    X, st = _freq_solve(f, B, om_col)
    assert all(s[3] for s in st) and max(s[2] for s in st) <= 1e-12
    cm, ck, c2 = O.coeffs(*sol)
    for fi, w in enumerate(W):
        cols = np.flatnonzero(om_col == w)
        # the same solver type solving this frequency alone (Reset at w)
        g = _solver(gm, sol, mode, precon)
        g.Reset(T.Solution(*sol), w)
        dB = L.DeviceArray.from_host(np.ascontiguousarray(B[:, cols]))
        dX = L.DeviceArray((B.shape[0], len(cols)), np.complex128)
        g.CalcFieldsDevice(len(cols), dB.ptr, dX.ptr)
        Xs = dX.to_host()
        rel = np.linalg.norm(X[:, cols] - Xs, axis=0) / np.linalg.norm(Xs, axis=0)
        assert rel.max() <= 1e-9, (FREQS[fi], rel.max())
        # the oracle at this frequency, two columns
        rp, ci, val = om.sysmat(cm, ck, c2, w)
        for c in cols[:2]:
            x, it, err = O.bicgstab(rp, ci, val, B[:, c].copy(), 1e-15, 20000, True)
            r = np.linalg.norm(X[:, c] - x) / np.linalg.norm(x)
            assert r <= 1e-8, (FREQS[fi], c, r)


@pytest.mark.parametrize("variant", ["rowblock", "rowblock_mfma", "direct", "lds"])
def test_multifreq_64_column_tiles(cyl2, variant):
    """64 columns in one tile (16 sources x 4 frequencies): the row-block
    SpMM or the per-row 64-column kernel with per-column frequencies, then
    the narrow tiles of compaction as columns finish."""
    case, gm, om, sol, B, src, om_col = cyl2
    Qd = np.ascontiguousarray(gm.Qvec(width=2.0).toarray().astype(np.complex128))
    s64 = np.arange(64) % 16 * 5
    o64 = W[np.arange(64) // 16]
    B64 = np.ascontiguousarray(Qd[:, s64])
    print("multifreq 64 columns variant=%s" % variant, flush=True)
    f = _solver(gm, sol, "mixed", "DIAG", variant=variant)
    X, st = _freq_solve(f, B64, o64)
    assert all(s[3] for s in st)
    cm, ck, c2 = O.coeffs(*sol)
    for c in (0, 21, 42, 63):   # one column of each frequency
        rp, ci, val = om.sysmat(cm, ck, c2, o64[c])
        x, it, err = O.bicgstab(rp, ci, val, B64[:, c].copy(), 1e-15, 20000, True)
        r = np.linalg.norm(X[:, c] - x) / np.linalg.norm(x)
        assert r <= 1e-8, (variant, c, r)


def test_multifreq_graph_replay_identical(cyl2):
    case, gm, om, sol, B, src, om_col = cyl2
    out = {}
    for graphs in (True, False):
        f = _solver(gm, sol, "mixed", "DIAG", graphs=graphs, tol=1e-10)
        X1, st1 = _freq_solve(f, B, om_col)
        X2, st2 = _freq_solve(f, B, om_col)   # replays
        assert np.array_equal(X1, X2) and st1 == st2
        out[graphs] = (X1, st1, f.LastCounts())
    assert np.array_equal(out[True][0], out[False][0]) and out[True][1] == out[False][1]
    assert out[True][2][1] > 0 and out[False][2][1] == 0


def test_multifreq_refusals(cyl2):
    """BiCGSTAB, the single-column FP64 mode and non-finite frequencies are
    refused with a status code."""
    case, gm, om, sol, B, src, om_col = cyl2
    dB = L.DeviceArray.from_host(B)
    dX = L.DeviceArray(B.shape, np.complex128)
    f = T.FwdSolver(gm, "BICGSTAB", 1e-10, complex=True)
    f.Allocate()
    f.Reset(T.Solution(*sol), 0.5)
    with pytest.raises(L.TmiError, match="COCG"):
        f.CalcFieldsFreqsDevice(B.shape[1], dB.ptr, dX.ptr, om_col)
    g = _solver(gm, sol, "fp64", "DIAG")
    with pytest.raises(L.TmiError, match="block modes"):
        g.CalcFieldsFreqsDevice(B.shape[1], dB.ptr, dX.ptr, om_col)
    h = _solver(gm, sol, "mixed", "DIAG")
    bad = om_col.copy()
    bad[3] = np.nan
    with pytest.raises(L.TmiError, match="not finite"):
        h.CalcFieldsFreqsDevice(B.shape[1], dB.ptr, dX.ptr, bad)
