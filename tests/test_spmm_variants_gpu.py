"""The 64-column SpMM variants (direct gathers, LDS staging through
registers, LDS DMA with double buffering, row blocks with shared
gathers) inside the mixed-precision block solve: each against the FP64
oracle on the 3D golden case, and against each other on a mesh large enough that every workgroup of the DMA kernel walks
several row blocks and owns more than one reduction slot.  Through the C-ABI
on the GPU."""
import numpy as np
import pytest
from scipy import sparse

from oracle import oracle as O
from tests.cases import CASES

pytestmark = pytest.mark.gpu

TMI = pytest.importorskip("toastpp_amd")
from toastpp_amd import _lib as L  # noqa: E402
from toastpp_amd import meshgen as G  # noqa: E402

VARIANTS = ["lds", "direct", "dma", "rowblock", "rowblock_pipelined", "rowblock_mfma"]


@pytest.fixture(scope="module")
def golden_3d():
    case = CASES["3d_cplx_direct"]
    om = O.Mesh(case["mesh"])
    n = om.n
    cm, ck, c2 = O.coeffs(np.full(n, case["mua"]), np.full(n, case["mus"]), np.full(n, case["n"]))
    rp, ci, val = om.sysmat(cm, ck, c2, case["freq"] * 2 * np.pi * 1e-6)
    Q = om.read_qm(case["qm"])[0]
    qv = om.qvec_gaussian(Q[:4], 2.0)
    ref = [O.bicgstab(rp, ci, val, qv[i].astype(np.complex128), 1e-15, 20000, True)[0] for i in range(4)]
    return case, ref


@pytest.mark.parametrize("variant", VARIANTS)
def test_variant_fields_vs_oracle_3d(golden_3d, variant):
    case, ref = golden_3d
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    n = gm.NodeCount()
    fws = TMI.FwdSolver(gm, "BICGSTAB", 1e-12, complex=True)
    fws.SetMode("mixed")
    fws.SetPrecon("DIAG")
    fws.SetSpmmVariant(variant)
    fws.Allocate()
    fws.Reset(TMI.Solution(np.full(n, case["mua"]), np.full(n, case["mus"]), np.full(n, case["n"])),
              case["freq"] * 2 * np.pi * 1e-6)
    phi = fws.CalcFields(sparse.csr_matrix(gm.Qvec(width=2.0).T))
    assert all(s[3] for s in fws.stats)
    for i in range(4):
        rel = np.linalg.norm(phi[i] - ref[i]) / np.linalg.norm(ref[i])
        assert rel <= 1e-8, "%s column %d rel L2 %.3g" % (variant, i, rel)


def test_variants_agree_on_cylinder():
    """31.5k-node cylinder, 64 sources at 100 MHz, tolerance 1e-10: every
    variant's fields agree with the direct-gather variant's to 1e-8
    relative L2 per source."""
    nodes, el = G.cylinder(25.0, 50.0, 20, 24)
    Q = np.vstack([G.ring(32, 25.0, z) for z in (-12.5, 12.5)])
    mesh = TMI.Mesh(data=(nodes, el))
    mesh.SetQM(Q, Q[:1])
    mua, mus, ref = G.params(nodes, 3)
    B = np.ascontiguousarray(mesh.Qvec(width=2.0).toarray())
    n, nq = B.shape
    assert n > 4096 and nq == 64
    dB = L.DeviceArray.from_host(B)
    out = {}
    for variant in VARIANTS:
        f = TMI.FwdSolver(mesh, "BICGSTAB", 1e-10, complex=True)
        f.SetMode("mixed")
        f.SetPrecon("DIAG")
        f.SetSpmmVariant(variant)
        f.Allocate()
        f.Reset(TMI.Solution(mua, mus, ref), 100.0 * 2 * np.pi * 1e-6)
        dX = L.DeviceArray((n, nq), np.complex128)
        stats = f.CalcFieldsDevice(nq, dB.ptr, dX.ptr)
        assert all(s[3] for s in stats), variant
        out[variant] = dX.to_host()
    for variant in ("lds", "dma", "rowblock", "rowblock_pipelined", "rowblock_mfma"):
        rel = np.linalg.norm(out[variant] - out["direct"], axis=0) / np.linalg.norm(out["direct"], axis=0)
        assert rel.max() <= 1e-8, "%s vs direct: worst column rel L2 %.3g" % (variant, rel.max())
    # the pipelined row-block kernel sums every row in the row-block kernel's
    # order: identical products, identical fields
    assert np.array_equal(out["rowblock_pipelined"], out["rowblock"])


@pytest.mark.parametrize("mode", ["mixed", "fp64_block"])
@pytest.mark.parametrize("tile", [2, 4, 8, 16, 32])
def test_narrow_tiles_vs_oracle_3d(golden_3d, mode, tile):
    """Tiles narrower than 64 columns (the narrow-SpMM kernel, complex FP32
    and FP64, and its FP64 residual form) against the FP64 oracle."""
    case, ref = golden_3d
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    n = gm.NodeCount()
    fws = TMI.FwdSolver(gm, "BICGSTAB", 1e-12, complex=True)
    fws.SetMode(mode)
    fws.SetPrecon("DIAG")
    fws.SetTile(tile)
    fws.Allocate()
    fws.Reset(TMI.Solution(np.full(n, case["mua"]), np.full(n, case["mus"]), np.full(n, case["n"])),
              case["freq"] * 2 * np.pi * 1e-6)
    qv = sparse.csr_matrix(gm.Qvec(width=2.0).T)[:4]
    phi = fws.CalcFields(qv)
    assert all(s[3] for s in fws.stats)
    for i in range(4):
        rel = np.linalg.norm(phi[i] - ref[i]) / np.linalg.norm(ref[i])
        assert rel <= 1e-8, "%s tile %d column %d rel L2 %.3g" % (mode, tile, i, rel)


@pytest.mark.parametrize("solver", ["BICGSTAB", "COCG"])
@pytest.mark.parametrize("tile", [16, 32, 64])
def test_matrix_core_rowblock_tiles_vs_oracle_3d(golden_3d, solver, tile):
    """The matrix-core row-block kernel on its three tile widths (16, 32 and
    64 columns; 64 with the packed shadow residual under BiCGSTAB), both
    complex solvers, against the FP64 oracle."""
    case, ref = golden_3d
    gm = TMI.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    n = gm.NodeCount()
    fws = TMI.FwdSolver(gm, solver, 1e-12, complex=True)
    fws.SetMode("mixed")
    fws.SetPrecon("DIAG")
    fws.SetSpmmVariant("rowblock_mfma")
    fws.SetTile(tile)
    fws.Allocate()
    fws.Reset(TMI.Solution(np.full(n, case["mua"]), np.full(n, case["mus"]), np.full(n, case["n"])),
              case["freq"] * 2 * np.pi * 1e-6)
    qv = sparse.csr_matrix(gm.Qvec(width=2.0).T)
    phi = fws.CalcFields(qv)
    assert all(s[3] for s in fws.stats)
    for i in range(4):
        rel = np.linalg.norm(phi[i] - ref[i]) / np.linalg.norm(ref[i])
        assert rel <= 1e-8, "%s tile %d column %d rel L2 %.3g" % (solver, tile, i, rel)
