"""tmi_solver_project_all_device_noalloc: ProjectAll on resident fields with
the library's fixed device scratch instead of per-call allocations.  Checked
against tmi_solver_project_all_device (the same projection kernel, so the
results are bit-identical) on two calls in a row, the second smaller than
the first, complex and real, linear and log; sizes beyond the scratch are
refused before any device work.  Through the C-ABI on the GPU."""
import ctypes as C

import numpy as np
import pytest
from scipy import sparse

from tests.cases import CASES

pytestmark = pytest.mark.gpu

T = pytest.importorskip("toastpp_amd")
from toastpp_amd import _lib as L  # noqa: E402


def _case(name, cplx, seed):
    case = CASES[name]
    gm = T.Mesh(case["mesh"])
    gm.ReadQM(case["qm"])
    n = gm.NodeCount()
    nq, nm, nqm = gm.qm_dims()
    f = T.FwdSolver(gm, "BICGSTAB" if cplx else "CG", 1e-10, complex=cplx)
    rng = np.random.default_rng(seed)
    ld = nq + 3   # row stride wider than the source count
    dt = np.complex128 if cplx else np.float64
    phi = rng.standard_normal((n, ld)) + (1j * rng.standard_normal((n, ld)) if cplx else 0.0)
    phi = np.ascontiguousarray(np.abs(phi) + 0.1 if not cplx else phi, dt)
    mvec = sparse.csr_matrix(gm.Mvec(width=2.0, ref=1.4).T)
    if not cplx:
        mvec = sparse.csr_matrix(mvec.real)
    return gm, f, L.DeviceArray.from_host(phi), ld, mvec, nqm, dt


def _project(entry, f, mvec, dphi, ld, nqm, dt, dscale):
    rp = np.ascontiguousarray(mvec.indptr, np.int32)
    ci = np.ascontiguousarray(mvec.indices, np.int32)
    val = np.ascontiguousarray(mvec.data, dt).view(np.float64)
    out = np.full(nqm, np.nan, dt)
    L.call(entry, f._h, mvec.shape[0], rp, ci, val, C.c_void_p(dphi.ptr), ld, dscale, out.view(np.float64))
    return out


@pytest.mark.parametrize("cplx", [True, False])
def test_noalloc_projections_equal_allocating_entry_point(cplx):
    # first call: the 3D golden case (80 sources x 80 detectors); second: the
    # smaller 2D golden case (32 x 32) on the same GPU scratch
    for name, seed in (("3d_cplx_direct", 1), ("2d_real_cg", 2)):
        gm, f, dphi, ld, mvec, nqm, dt = _case(name, cplx, seed)
        for dscale in (L.DATA_LIN, L.DATA_LOG):
            a = _project("tmi_solver_project_all_device", f, mvec, dphi, ld, nqm, dt, dscale)
            b = _project("tmi_solver_project_all_device_noalloc", f, mvec, dphi, ld, nqm, dt, dscale)
            assert np.all(np.isfinite(b))
            assert np.array_equal(a, b), (name, dscale)


def test_noalloc_refuses_sizes_beyond_the_scratch():
    gm, f, dphi, ld, mvec, nqm, dt = _case("2d_real_cg", True, 3)
    nm = mvec.shape[0]
    rp = np.zeros(nm + 1, np.int32)
    rp[-1] = 8 << 20   # 8M mvec entries: ~100 MiB of staging; refused before the CSR arrays are read
    ci = np.zeros(1, np.int32)
    val = np.zeros(2)
    out = np.zeros(nqm, np.complex128)
    st = L.lib().tmi_solver_project_all_device_noalloc(f._h, nm, rp, ci, val, C.c_void_p(dphi.ptr), ld, L.DATA_LIN,
                                                       out.view(np.float64))
    assert st == L.TMI_ERR_INVALID_ARG
    assert b"scratch" in L.lib().tmi_last_error()
    # the scratch still serves a call that fits
    a = _project("tmi_solver_project_all_device", f, mvec, dphi, ld, nqm, dt, L.DATA_LOG)
    b = _project("tmi_solver_project_all_device_noalloc", f, mvec, dphi, ld, nqm, dt, L.DATA_LOG)
    assert np.array_equal(a, b)
