"""bin/fwdfem and bin/fwdfem_cw on the GPU, driven exactly as the
reference's own regression suite drives fwdfem / fwdfem_cw:
test/2D/fwdfem/Makefile and test/3D/fwdfem/Makefile pipe a fixed answer
line into the program, which writes fmod.fem / farg.fem, and compare those
with the vectors in testdata.zip at the Makefile's `cmpvec -r` tolerances.
Every target of those Makefiles whose mesh, solver and preconditioner the
library provides is run here (the ICH and DILU targets and the Tri10 mesh are
refused, see test_data_io.py).  A parameter-file run with nodal
parameters from the reference's NIM files is checked against the oracle."""
import os
import subprocess

import numpy as np
import pytest
from scipy import sparse
from scipy.sparse import linalg as spla

from oracle import oracle as O
from tests.cases import cmpvec, golden_vector
from tests.conftest import ROOT

pytestmark = pytest.mark.gpu

G2 = os.path.join(ROOT, "tests", "golden", "ref2d")
G3 = os.path.join(ROOT, "tests", "golden", "ref3d")
BIN = os.path.join(ROOT, "toastpp_amd", "bin")

M2 = "../meshes/circle25_32.msh ../meshes/circle25_32x32.qm "
M3 = "../meshes/cyl2.msh ../meshes/cyl_5ring.qm "
# target: (suite dir, program, answer line, {output file: (golden file, tol)})
TARGETS = {
    # test/2D/fwdfem/Makefile:35-57
    "2d_fwd_real_cg": (G2, "fwdfem_cw", M2 + "2 2 1e-15 10000 0 1 2 2 2 2 2 0.025 2 2 2 1.4 0 1",
                       {"fmod.fem": ("fmod_real_cg.fem", 1e-8)}),
    "2d_fwd_real_cg_diag": (G2, "fwdfem_cw", M2 + "2 2 1e-15 10000 1 1 2 2 2 2 2 0.025 2 2 2 1.4 0 1",
                            {"fmod.fem": ("fmod_real_cg_diag.fem", 1e-8)}),
    # test/2D/fwdfem/Makefile:80-100
    "2d_fwd_cplx_bicgstab": (G2, "fwdfem", M2 + "2 4 1e-15 10000 0 1 2 2 2 2 100 2 0.025 2 2 2 1.4 0 1",
                             {"fmod.fem": ("fmod_cplx_bicgstab.fem", 1e-7),
                              "farg.fem": ("farg_cplx_bicgstab.fem", 1e-6)}),
    "2d_fwd_cplx_bicgstab_diag": (G2, "fwdfem", M2 + "2 4 1e-15 10000 1 1 2 2 2 2 100 2 0.025 2 2 2 1.4 0 1",
                                  {"fmod.fem": ("fmod_cplx_bicgstab_diag.fem", 1e-7),
                                   "farg.fem": ("farg_cplx_bicgstab_diag.fem", 1e-6)}),
    # test/2D/fwdfem/Makefile:128-148
    "2d_fwd_cplx_gmres": (G2, "fwdfem", M2 + "2 5 1e-15 10000 0 1 2 2 2 2 100 2 0.025 2 2 2 1.4 0 1",
                          {"fmod.fem": ("fmod_cplx_gmres.fem", 1e-7), "farg.fem": ("farg_cplx_gmres.fem", 1e-7)}),
    "2d_fwd_cplx_gmres_diag": (G2, "fwdfem", M2 + "2 5 1e-15 10000 1 1 2 2 2 2 100 2 0.025 2 2 2 1.4 0 1",
                               {"fmod.fem": ("fmod_cplx_gmres_diag.fem", 1e-7),
                                "farg.fem": ("farg_cplx_gmres_diag.fem", 1e-7)}),
    # test/3D/fwdfem/Makefile:10-30
    "3d_fwd_real_direct": (G3, "fwdfem_cw", M3 + "2 1 1 2 2 2 2 1 0.01 1 1 1 1.4 0 1",
                           {"fmod.fem": ("fmod_real_direct.fem", 1e-8)}),
    "3d_fwd_real_cg": (G3, "fwdfem_cw", M3 + "2 2 1e-15 10000 0 1 2 2 2 2 1 0.01 1 1 1 1.4 0 1",
                       {"fmod.fem": ("fmod_real_cg.fem", 1e-8)}),
    # test/3D/fwdfem/Makefile:43-63
    "3d_fwd_cplx_direct": (G3, "fwdfem", M3 + "2 1 1 2 2 2 2 100 1 0.01 1 1 1 1.4 0 1",
                           {"fmod.fem": ("fmod_cplx_direct.fem", 1e-8), "farg.fem": ("farg_cplx_direct.fem", 1e-8)}),
    "3d_fwd_cplx_bicgstab": (G3, "fwdfem", M3 + "2 4 1e-15 10000 0 1 2 2 2 2 100 1 0.01 1 1 1 1.4 0 1",
                             {"fmod.fem": ("fmod_cplx_bicgstab.fem", 1e-8),
                              "farg.fem": ("farg_cplx_bicgstab.fem", 1e-8)}),
}


def _suite_dir(tmp_path, meshes):
    """test/<2D|3D>/fwdfem layout: the program runs in fwdfem/, meshes in ../meshes."""
    os.symlink(meshes, tmp_path / "meshes")
    (tmp_path / "fwdfem").mkdir()
    return str(tmp_path / "fwdfem")


def _run(prog, answers, cwd, args=()):
    print("running %s %s <<< %r" % (prog, " ".join(args), answers), flush=True)
    r = subprocess.run([os.path.join(BIN, prog)] + list(args), input=answers, capture_output=True, text=True,
                       cwd=cwd, timeout=110)
    assert r.returncode == 0, "%s failed (%d):\n%s\n%s" % (prog, r.returncode, r.stdout[-2000:], r.stderr)
    return r.stdout


def _read_fem(path):
    txt = open(path).read()
    return np.array(txt.replace("[", " ").replace("]", " ").split(), dtype=np.float64)


@pytest.mark.parametrize("target", sorted(TARGETS))
def test_reference_suite_target(tmp_path, target):
    meshes, prog, answers, outs = TARGETS[target]
    cwd = _suite_dir(tmp_path, meshes)
    _run(prog, answers, cwd)
    zpath = os.path.join(meshes, "testdata.zip")
    for out, (gold, tol) in outs.items():
        got = _read_fem(os.path.join(cwd, out))
        ref = golden_vector(zpath, gold)
        assert got.shape == ref.shape, (out, got.shape, ref.shape)
        d = cmpvec(got, ref)
        print("%s %s vs %s: cmpvec %.3g (tol %g)" % (target, out, gold, d, tol))
        assert d <= tol, "%s: %s vs %s rel diff %.3g > %g" % (target, out, gold, d, tol)
    assert os.path.exists(os.path.join(cwd, "fwdfem.out"))


def _nim_by_text(path):
    toks = open(path).read().split("EndHeader", 1)[1].split()
    return np.array([float(t) for t in toks[2:]])


@pytest.mark.parametrize("prog", ["fwdfem", "fwdfem_cw"])
def test_param_file_nim_parameters_vs_oracle(tmp_path, prog):
    """fwdfem.prm-style run (test/3D/fwdfem/fwdfem.prm keys) on cyl2 with
    mua and mus from the reference's NIM images: data files and nodal field
    images against the oracle (its assembly, profiles and projection; the
    solve by sparse LU of the oracle's FP64 matrix)."""
    cplx = prog == "fwdfem"
    cwd = _suite_dir(tmp_path, G3)
    prm = os.path.join(cwd, "nim.prm")
    with open(prm, "w") as f:
        # tolerance and limit of the suite's iterative targets: far detectors
        # see ~e-13 of the peak amplitude, so their per-element accuracy
        # needs the relative residual at the FP64 floor
        f.write("LINSOLVER = %s\nLINSOLVER_TOL = 1e-15\nLINSOLVER_MAXIT = 10000\nLINSOLVER_PRECON = DIAG\n"
                "DATASCALE = LOG\nMESHFILE = ../meshes/cyl2.msh\nQMFILE = ../meshes/cyl_5ring.qm\n"
                "SOURCETYPE = NEUMANN\nSOURCEPROFILE = GAUSSIAN\nSOURCEWIDTH = 2\n"
                "MEASUREMENTPROFILE = GAUSSIAN\nMEASUREMENTWIDTH = 2\n%s"
                "RESET_MUA = NIM ../meshes/mua_tgt_cyl2.nim\nRESET_MUS = NIM ../meshes/mus_tgt_cyl2.nim\n"
                "RESET_N = HOMOG 1.400000\n" % ("BICGSTAB" if cplx else "CG", "FREQ = 100\n" if cplx else ""))
    _run(prog, "1 1", cwd, [prm, "nim.log"])

    om = O.Mesh(os.path.join(G3, "cyl2.msh"))
    n = om.n
    mua = _nim_by_text(os.path.join(G3, "mua_tgt_cyl2.nim"))
    mus = _nim_by_text(os.path.join(G3, "mus_tgt_cyl2.nim"))
    assert mua.size == n and mus.size == n
    cm, ck, c2 = O.coeffs(mua, mus, np.full(n, 1.4))
    omega = 100.0 * 2 * np.pi * 1e-6 if cplx else 0.0
    rp, ci, val = om.sysmat(cm, ck, c2, omega)
    Q, M, lp, li = om.read_qm(os.path.join(G3, "cyl_5ring.qm"))
    qv = om.qvec_gaussian(Q, 2.0)
    mv = om.qvec_gaussian(M, 2.0) * c2[None, :]
    if not cplx:
        assert np.all(val.imag == 0)
        val = np.ascontiguousarray(val.real)
    K = sparse.csr_matrix((val, ci, rp), shape=(n, n)).tocsc()
    phi = spla.splu(K).solve(np.ascontiguousarray(qv.T)).T
    proj = O.project(phi.astype(np.complex128), mv, lp, li, logscale=True)

    fmod = _read_fem(os.path.join(cwd, "fmod.fem"))
    d = cmpvec(fmod, proj.real)
    print("%s NIM parameters: fmod cmpvec %.3g" % (prog, d))
    assert d <= 1e-8
    if cplx:
        d = cmpvec(_read_fem(os.path.join(cwd, "farg.fem")), proj.imag)
        print("%s NIM parameters: farg cmpvec %.3g" % (prog, d))
        assert d <= 1e-8
    # nodal field images, source 0 and the last source
    from toastpp_amd import ReadNim
    if cplx:
        lnmod = os.path.join(cwd, "phi_lnmod.nim")
        arg = os.path.join(cwd, "phi_arg.nim")
        for q in (0, len(Q) - 1):
            assert cmpvec(ReadNim(lnmod, q), np.log(np.abs(phi[q]))) <= 1e-8
            # phases compared on the unit circle (no branch-cut flips at +-pi)
            assert np.max(np.abs(np.exp(1j * ReadNim(arg, q)) - np.exp(1j * np.angle(phi[q])))) <= 1e-8
    else:
        lphi = os.path.join(cwd, "lphi.nim")
        for q in (0, len(Q) - 1):
            pos = phi[q] > 0
            assert pos.mean() > 0.99
            assert cmpvec(ReadNim(lphi, q)[pos], np.log(phi[q][pos])) <= 1e-8
    log = open(os.path.join(cwd, "nim.log")).read()
    assert "RESET_MUA = NIM ../meshes/mua_tgt_cyl2.nim" in log and "T(solve)" in log
