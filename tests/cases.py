"""Golden known-answer cases of the reference's own regression suite.

test/2D/fwdfem/Makefile and test/3D/fwdfem/Makefile run fwdfem / fwdfem_cw
with canned answers and compare fmod.fem / farg.fem (log amplitude / phase
of every linked projection) with the vectors in testdata.zip using
`cmpvec -r=<tol>` (per-element relative difference).  The meshes, QM files
and zip archives are copied unchanged under tests/golden/ref2d and ref3d.
"""
import os
import zipfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
G2 = os.path.join(HERE, "golden", "ref2d")
G3 = os.path.join(HERE, "golden", "ref3d")

# name: mesh, qm, zip, mua, mus, n, freq [MHz], solver, precon, {file: tol}
CASES = {
    # test/2D/fwdfem/Makefile:35-63 (fwdfem_cw, CG, tol 1e-15)
    "2d_real_cg": dict(mesh=os.path.join(G2, "circle25_32.msh"), qm=os.path.join(G2, "circle25_32x32.qm"),
                       zip=os.path.join(G2, "testdata.zip"), mua=0.025, mus=2.0, n=1.4, freq=0.0, solver="cg",
                       precon=None, files={"fmod_real_cg.fem": 1e-8}),
    "2d_real_cg_diag": dict(mesh=os.path.join(G2, "circle25_32.msh"), qm=os.path.join(G2, "circle25_32x32.qm"),
                            zip=os.path.join(G2, "testdata.zip"), mua=0.025, mus=2.0, n=1.4, freq=0.0, solver="cg",
                            precon="diag", files={"fmod_real_cg_diag.fem": 1e-8}),
    # test/2D/fwdfem/Makefile:80-111 (fwdfem, BiCGSTAB, 100 MHz, tol 1e-15)
    "2d_cplx_bicgstab": dict(mesh=os.path.join(G2, "circle25_32.msh"), qm=os.path.join(G2, "circle25_32x32.qm"),
                             zip=os.path.join(G2, "testdata.zip"), mua=0.025, mus=2.0, n=1.4, freq=100.0,
                             solver="bicgstab", precon=None,
                             files={"fmod_cplx_bicgstab.fem": 1e-7, "farg_cplx_bicgstab.fem": 1e-6}),
    "2d_cplx_bicgstab_diag": dict(mesh=os.path.join(G2, "circle25_32.msh"), qm=os.path.join(G2, "circle25_32x32.qm"),
                                  zip=os.path.join(G2, "testdata.zip"), mua=0.025, mus=2.0, n=1.4, freq=100.0,
                                  solver="bicgstab", precon="diag",
                                  files={"fmod_cplx_bicgstab_diag.fem": 1e-7, "farg_cplx_bicgstab_diag.fem": 1e-6}),
    # test/2D/fwdfem/Makefile:128-148
    "2d_cplx_gmres": dict(mesh=os.path.join(G2, "circle25_32.msh"), qm=os.path.join(G2, "circle25_32x32.qm"),
                          zip=os.path.join(G2, "testdata.zip"), mua=0.025, mus=2.0, n=1.4, freq=100.0,
                          solver="gmres", precon=None,
                          files={"fmod_cplx_gmres.fem": 1e-7, "farg_cplx_gmres.fem": 1e-7}),
    "2d_cplx_gmres_diag": dict(mesh=os.path.join(G2, "circle25_32.msh"), qm=os.path.join(G2, "circle25_32x32.qm"),
                               zip=os.path.join(G2, "testdata.zip"), mua=0.025, mus=2.0, n=1.4, freq=100.0,
                               solver="gmres", precon="diag",
                               files={"fmod_cplx_gmres_diag.fem": 1e-7, "farg_cplx_gmres_diag.fem": 1e-7}),
    # test/3D/fwdfem/Makefile:10-74 (cyl2 Tet4, mua 0.01, mus 1, 100 MHz, tol 1e-8)
    "3d_real_direct": dict(mesh=os.path.join(G3, "cyl2.msh"), qm=os.path.join(G3, "cyl_5ring.qm"),
                           zip=os.path.join(G3, "testdata.zip"), mua=0.01, mus=1.0, n=1.4, freq=0.0, solver="cg",
                           precon="diag", files={"fmod_real_direct.fem": 1e-8, "fmod_real_cg.fem": 1e-8}),
    "3d_cplx_direct": dict(mesh=os.path.join(G3, "cyl2.msh"), qm=os.path.join(G3, "cyl_5ring.qm"),
                           zip=os.path.join(G3, "testdata.zip"), mua=0.01, mus=1.0, n=1.4, freq=100.0,
                           solver="bicgstab", precon="diag",
                           files={"fmod_cplx_direct.fem": 1e-8, "farg_cplx_direct.fem": 1e-8,
                                  "fmod_cplx_bicgstab.fem": 1e-8, "farg_cplx_bicgstab.fem": 1e-8}),
}


def golden_vector(zpath, name):
    with zipfile.ZipFile(zpath) as zf:
        txt = zf.read(name).decode()
    return np.array(txt.replace("[", " ").replace("]", " ").split(), dtype=np.float64)


def cmpvec(a, b):
    """max per-element |a-b| / max(|a|,|b|) (src/bintools/cmpvec.cc)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(a), np.abs(b))
    den[den == 0] = 1.0
    return float(np.max(np.abs(a - b) / den))


def compare(case, proj):
    """[(file, rel_diff, tol)] of the projections against the golden files."""
    out = []
    for fname, tol in case["files"].items():
        ref = golden_vector(case["zip"], fname)
        got = proj.real if fname.startswith("fmod") else proj.imag
        out.append((fname, cmpvec(got, ref), tol))
    return out
