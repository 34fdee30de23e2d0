"""Data files and the command-line forward model, host side (no GPU):
NIM images (toast.ReadNim / WriteNim, fwdfem's ReadNim / WriteNIM), the
bracketed vector text of the .fem files, and the dialogue of bin/fwdfem and
bin/fwdfem_cw up to the point where GPU work would start (every refusal of
an option the library does not provide happens before it)."""
import os
import subprocess
import zipfile

import numpy as np
import pytest

from tests.cases import CASES, golden_vector
from tests.conftest import ROOT

TMI = pytest.importorskip("toastpp_amd")
from toastpp_amd import _lib as L  # noqa: E402

G3 = os.path.join(ROOT, "tests", "golden", "ref3d")
G2 = os.path.join(ROOT, "tests", "golden", "ref2d")
BIN = os.path.join(ROOT, "toastpp_amd", "bin")


def _nim_values_by_text(path):
    """Independent parse of a one-image NIM file: the tokens after the
    'Image 0' line."""
    body = open(path).read().split("EndHeader", 1)[1]
    toks = body.split()
    assert toks[0] == "Image" and toks[1] == "0"
    return np.array([float(t) for t in toks[2:]])


@pytest.mark.parametrize("name", ["mua_tgt_cyl2.nim", "mus_tgt_cyl2.nim"])
def test_read_reference_nim(name):
    path = os.path.join(G3, name)
    img = TMI.ReadNim(path)
    ref = _nim_values_by_text(path)
    assert img.shape == (9043,)
    assert np.array_equal(img, ref)
    # the 0-based index and "past the end" select the same single image
    assert np.array_equal(TMI.ReadNim(path, 0), ref)
    assert np.array_equal(TMI.ReadNim(path, 5), ref)


def test_nim_series_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    imgs = rng.standard_normal((3, 257)) * 10.0 ** rng.integers(-6, 6, (3, 257))
    for style, rtol in ((L.TMI_NIM_TOASTMOD, 5e-13), (L.TMI_NIM_FWDFEM, 5e-10)):
        p = str(tmp_path / ("s%d.nim" % style))
        L.call("tmi_write_nim", p.encode(), b"mesh.msh", 257, 3, np.ascontiguousarray(imgs), style)
        for idx, want in ((0, 0), (1, 1), (2, 2), (-1, 2), (9, 2)):
            got = TMI.ReadNim(p, idx)
            np.testing.assert_allclose(got, imgs[want], rtol=rtol, atol=0)
        head = open(p).read().splitlines()[:6]
        assert head == ["NIM", "Mesh = mesh.msh", "SolutionType = N/A", "ImageSize = 257", "EndHeader", "Image 0"]


def test_write_nim_python(tmp_path):
    p = str(tmp_path / "w.nim")
    v = np.linspace(-1.0, 2.0, 11)
    TMI.WriteNim(p, "x.msh", v)
    np.testing.assert_allclose(TMI.ReadNim(p), v, rtol=1e-12, atol=1e-300)
    assert "%.12e" % v[3] in open(p).read()


def test_nim_errors(tmp_path):
    bad = tmp_path / "bad.nim"
    bad.write_text("NOTNIM\n")
    with pytest.raises(L.TmiError):
        TMI.ReadNim(str(bad))
    short = tmp_path / "short.nim"
    short.write_text("NIM\nImageSize = 4\nEndHeader\nImage 0\n1 2 3\n")
    with pytest.raises(L.TmiError):
        TMI.ReadNim(str(short))
    with pytest.raises(L.TmiError):
        TMI.ReadNim(str(tmp_path / "missing.nim"))


def _read_vector(path):
    import ctypes as C
    n = C.c_int()
    L.call("tmi_read_vector", path.encode(), C.byref(n), None)
    v = np.zeros(n.value)
    L.call("tmi_read_vector", path.encode(), C.byref(n), L.ptr(v))
    return v


def test_fem_vector_files(tmp_path):
    case = CASES["3d_cplx_direct"]
    with zipfile.ZipFile(case["zip"]) as zf:
        zf.extract("fmod_cplx_direct.fem", tmp_path)
    p = str(tmp_path / "fmod_cplx_direct.fem")
    got = _read_vector(p)
    assert np.array_equal(got, golden_vector(case["zip"], "fmod_cplx_direct.fem"))
    out = str(tmp_path / "w.fem")
    L.call("tmi_write_vector", out.encode(), got.size, got)
    txt = open(out).read()
    assert txt.startswith("[") and txt.rstrip().endswith("]")
    np.testing.assert_allclose(_read_vector(out), got, rtol=1e-10, atol=0)


def _run(prog, answers, cwd, args=()):
    r = subprocess.run([os.path.join(BIN, prog)] + list(args), input=answers, capture_output=True, text=True,
                       cwd=cwd, timeout=60)
    return r.returncode, r.stdout, r.stderr


@pytest.fixture()
def fwd_dir(tmp_path):
    """The layout of test/2D/fwdfem and test/3D/fwdfem: meshes one level up."""
    (tmp_path / "fwdfem").mkdir()
    os.symlink(G2, tmp_path / "meshes")
    os.symlink(G3, tmp_path / "meshes3")
    return str(tmp_path / "fwdfem")


@pytest.mark.parametrize("answers,msg", [
    # test/2D/fwdfem/Makefile targets the library does not provide
    ("../meshes/circle25_32.msh ../meshes/circle25_32x32.qm 2 3 1e-15 10000 0 1 2 2 2 2 100 2 0.025 2 2 2 1.4 0 1",
     "BICG is not provided"),
    ("../meshes/circle25_32.msh ../meshes/circle25_32x32.qm 2 4 1e-15 10000 2 1 2 2 2 2 100 2 0.025 2 2 2 1.4 0 1",
     "ICH is not provided"),
    ("../meshes/circle25_32.msh ../meshes/circle25_32x32.qm 2 4 1e-15 10000 3 1 2 2 2 2 100 2 0.025 2 2 2 1.4 0 1",
     "DILU is not provided"),
    ("../meshes/circle25_32.msh ../meshes/circle25_32x32.qm 2 2 1e-15 10000 0", "CG is not defined for the complex"),
    ("../meshes/circle25_32.msh ../meshes/circle25_32x32.qm 2 4 1e-15 0 1 1 3 2", "COSINE is not provided"),
    ("../meshes/circle25_32.msh ../meshes/circle25_32x32.qm 2 4 1e-15 0 1 1 2 2 2 2 100 2 0.025 2",
     "input ended"),
    ("../meshes/nothere.msh", "reading mesh"),
])
def test_fwdfem_refusals(fwd_dir, answers, msg):
    rc, out, err = _run("fwdfem", answers, fwd_dir)
    assert rc == 1, (out, err)
    assert msg in err, err


def test_fwdfem_param_file_and_log(fwd_dir):
    """A parameter file answers the dialogue (fwdfem.prm syntax); the log
    records every choice in the same syntax up to the refusal."""
    prm = os.path.join(fwd_dir, "run.prm")
    with open(prm, "w") as f:
        f.write("LINSOLVER = BICG ; not provided\nMESHFILE = ../meshes3/cyl2.msh\nQMFILE = ../meshes3/cyl_5ring.qm\n"
                "DATASCALE = LOG\nLINSOLVER_TOL = 1e-12\nLINSOLVER_MAXIT = 0\nLINSOLVER_PRECON = DIAG\n")
    rc, out, err = _run("fwdfem", "", fwd_dir, [prm, "run.log"])
    assert rc == 1 and "BICG is not provided" in err
    log = open(os.path.join(fwd_dir, "run.log")).read().splitlines()
    assert log[2:5] == ["MESHFILE = ../meshes3/cyl2.msh", "QMFILE = ../meshes3/cyl_5ring.qm", "DATASCALE = LOG"]


def test_fwdfem_cw_has_no_frequency(fwd_dir):
    # fwdfem_cw asks no modulation frequency: "2" after the measurement width
    # is already the mua choice, whose region value is missing here
    rc, out, err = _run("fwdfem_cw", "../meshes/circle25_32.msh ../meshes/circle25_32x32.qm 2 2 1e-15 10000 0 1 2 2 2 2 2",
                        fwd_dir)
    assert rc == 1 and "Value for region 0" in err, err
    assert "frequency" not in out
