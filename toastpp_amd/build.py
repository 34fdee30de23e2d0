"""Build libtoastpp_amd.so (HIP kernels + C-ABI) for gfx950, in tree.

Usage: python -m toastpp_amd.build  (or toastpp_amd.build.build())
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libtoastpp_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("TMI_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "--offload-arch=" + ARCH, "-I" + os.path.join(HERE, "..", "include"),
          "-Wall", "-Wno-unused-function",
          # packed FP32 instructions stay off (compiler-generated v_pk_* FP32 math)
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
# files whose FP64 arithmetic must follow the reference operation by operation
NO_CONTRACT = {"mesh_kernels.hip", "mesh_io.cpp", "jacobian_kernels.hip", "raster.cpp", "raster_kernels.hip"}
SOURCES = ["mesh_kernels.hip", "krylov_kernels.hip", "jacobian_kernels.hip", "raster_kernels.hip", "gmres.hip", "mesh_io.cpp",
           "order.cpp", "solver.cpp", "jacobian_api.cpp", "serve.cpp", "data_io.cpp", "api.cpp", "noalloc.hip"]


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    flags = list(COMMON)
    flags.append("-ffp-contract=off" if src in NO_CONTRACT else "-ffp-contract=fast")
    cmd = [HIPCC] + flags + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


BIN = os.path.join(HERE, "bin")
CXX = os.environ.get("CXX", "g++")
# command-line programs on the C-ABI (host code only): name -> extra flags
TOOLS = {"fwdfem": [], "fwdfem_cw": ["-DTMI_FWDFEM_CW"]}


def _build_tool(name, flags):
    out = os.path.join(BIN, name)
    cmd = [CXX, "-O2", "-std=c++17", "-Wall"] + flags + [os.path.join(CSRC, "fwdfem.cpp"), "-o", out + ".tmp",
                                                         "-L" + HERE, "-l:libtoastpp_amd.so",
                                                         "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath-link,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("tool build failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(out + ".tmp", out)
    return out


def build(verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    os.makedirs(BIN, exist_ok=True)
    tools = [_build_tool(name, flags) for name, flags in TOOLS.items()]
    if verbose:
        print("built", LIB, *tools)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
