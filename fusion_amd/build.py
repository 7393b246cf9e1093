"""Build recipe: the gfx950 HIP library, the host scene generator and the CPU oracle.

``build()`` compiles every time it is called (hipcc cross-compiles without a GPU):
  fusion_amd/lib/libfusion_amd.so   HIP kernels + C ABI (include/fusion_amd.h)
  fusion_amd/lib/libfusion_amd_cxx.so  the C++ drop-in layer (include/fusion_amd_compat.hpp)
  fusion_amd/lib/libfu_scene.so     synthetic scenes (include/fusion_scene.h), g++ + OpenMP
  tests/cxx/_build/test_compat      the C++ layer's test program (run by tests/test_gpu_compat.py)
  oracle/_build/liboracle.so        CPU oracle (test infrastructure; oracle/Makefile)
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "fusion_amd")
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "lib")
OBJ = os.path.join(PKG, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# strict IEEE float order (no contraction) so the kernels match the CPU oracle bit for bit;
# packed-FP32 instructions disabled
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
HIP_SOURCES = ["pyramid.hip", "map.hip", "raycast.hip", "mesh.hip", "icp.hip", "capi.hip", "server.hip",
               "snapshot.hip", "selftest.hip", "noalloc.hip", "keymap.hip",
               "reloc.hip"]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build step failed: %s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))


def build_hip(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    objs = []

    def one(src):
        obj = os.path.join(OBJ, src.replace(".hip", ".o"))
        _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj], verbose)
        return obj

    with ThreadPoolExecutor(max_workers=min(6, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    out = os.path.join(LIB, "libfusion_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, out)
    return out


def build_cxx(verbose=False):
    """the C++ drop-in layer over the C ABI, and its test program"""
    out = os.path.join(LIB, "libfusion_amd_cxx.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-std=c++17", "-O2", "-fPIC", "-shared", os.path.join(CSRC, "compat.cpp"), "-L" + LIB,
          "-lfusion_amd", "-Wl,-rpath,$ORIGIN", "-o", tmp], verbose)
    os.replace(tmp, out)
    tdir = os.path.join(ROOT, "tests", "cxx", "_build")
    os.makedirs(tdir, exist_ok=True)
    exe = os.path.join(tdir, "test_compat")
    _run([HIPCC, "-std=c++17", "-O2", os.path.join(ROOT, "tests", "cxx", "test_compat.cpp"), "-L" + LIB,
          "-lfusion_amd_cxx", "-lfusion_amd", "-lfu_scene", "-Wl,-rpath,$ORIGIN/../../../fusion_amd/lib", "-o",
          exe + ".tmp"], verbose)
    os.replace(exe + ".tmp", exe)
    return out


def build_scene(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libfu_scene.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
          os.path.join(CSRC, "scene.cpp"), "-o", tmp], verbose)
    os.replace(tmp, out)
    return out


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "-j4"], verbose)


def build(verbose=False):
    build_scene(verbose)
    build_oracle(verbose)
    build_hip(verbose)
    build_cxx(verbose)


if __name__ == "__main__":
    build(verbose=True)
