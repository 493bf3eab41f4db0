"""Build libvpmx.so (HIP, gfx950) and the Fortran drop-in module in-tree.

    python -m vpm_free_amd.build            # HIP library + Fortran module
    python -m vpm_free_amd.build --hip-only

Every HIP translation unit is compiled with hipcc for gfx950 only, in parallel, then
linked into vpm_free_amd/lib/libvpmx.so.  The Fortran module (vpm_lib, the reference's
entry points over the C ABI) is built with flang into vpm_free_amd/lib/libvpm_lib.so.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "lib")
OBJ = os.path.join(HERE, "_build")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
ARCH = "gfx950"

HIP_SOURCES = ["vpmx_api.hip", "vpmx_batch.hip", "tune.hip", "k_adjoint.hip", "k_particles.hip", "k_poisson.hip", "k_boundary.hip", "k_bspec.hip", "k_dgemm.hip", "k_grid.hip",
               "k_tree.hip", "k_panel.hip"]
CXX_SOURCES = ["geometry.cpp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Wno-unused-variable", "-Wno-unused-result",
    # keep packed FP32 instructions out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def _compile_one(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = ["g++"] + CXX_FLAGS + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


def build_hip(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile_one, HIP_SOURCES + CXX_SOURCES))
    out = os.path.join(LIB, "libvpmx.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, out)
    return out


def build_fortran():
    """vpm_lib.f90 -> libvpm_lib.so (links libvpmx.so)."""
    src = os.path.join(FSRC, "vpm_lib.f90")
    if not os.path.exists(src) or not os.path.exists(FLANG):
        return None
    mod = os.path.join(LIB, "mod")
    os.makedirs(mod, exist_ok=True)
    out = os.path.join(LIB, "libvpm_lib.so")
    _run([FLANG, "-O2", "-fPIC", "-shared", "-module-dir", mod, "-o", out + ".tmp", src,
          "-L" + LIB, "-lvpmx", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    # example application (the reference driver's loop) linked against the drop-in
    drv = os.path.join(LIB, "ring_driver")
    _run([FLANG, "-O2", "-I", mod, "-o", drv + ".tmp", os.path.join(FSRC, "ring_driver.f90"),
          "-L" + LIB, "-lvpm_lib", "-lvpmx", "-Wl,-rpath,$ORIGIN"])
    os.replace(drv + ".tmp", drv)
    return out


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    lib = build_hip()
    print("built", lib)
    if "--hip-only" not in argv:
        f = build_fortran()
        if f:
            print("built", f)
    if os.path.isdir(OBJ):
        shutil.rmtree(OBJ, ignore_errors=True)


if __name__ == "__main__":
    main()
