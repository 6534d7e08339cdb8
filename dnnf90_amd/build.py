"""Compile the gfx950 library libdnnf90_amd.so in-tree with hipcc.

    python -m dnnf90_amd.build          # build (always recompiles)

The library holds the host table builder, the HIP kernels and the C ABI of
include/dnnf90.h and include/dnnf90_amd.h.  Packed-FP32 instructions are
disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libdnnf90_amd.so")
SOURCES = ["tables.cpp", "codegen.cpp", "bell.hip", "gemm.hip", "gemm_x6.hip", "ngd.hip", "misc.hip", "kalman.hip", "engine.hip"]
ARCH = os.environ.get("DNNF90_ARCH", "gfx950")


def hipcc() -> str:
    p = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(p):
        raise RuntimeError("hipcc not found")
    return p


def build(verbose: bool = False) -> str:
    objs = []
    common = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-I", os.path.join(ROOT, "include"), "-Wall", "-Wno-unused-function"]
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [hipcc()] + common + lang + ["-c", os.path.join(CSRC, src), "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        msg = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                        if "packed-fp32-ops" not in l and l.strip())
        return src, obj, cmd, r.returncode, msg

    # one hipcc per source, a few at a time
    from concurrent.futures import ThreadPoolExecutor
    jobs = max(1, min(len(SOURCES), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 1), 8))
    with ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(compile_one, SOURCES))
    for src, obj, cmd, rc, msg in results:
        if verbose:
            print(" ".join(cmd), flush=True)
        if rc != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{msg}")
        if verbose and msg:
            print(msg)
        objs.append(obj)
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + ["-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, LIB)
    build_fortran(verbose)
    return LIB


FORTRAN_EXAMPLE = os.path.join(HERE, "fortran", "example_fd")


def build_fortran(verbose: bool = False) -> str | None:
    """The Fortran binding (fortran/dnnf90_gpu.f90) and its example program,
    compiled with amdflang and linked against the library."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        return None
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    fdir = os.path.join(HERE, "fortran")
    cmds = [[fc, "-O2", "-module-dir", bdir, "-c", os.path.join(fdir, "dnnf90_gpu.f90"),
             "-o", os.path.join(bdir, "dnnf90_gpu.o")],
            [fc, "-O2", "-I", bdir, os.path.join(fdir, "example_fd.f90"), os.path.join(bdir, "dnnf90_gpu.o"),
             "-L", HERE, "-ldnnf90_amd", "-Wl,-rpath,$ORIGIN/..", "-o", FORTRAN_EXAMPLE]]
    for cmd in cmds:
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("amdflang failed:\n" + r.stdout + r.stderr)
    return FORTRAN_EXAMPLE


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
