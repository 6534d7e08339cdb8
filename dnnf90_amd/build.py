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


FORTRAN_DIR = os.path.join(HERE, "fortran")
# the Fortran modules with the reference's names, in dependency order
FORTRAN_MODULES = ["dnnf90_capi.f90", "multi_index_bell_module.f90", "net_module.f90", "train_module.f90"]
FORTRAN_LIB = os.path.join(HERE, "_build", "libdnnf90_fortran.a")
# reference host programs compiled unchanged against those modules (from
# /root/reference/tools, where that tree is present); the binaries go to
# fortran/_ref/ and are never committed
REF_TOOLS = ["dbg_backward_fd", "example_multiout", "example_multiout_adj", "example_batch_act",
             "example_act_roundtrip"]
# ...and the stock trainer (app/main.f90) with fortran/trainer/lib_net_module.f90
# as its network bridge: REF_TOOL_DIR/serial_gpu
REF_TOOL_DIR = os.path.join(FORTRAN_DIR, "_ref")
REFERENCE = os.environ.get("DNNF90_REFERENCE", "/root/reference")


def build_fortran(verbose: bool = False) -> str | None:
    """The Fortran modules multi_index_bell_module, net_module and
    train_module (over the C ABI, fortran/*.f90), compiled with amdflang into
    _build/ (objects, .mod files, libdnnf90_fortran.a); then, where the
    reference tree exists, its tools/*.f90 host programs listed in REF_TOOLS,
    compiled as they are against those modules and linked with the library."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        return None
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    objs = []
    for src in FORTRAN_MODULES:
        obj = os.path.join(bdir, os.path.splitext(src)[0] + ".o")
        cmd = [fc, "-O2", "-fPIC", "-module-dir", bdir, "-I", bdir, "-c", os.path.join(FORTRAN_DIR, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed on {src}:\n" + r.stdout + r.stderr)
        objs.append(obj)
    if os.path.exists(FORTRAN_LIB):
        os.remove(FORTRAN_LIB)
    r = subprocess.run(["ar", "rcs", FORTRAN_LIB] + objs, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("ar failed:\n" + r.stdout + r.stderr)
    tools = os.path.join(REFERENCE, "tools")
    if not os.path.isdir(tools):
        return FORTRAN_LIB
    os.makedirs(REF_TOOL_DIR, exist_ok=True)
    for name in REF_TOOLS:
        src = os.path.join(tools, name + ".f90")
        if not os.path.exists(src):
            continue
        out = os.path.join(REF_TOOL_DIR, name)
        cmd = [fc, "-O2", "-I", bdir, src, FORTRAN_LIB, "-L", HERE, "-ldnnf90_amd",
               "-Wl,-rpath,$ORIGIN/../..", "-o", out]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True, cwd=bdir)
        if r.returncode != 0:
            raise RuntimeError(f"amdflang failed on the reference's {name}.f90:\n" + r.stdout + r.stderr)
    # the stock trainer with the GPU bridge (fortran/trainer/Makefile)
    if os.path.isdir(os.path.join(REFERENCE, "app")):
        r = subprocess.run(["make", "-s", "-B", "-C", os.path.join(FORTRAN_DIR, "trainer"), "gpu",
                            f"REF={REFERENCE}"], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("trainer bridge build failed:\n" + r.stdout[-4000:] + r.stderr[-4000:])
    return FORTRAN_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
