"""The library's host code under AddressSanitizer and UndefinedBehaviorSanitizer,
on the CPU: every source compiled as the library is, with the sanitizers on
the host side only (each -fsanitize after -Xarch_host; device code is built
plainly and never launched here), linked with the driver
tests/host/sanitize_driver.cpp, which runs the table builder for every
family shape, the kernel-source generator, the handle registries to their
limits, the weight-file writer and reader (good, missing and truncated
files) and the argument checks of the device entry points.  Any sanitizer
report aborts the driver with a non-zero status, and leak checking is on.
"""
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ROCM = "/opt/rocm"
CLANG = os.path.join(ROCM, "lib", "llvm", "bin", "clang++")
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]


def _hipcc():
    return shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")


@pytest.mark.skipif(not os.path.exists(os.path.join(ROCM, "bin", "hipcc")) or not os.path.exists(CLANG),
                    reason="needs hipcc and ROCm's clang++")
def test_host_code_under_asan_ubsan(tmp_path):
    from dnnf90_amd.build import ARCH, CSRC, SOURCES
    common = ["-O1", "-g", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
              "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
              "-I", os.path.join(ROOT, "include")] + SAN

    def compile_one(src):
        obj = str(tmp_path / (os.path.splitext(src)[0] + ".o"))
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        r = subprocess.run([_hipcc()] + common + lang + ["-c", os.path.join(CSRC, src), "-o", obj],
                           capture_output=True, text=True)
        return src, obj, r

    jobs = max(1, min(len(SOURCES), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(jobs) as ex:
        built = list(ex.map(compile_one, SOURCES))
    for src, _, r in built:
        assert r.returncode == 0, f"{src}:\n{r.stderr[-4000:]}"
    drv = str(tmp_path / "driver.o")
    r = subprocess.run([CLANG, "-O1", "-g", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"),
                        "-I", os.path.join(ROOT, "include"), "-fsanitize=address,undefined",
                        "-fno-sanitize-recover=all", "-c", os.path.join(HERE, "host", "sanitize_driver.cpp"),
                        "-o", drv], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    exe = str(tmp_path / "driver")
    r = subprocess.run([CLANG, "--hip-link", "-fsanitize=address,undefined", "-o", exe, drv]
                       + [o for _, o, _ in built] + ["-L", os.path.join(ROCM, "lib"), "-lhiprtc", "-lamdhip64",
                                                     f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    out = tmp_path / "out"
    out.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, str(out)], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout, r.stderr[-8000:])
    assert r.returncode == 0, r.stderr[-8000:]
    assert "all checks passed" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    order = (out / "hod_alpha_order.dat").read_text().splitlines()
    assert order[0].startswith("#") and len(order) == 2 + 41
