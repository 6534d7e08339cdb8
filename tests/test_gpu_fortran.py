"""A Fortran host on the GPU library: the reference's adjoint FD recipe
(tools/dbg_backward_fd.f90:10-66, :135-140) compiled with amdflang against
dnnf90_amd/fortran/dnnf90_gpu.f90 (the iso_c_binding module that mirrors
lib/net_module.f90 and lib/train_module.f90's routine names)."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def test_fortran_fd_recipe():
    from dnnf90_amd.build import FORTRAN_EXAMPLE
    if not os.path.exists(FORTRAN_EXAMPLE):
        pytest.skip("Fortran example not built (amdflang missing at build time)")
    print("running", FORTRAN_EXAMPLE, flush=True)
    r = subprocess.run([FORTRAN_EXAMPLE], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert "all checks passed" in r.stdout
