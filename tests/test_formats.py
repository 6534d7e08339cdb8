"""File formats shared with the reference's tools (host only, no GPU):
hod_alpha_order.dat against the reference's own files, and the Fortran
Ew.d rendering used by c_ref.dat."""
import os

import numpy as np
import pytest

from helpers import GOLDEN

D = pytest.importorskip("dnnf90_amd")


@pytest.mark.parametrize("case,d0,k,seeds", [
    ("hod_4d_k3", 4, 3, None),
])
def test_write_order_matches_reference(tmp_path, case, d0, k, seeds):
    ts = D.TableSet(d0, k, seeds)
    out = tmp_path / "hod_alpha_order.dat"
    ts.write_order(str(out))
    ref = os.path.join(GOLDEN, case, "hod_alpha_order.dat")
    assert out.read_bytes() == open(ref, "rb").read()


def test_fortran_e():
    from dnnf90_amd.formats import fortran_e
    assert fortran_e(-3.0) == "  -0.30000000000000000E+01"
    assert fortran_e(0.0) == "   0.00000000000000000E+00"
    assert fortran_e(0.1) == "   0.10000000000000001E+00"   # 0.1d0 = 0.1000000000000000055...
    assert fortran_e(1.0) == "   0.10000000000000000E+01"
    assert fortran_e(1e-300) == "   0.10000000000000000-299"
    rng = np.random.default_rng(0)
    for v in rng.standard_normal(200) * 10.0 ** rng.integers(-20, 20, 200):
        s = fortran_e(v)
        assert len(s) == 26 and float(s.replace("E", "e")) == v


def test_c_ref_points():
    from dnnf90_amd.formats import c_ref_points
    x = c_ref_points()
    assert x.shape == (25, 2)
    assert x[0, 0] == -3.0 and x[-1, 0] == 3.0
    assert x[0, 1] == 0.1 and x[9, 1] == 0.0 and x[10, 1] == 0.1
