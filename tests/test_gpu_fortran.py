"""Fortran hosts on the GPU library: the reference's own host programs
(tools/*.f90, compiled unchanged with amdflang against this repository's
multi_index_bell_module, net_module and train_module, which keep the
reference's names over the C ABI; dnnf90_amd/build.py builds them into
dnnf90_amd/fortran/_ref/ where the reference tree is present) run to their
own pass criteria:

  * dbg_backward_fd: point adjoint vs central differences, the value-only
    batch path vs the per-point engine (1e-12) and vs central differences,
    net_eval_hod vs the training forward (1e-12); "ALL PASSED", exit 0;
  * example_multiout: every output of a four-output network, net_eval_hod_multi
    against net_eval_hod and first and second input derivatives against
    central differences;
  * example_multiout_adj: net_backward_point_multi against central
    differences of a loss over every output and slot;
  * example_batch_act: net_grad_batch against per-point gradients for every
    activation of the reference (1e-8), "passed";
  * example_act_roundtrip: net_save / net_load for every activation, "passed".
"""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def _run(name):
    from dnnf90_amd.build import REF_TOOL_DIR
    exe = os.path.join(REF_TOOL_DIR, name)
    if not os.path.exists(exe):
        pytest.skip(f"{name} not built (needs amdflang and the reference tree at build time)")
    print("running", exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    return r.stdout


def _values(text, label):
    return [float(v) for v in re.findall(re.escape(label) + r"\s*([-+0-9.Ee]+)", text)]


def test_dbg_backward_fd():
    out = _run("dbg_backward_fd")
    assert "gradient checks: ALL PASSED" in out
    assert _values(out, "point adjoint vs FD          :")[0] <= 1e-6
    assert _values(out, "batch vs per-point engine    :")[0] <= 1e-12
    assert _values(out, "net_eval_hod vs forward      :")[0] <= 1e-12


def test_example_multiout():
    out = _run("example_multiout")
    assert _values(out, "component 1 vs net_eval_hod, max |diff| =")[0] <= 1e-13
    fd = _values(out, "outputs vs FD, max rel =") + _values(out, "pure second derivatives vs FD, max rel =")
    assert len(fd) == 2 and max(fd) <= 1e-5


def test_example_multiout_adj():
    out = _run("example_multiout_adj")
    v = _values(out, "weights, max rel =")
    assert len(v) == 1 and v[0] <= 1e-6


def test_example_batch_act():
    out = _run("example_batch_act")
    assert "passed" in out and "FAILED" not in out
    assert len(_values(out, "batch vs per-point, max rel =")) == 5


def test_example_act_roundtrip():
    out = _run("example_act_roundtrip")
    assert "passed" in out and "FAILED" not in out
