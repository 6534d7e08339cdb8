"""The trainer bridge: the reference's stock trainer (app/main.f90 and the
modules of its own build) with dnnf90_amd/fortran/trainer/lib_net_module.f90
in place of app/lib_net_module.f90, so that every per-point forward, seeded
adjoint and value-batch gradient the trainer asks for (app/sgd_batch_module.f90:
593-757) runs on the GPU library.  It runs the reference's bench/kdv case
(KdV PINN, [2,16,16,1] tanh, 100 data + 200 collocation points, Adam) for
1000 epochs and must reproduce the reference trainer's own run of the same
build (tests/golden/trainer_kdv, written by tests/golden/gen_trainer_kdv.py):
the HOD self-check at start-up passes, every history row (epoch, cost_train,
cost_val, ...; 5 printed digits) agrees to 1e-4 relative (measured: equal),
and the final weights agree to 1e-12 relative after the 1000 Adam steps
(measured 1.2e-15)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = os.path.join(HERE, "golden", "trainer_kdv")
EPOCHS = 1000


def _history(path):
    rows = [l.split() for l in open(path) if l.strip() and not l.startswith("#")]
    return np.array([[float(v) for v in r] for r in rows])


def _weights(path):
    lines = open(path).read().split("\n")
    nl = int(lines[3])
    vals = []
    for l in lines[4 + nl:]:
        if l.strip() and not l.strip().startswith("#"):
            vals += [float(v) for v in l.split()]
    return np.array(vals)


def test_stock_trainer_on_gpu(tmp_path):
    from dnnf90_amd.build import REF_TOOL_DIR
    exe = os.path.join(REF_TOOL_DIR, "serial_gpu")
    if not os.path.exists(exe):
        pytest.skip("trainer bridge not built (needs amdflang and the reference tree at build time)")
    for f in ("input_nn.dat", "data.dat", "colloc.dat"):
        shutil.copyfile(os.path.join(FIX, f), tmp_path / f)
    p = tmp_path / "input_nn.dat"
    p.write_text(re.sub(r"(?m)^Epoch\s+\S+", f"Epoch      {EPOCHS}", p.read_text()))
    print("running the stock trainer on the GPU bridge, bench/kdv,", EPOCHS, "epochs", flush=True)
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=900)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0
    assert "HOD self-check: ALL PASSED" in r.stdout
    h = _history(tmp_path / "history_ep0000000.dat")
    href = _history(os.path.join(FIX, f"history_ref_e{EPOCHS}.dat"))
    assert h.shape == href.shape
    assert np.array_equal(h[:, :2], href[:, :2])          # epochs, patience
    rel = np.abs(h[:, 2:] - href[:, 2:]) / np.maximum(np.abs(href[:, 2:]), 1e-300)
    print("history: max rel diff", rel.max(), "rows", h.shape[0])
    assert rel.max() <= 1e-4
    w = _weights(tmp_path / "nn_weight.dat")
    wref = _weights(os.path.join(FIX, f"nn_weight_ref_e{EPOCHS}.dat"))
    ew = np.abs(w - wref).max() / np.abs(wref).max()
    print("final weights: max rel diff", ew)
    assert ew <= 1e-12
