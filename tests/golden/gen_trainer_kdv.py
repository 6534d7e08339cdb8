"""Fixture of the trainer-bridge test: the reference's bench/kdv case
(input_nn.dat, data.dat, colloc.dat copied as data) and what the reference's
stock trainer, compiled unchanged with amdflang (oracle/Makefile `trainer`,
oracle/_ref/serial_ref), writes for it over EPOCHS epochs: the history file
and the final weights.  The frozen history column of bench/kdv itself was
written by a gfortran build, whose random_number draws other initial weights,
so the comparison is against this build of the same trainer.

    python tests/golden/gen_trainer_kdv.py     # needs /root/reference
"""
import os
import re
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "trainer_kdv")
CASE = "/root/reference/bench/kdv"
EPOCHS = 1000
INPUTS = ("input_nn.dat", "data.dat", "colloc.dat")


def with_epochs(text, epochs):
    """input_nn.dat with the Epoch key set to `epochs`."""
    return re.sub(r"(?m)^Epoch\s+\S+", f"Epoch      {epochs}", text)


def main():
    exe = os.path.join(ROOT, "oracle", "_ref", "serial_ref")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "trainer"], check=True)
    os.makedirs(OUT, exist_ok=True)
    for f in INPUTS:
        shutil.copyfile(os.path.join(CASE, f), os.path.join(OUT, f))
    with tempfile.TemporaryDirectory() as td:
        for f in INPUTS:
            shutil.copy(os.path.join(OUT, f), td)
        p = os.path.join(td, "input_nn.dat")
        text = with_epochs(open(p).read(), EPOCHS)
        os.chmod(p, 0o644)
        with open(p, "w") as f:
            f.write(text)
        r = subprocess.run([exe], cwd=td, capture_output=True, text=True)
        if r.returncode != 0:
            sys.exit("reference trainer failed:\n" + r.stdout[-2000:] + r.stderr[-2000:])
        shutil.copy(os.path.join(td, "history_ep0000000.dat"), os.path.join(OUT, f"history_ref_e{EPOCHS}.dat"))
        shutil.copy(os.path.join(td, "nn_weight.dat"), os.path.join(OUT, f"nn_weight_ref_e{EPOCHS}.dat"))
    for f in os.listdir(OUT):
        os.chmod(os.path.join(OUT, f), 0o644)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
