"""Write ref_vectors.npz: outputs and gradients of the reference's own lib/
(compiled from /root/reference by `make -C oracle ref`) at the measured
workloads' network shapes, on a few points each.

    python tests/golden/gen_ref_vectors.py

Inputs are regenerated from each workload's seed (dnnf90_amd.workloads), so
only outputs are stored: T [npt][nout][na], the adjoint seeds used, and a
strided subset of the gradient (indices stored alongside).  The erf-GELU
workload is left out: the reference has no GELU activation.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    import oracle
    from dnnf90_amd.workloads import WORKLOADS
    if not oracle.RefLib.available():
        oracle.build(ref=True)
    out = {}
    for name, wl in WORKLOADS.items():
        if wl.act not in (0, 1, 2):
            continue
        npt = 8
        w, x = wl.inputs(npt)
        r = oracle.RefNet(wl.d0, wl.k, wl.seeds, wl.act, wl.dims, w)
        t = r.eval(x)
        seed = np.random.default_rng(11).standard_normal((npt, wl.dims[-1], r.na))
        g = r.grad(x, seed)
        stride = max(1, len(g) // 4096)
        idx = np.arange(0, len(g), stride)
        out[f"{name}_npt"] = np.array(npt)
        out[f"{name}_T"] = t
        out[f"{name}_seed"] = seed
        out[f"{name}_gidx"] = idx
        out[f"{name}_g"] = g[idx]
        print(name, "na", r.na, "nw", len(g), "stored", len(idx))
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)


if __name__ == "__main__":
    main()
