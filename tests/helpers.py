"""Shared helpers for the parity tests (not collected by pytest)."""
import math
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def check_weights(dims):
    """w(l,j,i) = 0.1 sin(1.7 l + 0.9 j + 0.3 i) + 0.05, Fortran-numbered
    (app/hod_check_module.f90:100-107), in the ABI block layout."""
    w = []
    for l in range(1, len(dims)):
        for j in range(1, dims[l] + 1):
            for i in range(0, dims[l - 1] + 1):
                w.append(0.1 * math.sin(1.7 * (l + 1) + 0.9 * j + 0.3 * i) + 0.05)
    return np.array(w)


def read_golden(case):
    lines = open(os.path.join(GOLDEN, case, "hod_golden.dat")).read().splitlines()
    x = np.array([float(v) for v in lines[0].split()])
    vals = {}
    for s in lines[1:]:
        f = s.split()
        if len(f) == 2:
            vals[int(f[0])] = float(f[1])
    return x, np.array([vals[i] for i in sorted(vals)])


GOLDEN_CASES = {
    # case: (d0, k, closure seeds, dims)
    "hod_4d_k3": (4, 3, None, [4, 10, 10, 1]),
    "hod_4d_k7": (4, 7, None, [4, 6, 6, 1]),
    "hod_4d_k7_active": (4, 7, [[7, 0, 0, 0], [0, 0, 0, 1], [2, 2, 1, 0]], [4, 6, 6, 1]),
}


def rel_by_order(x, ref, deg):
    """{order: max|x - ref| / max|ref|} over points and slots of each order;
    x, ref: [..., na]."""
    out = {}
    for p in sorted(set(int(d) for d in deg)):
        sel = np.asarray(deg) == p
        r = np.abs(ref[..., sel]).max()
        out[p] = float(np.abs(x[..., sel] - ref[..., sel]).max() / max(r, 1e-300))
    return out


def rel(x, ref):
    return float(np.abs(np.asarray(x) - np.asarray(ref)).max() / max(np.abs(ref).max(), 1e-300))
