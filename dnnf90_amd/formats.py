"""File formats shared with the reference's tools.

* hod_alpha_order.dat -- the slot order of a table family
  (lib/multi_index_bell_module.f90:451-471): `TableSet.write_order`, which
  calls dnnf90_tabset_write_order.
* nn_weight.dat -- weights (lib/net_module.f90:146-243): `Net.load` /
  dnnf90_net_load and dnnf90_net_save.
* c_ref.dat -- value and derivatives of a network at 25 fixed points, the
  file tools/gen_c_ref.f90 writes for the cross-language check of
  tools/example_c.c: `write_c_ref` here.
"""
from __future__ import annotations

import numpy as np


def fortran_e(v: float, w: int = 26, d: int = 17) -> str:
    """v in Fortran's Ew.d edit descriptor: 0.<d digits>E+ee, right-justified
    in w columns (E is dropped for three-digit exponents, as Fortran does)."""
    v = float(v)
    if v == 0.0:
        body = "0." + "0" * d + "E+00"
    else:
        s = format(abs(v), f".{d - 1}e")          # d significant digits, rounded once
        mant, ex = s.split("e")
        digits = mant.replace(".", "")
        e = int(ex) + 1                           # 0.ddd x 10^e
        body = "0." + digits + (f"E{e:+03d}" if abs(e) <= 99 else f"{e:+04d}")
        if v < 0:
            body = "-" + body
    return body.rjust(w)


def c_ref_points(npt: int = 25) -> np.ndarray:
    """The points of tools/gen_c_ref.f90: x1 = -3 + 6 (ip-1)/(npt-1),
    x2 = 0.1 mod(ip, 10), ip = 1..npt."""
    ip = np.arange(1, npt + 1)
    return np.stack([-3.0 + 6.0 * (ip - 1) / (npt - 1), 0.1 * (ip % 10)], axis=1)


def write_c_ref(net, path, npt: int = 25) -> np.ndarray:
    """c_ref.dat of a two-input network (tools/gen_c_ref.f90): the point
    count, then per point x(1:2) and the carried derivatives t(1:nderiv) of
    the first output in e26.17, through the per-point ABI (dnnf90_eval).
    Returns the values written [npt][nderiv]."""
    from .api import Work
    if net.dims[0] != 2:
        raise ValueError("c_ref.dat is written for two-input networks")
    x = c_ref_points(npt)
    wk = Work(net)
    t = np.array([wk.eval(xn) for xn in x])   # per point, as tools/gen_c_ref.f90 calls net_eval_hod
    wk.free()
    with open(path, "w") as f:
        f.write(f"{npt}\n")
        for n in range(npt):
            f.write("".join(fortran_e(v) for v in list(x[n]) + list(t[n])) + "\n")
    return t


def read_c_ref(path):
    """(x [npt][2], t [npt][nderiv]) of a c_ref.dat."""
    with open(path) as f:
        npt = int(f.readline())
        rows = []
        for _ in range(npt):
            line = f.readline().rstrip("\n")
            vals = [float(line[i:i + 26].replace("E", "e") if "E" in line[i:i + 26] else
                          _fix_exp(line[i:i + 26])) for i in range(0, len(line), 26)]
            rows.append(vals)
    a = np.array(rows)
    return a[:, :2], a[:, 2:]


def _fix_exp(s: str) -> str:
    # three-digit exponent without E: 0.1234+100 -> 0.1234e+100
    s = s.strip()
    k = max(s.rfind("+"), s.rfind("-"))
    return s[:k] + "e" + s[k:] if k > 0 else s
