"""The measured workloads, with their synthetic inputs.

Each workload fixes a table family (D0, K, closure seeds), an MLP shape, an
activation, a residual (the loss seed) and a point box.  Inputs come from
numpy's PCG64 generator: Xavier-normal weights, sigma = sqrt(2/(fan_in +
fan_out)), zero biases (the reference's Glorot-normal rule,
app/init_weight_module.f90:87, 245-265), drawn layer by layer, then the
points, uniform on the box, from the same generator.

Residuals are lists of (coeff, slot, nonlin): R = sum coeff * [u *] T_slot,
with slots 0-based in the family's carried order; the loss is the mean
square of R, so the seed is (2/N) R dR/dT (app/pinn_module.f90:50-65 with
fac = 2/N).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

ACT_TANH, ACT_GELU = 0, 5


@dataclass
class Workload:
    name: str
    d0: int
    k: int
    seeds: list | None          # closure seeds, None = dense
    dims: list
    act: int
    npt: int
    box: list                   # per-input (lo, hi)
    seed: int
    residual: list = field(default_factory=list)   # symbolic terms, resolved by terms()
    note: str = ""
    kf: tuple | None = None     # Kalman filter (p0, lam_ini, lam0, decoupled): npt observations in order

    @property
    def nweights(self):
        return int(sum(self.dims[l] * (self.dims[l - 1] + 1) for l in range(1, len(self.dims))))

    def weights(self, rng=None):
        rng = np.random.default_rng(self.seed) if rng is None else rng
        out = []
        for l in range(1, len(self.dims)):
            fi, fo = self.dims[l - 1], self.dims[l]
            sig = np.sqrt(2.0 / (fi + fo))
            W = rng.normal(0.0, sig, size=(fo, fi))
            blk = np.zeros((fo, fi + 1))
            blk[:, 1:] = W
            out.append(blk.reshape(-1))
        return np.concatenate(out)

    def inputs(self, npt=None):
        """(weights, points[npt][d0]) from the workload's generator."""
        rng = np.random.default_rng(self.seed)
        w = self.weights(rng)
        n = self.npt if npt is None else npt
        lo = np.array([b[0] for b in self.box])
        hi = np.array([b[1] for b in self.box])
        x = lo + (hi - lo) * rng.random((n, self.d0))
        return w, x

    def terms(self, alpha):
        """Residual terms with 0-based slots for the carried order `alpha`."""
        index = {tuple(int(v) for v in a): i for i, a in enumerate(alpha)}
        out = []
        for c, mi, nonlin in self.residual:
            if mi == "all":
                cs = np.random.default_rng(self.seed + 1).standard_normal(len(alpha))
                out += [(float(cs[i]), i, False) for i in range(len(alpha))]
            else:
                out.append((float(c), index[tuple(mi)], bool(nonlin)))
        return out


WORKLOADS = {
    # [2,32,32,32,1] tanh, closure{t, x, xxx} over (x, t); KdV residual
    "kdv": Workload("kdv", 2, 3, [[0, 1], [1, 0], [3, 0]], [2, 32, 32, 32, 1], ACT_TANH, 1024,
                    [(-10.0, 10.0), (0.0, 1.0)], 0,
                    [(1.0, (0, 1), False), (6.0, (1, 0), True), (1.0, (3, 0), False)],
                    "u_t + 6 u u_x + u_xxx"),
    # [4,128x4,1] tanh, dense |a| <= 3 (35 slots); random linear combination
    "dense4_k3": Workload("dense4_k3", 4, 3, None, [4, 128, 128, 128, 128, 1], ACT_TANH, 65536,
                          [(-1.0, 1.0)] * 4, 1, [(None, "all", False)],
                          "sum_a c_a T_a, c ~ N(0,1) from seed 2"),
    # [4,64x4,1] tanh, closure{t, x, x^7, x^3 y^2 z^2} over (x, y, z, t)
    "closure4_k7": Workload("closure4_k7", 4, 7, [[0, 0, 0, 1], [1, 0, 0, 0], [7, 0, 0, 0], [3, 2, 2, 0]],
                            [4, 64, 64, 64, 64, 1], ACT_TANH, 16384, [(-1.0, 1.0)] * 4, 3,
                            [(1.0, (0, 0, 0, 1), False), (1.0, (1, 0, 0, 0), True),
                             (1.0, (7, 0, 0, 0), False), (1.0, (3, 2, 2, 0), False)],
                            "u_t + u u_x + d_x^7 u + d_x^3 d_y^2 d_z^2 u"),
    # [3,256x4,1] erf-GELU, dense |a| <= 4 (35 slots); random linear combination
    "gelu3_k4": Workload("gelu3_k4", 3, 4, None, [3, 256, 256, 256, 256, 1], ACT_GELU, 131072,
                         [(-2.0, 2.0)] * 3, 4, [(None, "all", False)],
                         "sum_a c_a T_a, c ~ N(0,1) from seed 5"),
    # [2,64x3,1] tanh, closure{x^4, y^2, x}; NGD rows of u_xxxx + u_yy + u_x
    "ngd2_k4": Workload("ngd2_k4", 2, 4, [[4, 0], [0, 2], [1, 0]], [2, 64, 64, 64, 1], ACT_TANH, 4096,
                        [(-1.0, 1.0)] * 2, 5,
                        [(1.0, (4, 0), False), (1.0, (0, 2), False), (1.0, (1, 0), False)],
                        "u_xxxx + u_yy + u_x"),
    # the extended Kalman filter of bench/opt_kalman (input_nn.dat there:
    # [2,16,16,1] tanh, KdV residual with A = 3, GD_param 1e-3 0.995 0.9999,
    # dense filter), one collocation observation per point
    "kf_kdv": Workload("kf_kdv", 2, 3, [[0, 1], [1, 0], [3, 0]], [2, 16, 16, 1], ACT_TANH, 2000,
                       [(-8.0, 8.0), (0.0, 2.0)], 6,
                       [(1.0, (0, 1), False), (3.0, (1, 0), True), (1.0, (3, 0), False)],
                       "u_t + 3 u u_x + u_xxx", kf=(1e-3, 0.995, 0.9999, False)),
    # the dense filter on the closure4_k7 network (nw = 12865, covariance
    # 0.66 GB as a lower triangle), collocation observations of its residual
    "kf_closure4_k7": Workload("kf_closure4_k7", 4, 7, [[0, 0, 0, 1], [1, 0, 0, 0], [7, 0, 0, 0], [3, 2, 2, 0]],
                               [4, 64, 64, 64, 64, 1], ACT_TANH, 200, [(-1.0, 1.0)] * 4, 7,
                               [(1.0, (0, 0, 0, 1), False), (1.0, (1, 0, 0, 0), True),
                                (1.0, (7, 0, 0, 0), False), (1.0, (3, 2, 2, 0), False)],
                               "u_t + u u_x + d_x^7 u + d_x^3 d_y^2 d_z^2 u", kf=(1e-3, 0.995, 0.9999, False)),
}
