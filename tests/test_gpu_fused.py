"""The fused layer kernels of the FP32 collocation step against the oracle.

An FP32 work space whose network has at least two hidden layers, all 64
wide, and one output row takes dnnf90_residual_step_batch on the fused
kernels (point-blocked planes; each hidden layer's linear map inside its
Faa di Bruno forward, on BF16 three-way split products, which also writes
h; one adjoint kernel a layer reading h and forming the pair sums, the
weight gradient and T-bar of the layer below).  Work spaces with per-layer
recompute keep the layer kernels.  The step's
gradient and sum of squared residuals are checked against the oracle's FP64
pinn_step (the restatement of app/pinn_module.f90:32-65 over
lib/train_module.f90:241-450, pinned bitwise to the reference), at the FP32
modes' tolerance, for the headline family, a three-hidden-layer network, a
network of two hidden layers (the first and last hidden layer kernels meet)
and another activation; the layer-by-layer FP32 kernels take the same step
for comparison.
"""
import numpy as np
import pytest

from helpers import rel, rel_by_order

pytestmark = pytest.mark.gpu

TOL32 = 1.5e-6     # gradient, max |diff| / max |oracle| (the FP32 modes' bound)
TOL_RSQ = 2e-6     # sum of R^2, relative
# the step's output row through derivative order 4, per order: max |diff| /
# max |oracle| (the mixed modes' bar; the sin network's FP32 storage error,
# in the layer kernels as much as in the fused ones, is near 3e-6)
TOL_T = {"closure4_k7": 1e-6, "ngd2_k4": 1e-6, "dense3_k3_h2": 1e-6, "sin2_k4": 5e-6}


@pytest.fixture(scope="module")
def D():
    import dnnf90_amd
    return dnnf90_amd


def _cases():
    from dnnf90_amd.workloads import WORKLOADS, Workload
    return {
        "closure4_k7": (WORKLOADS["closure4_k7"], 1000),
        "ngd2_k4": (WORKLOADS["ngd2_k4"], 517),
        # two hidden layers: the last hidden layer's kernels feed the first's
        "dense3_k3_h2": (Workload("dense3_k3_h2", 3, 3, None, [3, 64, 64, 1], 0, 300, [(-1.0, 1.0)] * 3, 11,
                                  [(None, "all", False)]), 300),
        # sin activation, an odd point count (a partly filled block of 4 and of 64)
        "sin2_k4": (Workload("sin2_k4", 2, 4, [[4, 0], [0, 2], [1, 0]], [2, 64, 64, 64, 1], 1, 201,
                             [(-1.0, 1.0)] * 2, 12, [(1.0, (4, 0), False), (1.0, (0, 2), False),
                                                     (2.0, (1, 0), True)]), 201),
    }


def _step(D, net, x, terms, npt, fused, poison=False, recompute=False):
    from dnnf90_amd._lib import lib
    prev = lib().dnnf90_set_fused_layers(1 if fused else 0)
    prev_p = lib().dnnf90_set_poison(1 if poison else 0)
    try:
        bw = D.BatchWork(net, npt, D.PREC_FP32, budget=1 if recompute else 0)
        assert bw.info()["recompute"] == recompute
        g = D.Grad(net)
        rsq = bw.residual_step(x, D.Residual(terms), 2.0 / npt, g)
        out = g.get().copy()
        _step.tout = bw.step_output()
        bw.free()
        g.free()
    finally:
        lib().dnnf90_set_fused_layers(prev)
        lib().dnnf90_set_poison(prev_p)
    return out, rsq


@pytest.mark.parametrize("name", ["closure4_k7", "ngd2_k4", "dense3_k3_h2", "sin2_k4"])
def test_fused_step_vs_oracle(D, oracle_mod, name):
    from dnnf90_amd._lib import lib
    wl, npt = _cases()[name]
    print(f"case {name} dims {wl.dims} npt {npt}", flush=True)
    ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
    net = D.Net(ts, wl.dims)
    w, x = wl.inputs(npt)
    net.weights = w
    info = np.zeros(14, np.int32)
    assert lib().dnnf90_net_fused_info(net.id, info.ctypes.data) == 0
    print("fused info", info.tolist(), flush=True)
    assert info[0] == 1, "the fused kernels must take this step"
    assert info[13] > 0 and max(info[7:13]) <= 256
    o = oracle_mod.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
    terms = wl.terms(o.alpha())
    gr, rsq_r, _ = o.pinn_step(wl.dims, w, x, terms, 2.0 / npt)
    gf, rsq_f = _step(D, net, x, terms, npt, True)
    tf = _step.tout
    gl, rsq_l = _step(D, net, x, terms, npt, False)
    tl = _step.tout
    ef, el = rel(gf, gr), rel(gl, gr)
    # the output row the step formed (the fused last hidden layer's kernel
    # writes it) against the oracle's forward, order by order
    te = o.eval(wl.dims, w, x)
    deg = o.deg()
    etf = {p: e for p, e in rel_by_order(tf, te, deg).items() if p <= 4}
    etl = {p: e for p, e in rel_by_order(tl, te, deg).items() if p <= 4}
    print(f"{name}: fused grad err {ef:.3e} rsq err {abs(rsq_f - rsq_r) / abs(rsq_r):.3e}; "
          f"layer kernels grad err {el:.3e}; fused vs layer {rel(gf, gl):.3e}; "
          f"output row through order 4 fused {max(etf.values()):.3e} layer {max(etl.values()):.3e}", flush=True)
    assert ef <= TOL32 and abs(rsq_f - rsq_r) <= TOL_RSQ * abs(rsq_r)
    assert el <= TOL32
    assert max(etf.values()) <= TOL_T[name] and max(etl.values()) <= TOL_T[name]
    # the same step again, and with every plane and partial poisoned: the same bits
    gf2, rsq_f2 = _step(D, net, x, terms, npt, True)
    assert np.array_equal(gf, gf2) and rsq_f == rsq_f2
    gp, rsq_p = _step(D, net, x, terms, npt, True, poison=True)
    assert np.array_equal(gf, gp) and rsq_f == rsq_p
    # per-layer recompute takes the layer kernels: their bits
    gr2, rsq_r2 = _step(D, net, x, terms, npt, True, recompute=True)
    assert np.array_equal(gl, gr2) and rsq_l == rsq_r2
    net.free()


def test_fused_not_taken(D):
    """Widths other than 64 and FP64 work spaces keep the layer kernels."""
    from dnnf90_amd._lib import lib
    from dnnf90_amd.workloads import WORKLOADS
    for name in ["kdv", "gelu3_k4"]:
        wl = WORKLOADS[name]
        ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
        net = D.Net(ts, wl.dims)
        info = np.ones(14, np.int32)
        assert lib().dnnf90_net_fused_info(net.id, info.ctypes.data) == 0
        assert info[0] == 0, name
        net.free()


@pytest.mark.parametrize("name", ["closure4_k7", "ngd2_k4"])
def test_fused_full_size(D, name):
    """The benchmark sizes (closure4_k7 N = 16 384: 4 096 point blocks, eight
    a workgroup in the adjoint; ngd2_k4 N = 4 096): the fused FP32 step
    against the FP64 step (the FP64 path matches the oracle to 1e-12 at these
    sizes, tests/test_gpu_fullsize.py) at the FP32 modes' tolerance, against
    the layer-by-layer FP32 kernels, and additive over four quarter batches
    (only the order of the point sums differs)."""
    from dnnf90_amd.workloads import WORKLOADS
    import oracle
    wl = WORKLOADS[name]
    N = wl.npt
    print(f"case {name} N {N}", flush=True)
    ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
    net = D.Net(ts, wl.dims)
    w, x = wl.inputs()
    net.weights = w
    o = oracle.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
    terms = wl.terms(o.alpha())
    bw = D.BatchWork(net, N, D.PREC_FP64)
    g64 = D.Grad(net)
    rsq64 = bw.residual_step(x, D.Residual(terms), 2.0 / N, g64)
    gd = g64.get().copy()
    td = bw.step_output()
    bw.free()
    g64.free()
    gf, rsq_f = _step(D, net, x, terms, N, True)
    tf = _step.tout
    gl, rsq_l = _step(D, net, x, terms, N, False)
    quarters = np.zeros_like(gf)
    rq = 0.0
    for k in range(4):
        sl = slice(k * N // 4, (k + 1) * N // 4)
        gq, r = _step(D, net, x[sl], terms, N // 4, True)
        quarters += gq * 0.25   # each quarter's step is scaled by 2 / (N / 4)
        rq += r
    et = {p: e for p, e in rel_by_order(tf, td, o.deg()).items() if p <= 4}
    print(f"{name}: fused vs FP64 grad {rel(gf, gd):.3e} rsq {abs(rsq_f - rsq64) / rsq64:.3e} "
          f"output row through order 4 {max(et.values()):.3e}; "
          f"layer kernels vs FP64 {rel(gl, gd):.3e}; quarters {rel(quarters, gf):.3e}", flush=True)
    assert rel(gf, gd) <= TOL32 and abs(rsq_f - rsq64) <= TOL_RSQ * rsq64
    assert max(et.values()) <= TOL_T[name]
    assert rel(gl, gd) <= TOL32
    assert rel(quarters, gf) <= 1e-6 and abs(rq - rsq_f) <= 1e-6 * rsq_f
    net.free()
