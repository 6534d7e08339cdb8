"""Debug mode for reads of unwritten memory (dnnf90_set_poison): a batched
work space created in it fills its planes, operand scratch and gradient
partials with NaN before every forward, so any value a step reads without
having written it in that step turns the result into NaN.  Every step
below must give the same bits with and without the poison: forward,
collocation step (loss and gradient) and Gauss-Newton rows, in FP64 and
BF16x6, with every plane kept and with per-layer recompute, through the
generated kernels (with the edge and fused kernels) and the table-driven
ones, at a point count that leaves a partly filled 64-point block."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [("closure4", 4, 7, [[0, 0, 0, 1], [1, 0, 0, 0], [7, 0, 0, 0], [3, 2, 2, 0]], [4, 32, 64, 64, 1]),
         ("kdv", 2, 3, [[0, 1], [1, 0], [3, 0]], [2, 32, 32, 32, 1]),
         ("dense_two_out", 3, 3, None, [3, 24, 40, 2])]


@pytest.fixture(scope="module")
def D():
    import dnnf90_amd as D
    return D


def _step(D, net, x, terms, prec, budget, poison):
    import torch
    from dnnf90_amd._lib import lib
    prev = lib().dnnf90_set_poison(1 if poison else 0)
    try:
        bw = D.BatchWork(net, len(x), prec, budget=budget)
    finally:
        lib().dnnf90_set_poison(prev)
    g = D.Grad(net)
    xd = torch.from_numpy(x).cuda()
    tout = bw.forward(xd)
    bw.sync()   # device outputs are written on the work space's own stream
    t = np.array(tout.cpu().numpy(), copy=True)
    rsq = bw.residual_step(xd, D.Residual(terms), 1.0 / len(x), g)
    grad = np.array(g.get(), copy=True)
    rows = torch.zeros((net.nw, (len(x) + 63) // 64 * 64), dtype=torch.float32, device="cuda")
    bw.residual_rows(xd, D.Residual(terms), rows, rows.shape[1])
    torch.cuda.synchronize()
    r = rows.cpu().numpy()
    bw.free(); g.free()
    return t, rsq, grad, r


@pytest.mark.parametrize("codegen", [1, 0], ids=["generated", "table"])
@pytest.mark.parametrize("budget", [0, 1], ids=["keep", "recompute"])
@pytest.mark.parametrize("prec", [0, 3], ids=["fp64", "bf16x6"])
@pytest.mark.parametrize("case", CASES, ids=[c[0] for c in CASES])
def test_poisoned_planes_change_nothing(D, case, prec, budget, codegen):
    from dnnf90_amd._lib import lib
    name, d0, k, seeds, dims = case
    if codegen == 0 and prec == 3:
        pytest.skip("the table-driven kernels are checked in FP64")
    print(f"poison: {name} dims={dims} prec={prec} budget={budget} codegen={codegen}", flush=True)
    ts = D.TableSet(d0, k, seeds)
    net = D.Net(ts, dims)
    rng = np.random.default_rng(31)
    w = rng.standard_normal(net.nw) * 0.4
    net.weights = w
    x = rng.uniform(-1, 1, (150, d0))
    # R = T_0 + T_1 + T_2 + T_3 + 0.5 T_0 T_1 on output 0 (slots 0-based)
    terms = [(1.0, i, False) for i in range(min(4, ts.na))] + [(0.5, 1, True)]
    prev = lib().dnnf90_set_codegen(codegen)
    try:
        ref = _step(D, net, x, terms, prec, budget, False)
        got = _step(D, net, x, terms, prec, budget, True)
    finally:
        lib().dnnf90_set_codegen(prev)
    for a, b in zip(ref, got):
        a, b = np.asarray(a), np.asarray(b)
        assert np.all(np.isfinite(b))
        assert np.array_equal(a, b)
    net.free(); ts.free()
