"""Parity at the benchmark sizes (dense4_k3: N = 65 536, |A| = 35, width 128;
closure4_k7: N = 16 384, |A| = 41, width 64; gelu3_k4: N = 131 072, |A| = 35,
width 256, erf-GELU; kdv: N = 1 024, |A| = 5), through properties that hold
at any size, so that the oracle (the C restatement of
lib/train_module.f90:241-450) only runs on samples:

  * sampled points of the full FP64 batch against the oracle, 1e-12 relative
    per derivative order;
  * batch invariance: the sampled points as a batch of their own give the
    same bits (the forward has no cross-point sums; every point's column
    goes through the same kernels in the same order);
  * additivity of the gradient: the FP64 gradient of the collocation loss
    over the full batch equals the sum of the gradients of four disjoint
    quarter batches to 1e-12 relative (only the order of the point sums
    differs);
  * the mixed mode (BF16X6) against FP64 at full size: at most 1e-6 relative
    per derivative order through order 4 (the mixed mode's accuracy target),
    and the gradient within 1e-6 relative.
"""
import numpy as np
import pytest

from helpers import rel, rel_by_order

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL_MIXED = 1e-6


@pytest.fixture(scope="module")
def D():
    import dnnf90_amd
    return dnnf90_amd


@pytest.mark.parametrize("name", ["dense4_k3", "closure4_k7", "gelu3_k4", "kdv"])
def test_full_size(D, oracle_mod, name):
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[name]
    N = wl.npt
    ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
    net = D.Net(ts, wl.dims)
    w, x = wl.inputs()
    net.weights = w
    o = oracle_mod.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
    deg = o.deg()
    res = D.Residual(wl.terms(o.alpha()))

    bw = D.BatchWork(net, N, D.PREC_FP64)
    t = bw.forward(x)
    idx = np.sort(np.random.default_rng(17).choice(N, 96, replace=False))
    e_or = rel_by_order(t[idx], o.eval(wl.dims, w, x[idx]), deg)
    print(name, "N", N, "sampled points vs oracle, by order:", e_or)
    assert max(e_or.values()) <= TOL64

    bs = D.BatchWork(net, len(idx), D.PREC_FP64)
    ts_ = bs.forward(x[idx])
    print(name, "batch invariance: max |diff|", np.abs(ts_ - t[idx]).max())
    assert np.array_equal(ts_, t[idx])
    bs.free()

    g = D.Grad(net)
    bw.residual_step(x, res, 1.0, g, want_rsq=False)
    g_full = g.get()
    q = N // 4
    bq = D.BatchWork(net, q, D.PREC_FP64)
    gq = D.Grad(net)
    for k in range(4):
        bq.residual_step(x[k * q:(k + 1) * q], res, 1.0, gq, want_rsq=False)
    e_add = rel(gq.get(), g_full)
    print(name, "gradient: full batch vs four quarters", e_add)
    assert e_add <= TOL64
    bq.free(); gq.free()

    b6 = D.BatchWork(net, N, D.PREC_BF16X6)
    t6 = b6.forward(x)
    e6 = rel_by_order(t6, t, deg)
    print(name, "BF16X6 vs FP64 at full size, by order:", e6)
    assert max(v for p, v in e6.items() if p <= 4) <= TOL_MIXED
    g6 = D.Grad(net)
    b6.residual_step(x, res, 1.0, g6, want_rsq=False)
    e6g = rel(g6.get(), g_full)
    print(name, "BF16X6 gradient vs FP64", e6g)
    assert e6g <= TOL_MIXED
    b6.free(); g6.free(); g.free(); bw.free(); net.free()
