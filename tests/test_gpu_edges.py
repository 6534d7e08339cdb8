"""Edge cases of the batched path on the GPU, through the C ABI.

  * every activation of the reference (tanh, sin, erf, J0, J1:
    lib/multi_index_bell_module.f90:723-867) on a two-output network, forward
    and adjoint, against the reference's own lib/ compiled by `make ref`
    (oracle/_ref) and, for the activations the oracle restates, against the
    oracle.  J0 and J1: the reference takes J_n from Zhang and Jin's
    jyna/jy01b (lib/bessel_module.f90), whose sigma^(q), q <= 6, differ from
    the Bessel function's by 2e-9 to 6e-9 relative per order (measured
    against scipy.special.jv on [-4, 4]); this library evaluates J_n in
    double precision, so against the reference the bound is 1e-7 (those
    errors carried through two hidden layers), and the derivatives
    themselves are pinned against scipy's closed form
    d^q J_n = 2^-q sum_k (-1)^k C(q,k) J_(n-q+2k) to 1e-13;
  * zero points, one point, and batches that shrink and grow inside one work
    space (no state of an earlier call may leak into a later one);
  * a point count above the work space's maximum is refused with -1;
  * one input at the largest order the tables take (K = 15, 16 slots);
  * the BF16x6 kernels selected by dnnf90_set_split_kernels: the fixed-chunk
    linear maps of 32- and 64-wide layers and the one-block forward map of
    128 x 128 layers give the general kernel's bits (forward planes), the one-block weight gradient (128 x 128 tiles of
    wider blocks) the tiled kernel's
    gradient to 1e-6 of max |gradient| (the same products summed in another
    FP32 order).
FP64 mode elsewhere, 1e-12 relative per derivative order (forward) and to
max |gradient| (adjoint).
"""
import numpy as np
import pytest

from helpers import rel, rel_by_order

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
SEEDS = [[4, 0], [0, 2], [1, 1]]


@pytest.fixture(scope="module")
def D():
    import dnnf90_amd
    return dnnf90_amd


def _net(D, d0, k, seeds, act, dims, seed=1):
    ts = D.TableSet(d0, k, seeds, act)
    net = D.Net(ts, dims)
    rng = np.random.default_rng(seed)
    w = rng.uniform(-0.6, 0.6, net.nw)
    net.weights = w
    return ts, net, w


@pytest.mark.parametrize("act", [0, 1, 2, 3, 4], ids=["tanh", "sin", "erf", "J0", "J1"])
def test_activations_vs_reference(D, oracle_mod, act):
    if not oracle_mod.RefLib.available():
        pytest.skip("reference library not built")
    dims = [2, 12, 10, 2]
    ts, net, w = _net(D, 2, 4, SEEDS, act, dims)
    npt = 37
    rng = np.random.default_rng(2)
    x = rng.uniform(-1, 1, (npt, 2))
    seed = rng.standard_normal((npt, 2, ts.na))
    bw = D.BatchWork(net, npt, D.PREC_FP64)
    g = D.Grad(net)
    t = bw.forward(x)
    bw.backward(seed, g)
    ref = oracle_mod.RefNet(2, 4, SEEDS, act, dims, w)
    tr = ref.eval(x)
    gr = ref.grad(x, seed)
    deg = np.array([sum(a) for a in oracle_mod.Oracle(2, 4, SEEDS).alpha()])
    ef = max(rel_by_order(t, tr, deg).values())
    eg = rel(g.get(), gr)
    print(f"activation {act}: forward {ef:.2e} adjoint {eg:.2e} vs the compiled reference")
    tol = TOL64 if act in (0, 1, 2) else 1e-7   # the reference's own J_n error, see above
    assert ef <= tol and eg <= tol
    if act in (0, 1, 2):
        o = oracle_mod.Oracle(2, 4, SEEDS, act)
        assert max(rel_by_order(t, o.eval(dims, w, x), deg).values()) <= TOL64
        assert rel(g.get(), o.grad(dims, w, x, seed)) <= TOL64
    bw.free(); g.free(); net.free()


@pytest.mark.parametrize("act,n0", [(3, 0), (4, 1)], ids=["J0", "J1"])
def test_bessel_derivatives_vs_scipy(D, act, n0):
    """A [1,1,1] network with unit weights and zero biases outputs sigma^(q)(x)
    in slot q; here q = 0..8 for x in [-12, 12]."""
    from scipy.special import comb, jv
    ts = D.TableSet(1, 8, None, act)
    net = D.Net(ts, [1, 1, 1])
    net.weights = np.array([0.0, 1.0, 0.0, 1.0])
    x = np.linspace(-12.0, 12.0, 97)[:, None]
    bw = D.BatchWork(net, 97, D.PREC_FP64)
    t = bw.forward(x)[:, 0, :]
    ref = np.stack([sum((-1) ** k * comb(q, k) * jv(n0 - q + 2 * k, x[:, 0]) for k in range(q + 1)) / 2.0 ** q
                    for q in range(9)], 1)
    e = (np.abs(t - ref) / np.abs(ref).max(0)).max(0)
    print(f"J{n0}^(q), q = 0..8, vs scipy: max rel per order", e)
    assert e.max() <= 1e-13
    bw.free(); net.free()


def test_zero_points(D):
    ts, net, w = _net(D, 2, 4, SEEDS, 0, [2, 8, 8, 1])
    bw = D.BatchWork(net, 16, D.PREC_FP64)
    g = D.Grad(net)
    t = bw.forward(np.zeros((0, 2)))
    assert t.shape == (0, 1, ts.na)
    bw.backward(np.zeros((0, 1, ts.na)), g)
    assert not g.get().any()
    rsq = bw.residual_step(np.zeros((0, 2)), D.Residual([(1.0, 1, False), (2.0, 2, True)]), 1.0, g)
    assert rsq == 0.0 and not g.get().any()
    bw.free(); g.free(); net.free()


def test_batches_shrink_and_grow_in_one_work_space(D, oracle_mod):
    dims = [2, 16, 16, 1]
    ts, net, w = _net(D, 2, 4, SEEDS, 0, dims)
    o = oracle_mod.Oracle(2, 4, SEEDS)
    deg = o.deg()
    bw = D.BatchWork(net, 300, D.PREC_FP64)
    g = D.Grad(net)
    rng = np.random.default_rng(3)
    for npt in (300, 65, 1, 64, 129, 1, 300):
        x = rng.uniform(-1, 1, (npt, 2))
        seed = rng.standard_normal((npt, 1, o.na))
        g.zero()
        t = bw.forward(x)
        bw.backward(seed, g)
        ef = max(rel_by_order(t, o.eval(dims, w, x), deg).values())
        eg = rel(g.get(), o.grad(dims, w, x, seed))
        print(f"npt {npt}: forward {ef:.2e} adjoint {eg:.2e}")
        assert ef <= TOL64 and eg <= TOL64
    bw.free(); g.free(); net.free()


def test_npt_above_maxpt_refused(D):
    ts, net, w = _net(D, 2, 4, SEEDS, 0, [2, 8, 1])
    bw = D.BatchWork(net, 10, D.PREC_FP64)
    with pytest.raises(D.DNNF90Error) as e:
        bw.forward(np.zeros((11, 2)))
    assert e.value.rc == -1
    bw.free(); net.free()


@pytest.mark.parametrize("codegen", [1, 0], ids=["generated", "table"])
def test_one_input_order_15(D, oracle_mod, codegen):
    from dnnf90_amd._lib import lib
    prev = lib().dnnf90_set_codegen(codegen)
    try:
        dims = [1, 10, 10, 1]
        ts, net, w = _net(D, 1, 15, None, 0, dims, seed=4)
        o = oracle_mod.Oracle(1, 15, None)
        assert ts.na == 16 == o.na
        x = np.linspace(-0.9, 0.9, 50)[:, None]
        seed = np.random.default_rng(5).standard_normal((50, 1, 16))
        bw = D.BatchWork(net, 50, D.PREC_FP64)
        g = D.Grad(net)
        t = bw.forward(x)
        bw.backward(seed, g)
        ef = max(rel_by_order(t, o.eval(dims, w, x), o.deg()).values())
        eg = rel(g.get(), o.grad(dims, w, x, seed))
        print(f"d0 = 1, K = 15 ({'generated' if codegen else 'table'}): forward {ef:.2e} adjoint {eg:.2e}")
        assert ef <= TOL64 and eg <= TOL64
        bw.free(); g.free(); net.free()
    finally:
        lib().dnnf90_set_codegen(prev)


def test_profile_one_category(D):
    """dnnf90_bwork_profile(bid, 2 + c) brackets category c alone; values
    outside 0 .. 2 + ncategories - 1 are refused."""
    from dnnf90_amd._lib import lib
    ts, net, w = _net(D, 2, 3, None, 0, [2, 8, 8, 1])
    bw = D.BatchWork(net, 64, D.PREC_FP64)
    g = D.Grad(net)
    x = np.random.default_rng(9).uniform(-1, 1, (64, 2))
    seed = np.ones((64, 1, ts.na))
    bw.kernel_ms()
    bw.profile(True, only="bell_bwd")
    bw.forward(x, want=False)
    bw.backward(seed, g)
    bw.profile(False)
    k = bw.kernel_ms()
    print("one category:", {c: v for c, v in k.items() if v[1]})
    assert k["bell_bwd"][1] == 2 and k["bell_bwd"][0] > 0.0   # two hidden layers
    assert all(v[1] == 0 for c, v in k.items() if c != "bell_bwd")
    bw.profile(True)
    bw.forward(x, want=False)
    bw.profile(False)
    k = bw.kernel_ms()
    # the first hidden layer forms its S inside the Faa di Bruno kernel (edge
    # kernels): linear maps of the second hidden and the output layer only
    assert k["linear"][1] == 2 and k["bell_fwd"][1] == 2
    prev = lib().dnnf90_set_edge_kernels(0)
    try:
        bw.profile(True)
        bw.forward(x, want=False)
        bw.profile(False)
        k = bw.kernel_ms()
    finally:
        lib().dnnf90_set_edge_kernels(prev)
    assert k["linear"][1] == 3 and k["bell_fwd"][1] == 2
    n = len(D.BatchWork.KERNEL_CATEGORIES)
    assert lib().dnnf90_bwork_profile(bw.id, 2 + n) == -1
    assert lib().dnnf90_bwork_profile(bw.id, -1) == -1
    bw.free(); g.free(); net.free()


@pytest.mark.parametrize("dims", [[2, 64, 64, 64, 1], [3, 32, 32, 2], [2, 128, 128, 1], [3, 256, 200, 1]],
                         ids=["w64", "w32", "w128", "w256"])
def test_split_kernels_vs_general(D, dims):
    """dnnf90_set_split_kernels(0) routes every BF16x6 linear map through the
    general kernel and the weight gradient through the tiled kernel.  The
    fixed-chunk maps (whole 32-row chunks) must give the same bits: the same
    products in the same order.  The one-block gradient sums the same
    products in another FP32 order (each chunk in one accumulator, not four
    interleaved ones): within 1e-6 of max |gradient|."""
    from dnnf90_amd._lib import lib
    ts, net, w = _net(D, dims[0], 4, None, 0, dims, seed=7)
    rng = np.random.default_rng(8)
    npt = 300
    x = rng.uniform(-1, 1, (npt, dims[0]))
    seed = rng.standard_normal((npt, dims[-1], ts.na))
    out = []
    for on in (1, 0):
        prev = lib().dnnf90_set_split_kernels(on)
        try:
            bw = D.BatchWork(net, npt, D.PREC_BF16X6)
            g = D.Grad(net)
            t = np.array(bw.forward(x), copy=True)
            bw.backward(seed, g)
            out.append((t, np.array(g.get(), copy=True)))
            bw.free(); g.free()
        finally:
            lib().dnnf90_set_split_kernels(prev)
    (t1, g1), (t0, g0) = out
    assert np.all(np.isfinite(t1)) and np.all(np.isfinite(g1))
    assert np.array_equal(t1, t0)
    assert rel(g1, g0) <= 1e-6
    net.free()


@pytest.mark.parametrize("budget", [0, 1], ids=["keep", "recompute"])
@pytest.mark.parametrize("dims", [[2, 64, 64, 64, 1], [3, 32, 32, 32, 2], [2, 64, 40, 64, 64, 1], [4, 24, 64, 1]],
                         ids=["w64", "w32", "w40", "w24"])
def test_fused_grad_identical(D, dims, budget):
    """dnnf90_set_fused_grad: a BF16x6 hidden layer of 32 or 64 neurons over
    at most 64 inputs forms its weight gradient and the layer below's Tbar
    in one kernel that reads Sbar once.  Tbar takes the same products in the
    same order as the separate map, so everything computed from it -- the
    first hidden layer's gradient here -- is identical bit for bit; the fused
    layers' own gradients sum the same products over more column chunks
    (three workgroups a CU instead of two), within 1e-6 of max |gradient|.
    The step must launch fewer transposed maps (dnnf90_net_step_info says
    which layers)."""
    from dnnf90_amd._lib import lib
    ts, net, w = _net(D, dims[0], 4, None, 0, dims, seed=17)
    rng = np.random.default_rng(18)
    npt = 300
    x = rng.uniform(-1, 1, (npt, dims[0]))
    seed = rng.standard_normal((npt, dims[-1], ts.na))
    print(f"fused grad: dims={dims} budget={budget}", flush=True)
    info = np.zeros(5, np.int32)
    assert lib().dnnf90_net_step_info(net.id, D.PREC_BF16X6, info.ctypes.data) == 0
    out, nlin_t = [], []
    for on in (1, 0):
        prev = lib().dnnf90_set_fused_grad(on)
        try:
            bw = D.BatchWork(net, npt, D.PREC_BF16X6, budget=budget)
            g = D.Grad(net)
            t = np.array(bw.forward(x), copy=True)
            bw.kernel_ms()
            bw.profile(True)
            bw.backward(seed, g)
            bw.profile(False)
            nlin_t.append(bw.kernel_ms()["linear_t"][1])
            out.append((t, np.array(g.get(), copy=True)))
            bw.free(); g.free()
        finally:
            lib().dnnf90_set_fused_grad(prev)
    (t1, g1), (t0, g0) = out
    assert np.all(np.isfinite(g1))
    assert np.array_equal(t1, t0)
    n1 = dims[1] * (dims[0] + 1)   # the first hidden layer's weights lead the vector
    assert np.array_equal(g1[:n1], g0[:n1])
    assert rel(g1, g0) <= 1e-6
    nf = bin(int(info[4])).count("1")
    print("fused layers mask", int(info[4]), "linear_t launches", nlin_t)
    assert nlin_t[0] == nlin_t[1] - nf
    if dims[1:-1] == [64, 64, 64] or dims[1:-1] == [32, 32, 32]:
        assert nf == 2   # layers 2 and 3; the last hidden layer's Tbar comes from the output layer's Sbar
    net.free()


def test_stale_dependents_refused(D):
    """Handles of work spaces and accumulators built on a network that was
    freed are refused (-1) after the network's slot is reused: the
    generation counters of lib/c_api_module.f90:44-62."""
    import ctypes as C
    from dnnf90_amd._lib import lib
    L = lib()
    ts = D.TableSet(2, 3, [[0, 1], [1, 0], [3, 0]], 0)
    dims = (C.c_int * 3)(2, 8, 1)
    nid = L.dnnf90_net_create_ts(ts.id, 3, dims)
    assert nid > 0
    wid, tid, gid = L.dnnf90_work_create(nid), L.dnnf90_twork_create(nid), L.dnnf90_grad_create(nid)
    bid = L.dnnf90_bwork_create(nid, 4, 0)
    assert min(wid, tid, gid, bid) > 0
    x = np.array([0.3, 0.2])
    t = np.zeros(5)
    seed = np.ones(5)
    assert L.dnnf90_eval(nid, wid, x.ctypes.data, t.ctypes.data) == 0
    assert L.dnnf90_net_free(nid) == 0
    nid2 = L.dnnf90_net_create_ts(ts.id, 3, dims)
    assert nid2 == nid                     # the slot is reused
    assert L.dnnf90_eval(nid2, wid, x.ctypes.data, t.ctypes.data) == -1
    assert L.dnnf90_grad_point(nid2, tid, x.ctypes.data, seed.ctypes.data, gid) == -1
    assert L.dnnf90_forward_batch(nid2, bid, 1, x.ctypes.data, None, 0) == -1
    # fresh dependents of the new network work
    wid2 = L.dnnf90_work_create(nid2)
    assert wid2 > 0 and L.dnnf90_eval(nid2, wid2, x.ctypes.data, t.ctypes.data) == 0
    for f, h in ((L.dnnf90_work_free, wid), (L.dnnf90_work_free, wid2), (L.dnnf90_twork_free, tid),
                 (L.dnnf90_grad_free, gid), (L.dnnf90_bwork_free, bid)):
        assert f(h) == 0
    assert L.dnnf90_net_free(nid2) == 0


EDGE_CASES = [("closure4", 4, 7, [[0, 0, 0, 1], [1, 0, 0, 0], [7, 0, 0, 0], [3, 2, 2, 0]], [4, 24, 24, 24, 1]),
              ("one_hidden", 2, 4, SEEDS, [2, 40, 1]),
              ("two_out", 3, 3, None, [3, 16, 16, 2])]


@pytest.mark.parametrize("budget", [0, 1], ids=["keep", "recompute"])
@pytest.mark.parametrize("prec", [0, 3], ids=["fp64", "bf16x6"])
@pytest.mark.parametrize("case", EDGE_CASES, ids=[c[0] for c in EDGE_CASES])
def test_edge_kernels(D, oracle_mod, case, prec, budget):
    """dnnf90_set_edge_kernels: the first hidden layer forms S from the input
    plane (no linear map, no S plane) and the last hidden layer forms Tbar
    from the output layer's Sbar (no transposed map); one hidden layer takes
    both at once, two outputs sum Tbar over both rows.  Forward and gradient
    with the edge kernels on and off agree to rounding (1e-13 FP64, 1e-6
    BF16x6) and the FP64 run matches the oracle to 1e-12; so do the
    Gauss-Newton rows (the per-point adjoint takes the same edge kernels)."""
    from dnnf90_amd._lib import lib
    name, d0, k, seeds, dims = case
    ts, net, w = _net(D, d0, k, seeds, 0, dims, seed=11)
    rng = np.random.default_rng(12)
    npt = 200
    x = rng.uniform(-1, 1, (npt, d0))
    seed = rng.standard_normal((npt, dims[-1], ts.na))
    print(f"edge kernels: {name} dims={dims} prec={prec} budget={budget}", flush=True)
    out = []
    for on in (1, 0):
        prev = lib().dnnf90_set_edge_kernels(on)
        try:
            bw = D.BatchWork(net, npt, prec, budget=budget)
            g = D.Grad(net)
            t = np.array(bw.forward(x), copy=True)
            bw.backward(seed, g)
            out.append((t, np.array(g.get(), copy=True)))
            bw.free(); g.free()
        finally:
            lib().dnnf90_set_edge_kernels(prev)
    (t1, g1), (t0, g0) = out
    tol = 1e-13 if prec == 0 else 1e-6
    assert rel(t1, t0) <= tol and rel(g1, g0) <= tol
    if prec == 0:
        o = oracle_mod.Oracle(d0, k, seeds)
        assert rel(t1, o.eval(dims, w, x)) <= TOL64
        assert rel(g1, o.grad(dims, w, x, seed)) <= TOL64
    net.free()


def test_edge_kernels_rows(D, oracle_mod):
    """Gauss-Newton rows through the edge kernels against the oracle."""
    import torch
    ts, net, w = _net(D, 2, 4, SEEDS, 0, [2, 20, 20, 1], seed=13)
    rng = np.random.default_rng(14)
    npt = 100
    x = rng.uniform(-1, 1, (npt, 2))
    o = oracle_mod.Oracle(2, 4, SEEDS)
    al = [tuple(a) for a in o.alpha()]
    terms = [(1.0, al.index((4, 0)), False), (0.5, al.index((0, 2)), False), (2.0, al.index((1, 0)), True)]
    bw = D.BatchWork(net, npt, D.PREC_FP64)
    rows_t = torch.zeros((net.nw, 128), dtype=torch.float32, device="cuda")
    bw.residual_rows(x, D.Residual(terms), rows_t, 128)
    torch.cuda.synchronize()
    rows = rows_t[:, :npt].double().cpu().numpy().T
    assert rel(rows, o.pinn_rows([2, 20, 20, 1], w, x, terms)) <= 1e-6
    bw.free(); net.free()


def test_rows_stride_refused(D):
    """dnnf90_residual_rows_batch writes whole 64-point blocks: a row stride
    below npt rounded up to 64 is refused (-1), the rounded one accepted."""
    import torch
    ts, net, w = _net(D, 2, 4, SEEDS, 0, [2, 8, 1])
    o_terms = [(1.0, 1, False)]
    bw = D.BatchWork(net, 100, D.PREC_FP64)
    x = np.random.default_rng(3).uniform(-1, 1, (100, 2))
    small = torch.zeros((net.nw, 100), dtype=torch.float32, device="cuda")
    with pytest.raises(D.DNNF90Error):
        bw.residual_rows(x, D.Residual(o_terms), small, 100)
    ok = torch.zeros((net.nw, 128), dtype=torch.float32, device="cuda")
    bw.residual_rows(x, D.Residual(o_terms), ok, 128)
    torch.cuda.synchronize()
    assert float(ok[:, 100:].abs().max()) == 0.0
    bw.free(); net.free()
