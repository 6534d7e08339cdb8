"""GPU parity: the gfx950 path through the C ABI against the oracle.

The oracle (oracle/) is the C FP64 restatement of lib/train_module.f90:241-450
that agrees bitwise with the reference's own lib/ (tests/test_oracle.py).
FP64 mode must match it to 1e-12 relative, per derivative order, norm-wise
(max |diff| over points and slots of the order / max |oracle| of the order),
the metric of app/hod_check_module.f90:216, 302.
"""
import numpy as np
import pytest

from helpers import GOLDEN_CASES, check_weights, read_golden, rel, rel_by_order

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


@pytest.fixture(scope="module")
def D():
    import dnnf90_amd
    return dnnf90_amd


@pytest.fixture(params=[(1, 1), (1, 0), (0, 1)], ids=["generated", "generated_keep_h", "table"])
def codegen(request, D):
    """Run with the per-family generated kernels (each hidden layer keeping
    S, the adjoint recomputing h; then keeping h), then with the
    table-driven ones."""
    from dnnf90_amd._lib import lib
    on, store_s = request.param
    prev = lib().dnnf90_set_codegen(on)
    prev_s = lib().dnnf90_set_store_s(store_s)
    yield on
    lib().dnnf90_set_codegen(prev)
    lib().dnnf90_set_store_s(prev_s)


def _setup(D, oracle_mod, wl, npt):
    ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
    net = D.Net(ts, wl.dims)
    w, x = wl.inputs(npt)
    net.weights = w
    o = oracle_mod.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
    return ts, net, w, x, o


WL_SMALL = [("kdv", 300), ("dense4_k3", 197), ("closure4_k7", 130), ("gelu3_k4", 70), ("ngd2_k4", 257)]


@pytest.mark.parametrize("case", list(GOLDEN_CASES))
def test_golden_per_point_abi(D, case):
    """dnnf90_eval on the GPU against the reference's golden files."""
    d0, k, seeds, dims = GOLDEN_CASES[case]
    ts = D.TableSet(d0, k, seeds)
    net = D.Net(ts, dims)
    net.weights = check_weights(dims)
    x, gold = read_golden(case)
    wk = D.Work(net)
    t = wk.eval(x)
    print(case, "max|gpu - golden| =", np.abs(t - gold).max())
    assert np.abs(t - gold).max() <= 1e-13
    wk.free()
    net.free()


@pytest.mark.parametrize("name,npt", WL_SMALL)
def test_forward_fp64(D, oracle_mod, codegen, name, npt):
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[name]
    ts, net, w, x, o = _setup(D, oracle_mod, wl, npt)
    bw = D.BatchWork(net, npt, D.PREC_FP64)
    t = bw.forward(x)
    tr = o.eval(wl.dims, w, x)
    err = rel_by_order(t, tr, o.deg())
    print(name, "forward rel err by order", err)
    assert max(err.values()) <= TOL64


@pytest.mark.parametrize("name,npt", WL_SMALL)
def test_adjoint_fp64(D, oracle_mod, codegen, name, npt):
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[name]
    ts, net, w, x, o = _setup(D, oracle_mod, wl, npt)
    bw = D.BatchWork(net, npt, D.PREC_FP64)
    g = D.Grad(net)
    seed = np.random.default_rng(7).standard_normal((npt, wl.dims[-1], o.na))
    bw.forward(x, want=False)
    bw.backward(seed, g)
    gg = g.get()
    gr = o.grad(wl.dims, w, x, seed)
    e = rel(gg, gr)
    print(name, "gradient rel err", e)
    assert e <= TOL64


@pytest.mark.parametrize("name,npt", WL_SMALL)
def test_residual_step_fp64(D, oracle_mod, name, npt):
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[name]
    ts, net, w, x, o = _setup(D, oracle_mod, wl, npt)
    terms = wl.terms(o.alpha())
    bw = D.BatchWork(net, npt, D.PREC_FP64)
    g = D.Grad(net)
    rsq = bw.residual_step(x, D.Residual(terms), 2.0 / npt, g)
    gr, rsq_r, _ = o.pinn_step(wl.dims, w, x, terms, 2.0 / npt)
    print(name, "rsq", rsq, rsq_r, "grad rel", rel(g.get(), gr))
    assert abs(rsq - rsq_r) <= 1e-12 * abs(rsq_r)
    assert rel(g.get(), gr) <= TOL64


@pytest.mark.parametrize("name", ["kdv", "dense4_k3", "closure4_k7", "gelu3_k4", "ngd2_k4"])
def test_generated_kernels_active(D, name):
    """The measured families run the generated kernels, keeping S in every
    hidden layer, without scratch spills."""
    from dnnf90_amd._lib import lib
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[name]
    net = D.Net(D.TableSet(wl.d0, wl.k, wl.seeds, wl.act), wl.dims)
    for prec in (D.PREC_FP64, D.PREC_FP32):
        info = np.zeros(6, np.int32)
        assert lib().dnnf90_net_kernel_info(net.id, prec, info.ctypes.data) == 0
        print(name, "prec", prec, "generated", info[0], "vgprs fwd/bwd", info[1], info[2], "scratch", info[3], info[4])
        assert info[0] == 2
        assert info[3] == 0 and info[4] == 0


@pytest.mark.parametrize("prec", [0, 1, 3], ids=["fp64", "fp32", "bf16x6"])
@pytest.mark.parametrize("name,npt", [("dense4_k3", 197), ("closure4_k7", 130), ("kdv", 100)])
def test_store_s_identical(D, oracle_mod, name, npt, prec):
    """Keeping S and recomputing h in the adjoint gives the same forward
    planes and gradient, bit for bit, as keeping h (edge kernels off: they
    need S kept, and their maps round differently; tests/test_gpu_edges.py
    compares them)."""
    from dnnf90_amd._lib import lib
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[name]
    ts, net, w, x, o = _setup(D, oracle_mod, wl, npt)
    seed = np.random.default_rng(7).standard_normal((npt, wl.dims[-1], o.na))
    out = []
    prev = lib().dnnf90_set_store_s(1)
    prev_edge = lib().dnnf90_set_edge_kernels(0)
    try:
        for store_s in (1, 0):
            lib().dnnf90_set_store_s(store_s)
            info = np.zeros(6, np.int32)
            assert lib().dnnf90_net_kernel_info(net.id, prec, info.ctypes.data) == 0
            bw = D.BatchWork(net, npt, prec)
            g = D.Grad(net)
            t = bw.forward(x)
            bw.backward(seed, g)
            out.append((info[0], t, g.get()))
            bw.free()
            g.free()
    finally:
        lib().dnnf90_set_store_s(prev)
        lib().dnnf90_set_edge_kernels(prev_edge)
    (k1, t1, g1), (k0, t0, g0) = out
    print(name, "prec", prec, "kernels", k1, k0, "max |dT|", np.abs(t1 - t0).max(), "max |dg|", np.abs(g1 - g0).max())
    assert (k1, k0) == (2, 1)
    assert np.array_equal(t1, t0) and np.array_equal(g1, g0)


@pytest.mark.parametrize("name,dims", [("kdv", None), ("ngd2_k4", None), ("dense4_k3", [4, 24, 24, 1]),
                                       ("gelu3_k4", [3, 16, 16, 16, 1])])
def test_point_kernels_vs_layer_kernels(D, oracle_mod, name, dims):
    """The per-point ABI (dnnf90_eval, dnnf90_grad_point) with the one-point
    kernels and with the layer kernels, both against the oracle (1e-12) and
    against each other."""
    from dnnf90_amd._lib import lib
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[name]
    dims = dims or wl.dims
    ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
    net = D.Net(ts, dims)
    o = oracle_mod.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
    rng = np.random.default_rng(11)
    w = rng.uniform(-0.5, 0.5, net.nw)
    net.weights = w
    x = rng.uniform(-1, 1, (5, wl.d0))
    seed = rng.standard_normal((5, 1, o.na))
    tr = o.eval(dims, w, x)
    gr = o.grad(dims, w, x, seed)
    out = []
    prev = lib().dnnf90_set_point_kernels(1)
    try:
        for on in (1, 0):
            lib().dnnf90_set_point_kernels(on)
            wk = D.Work(net)
            tw = D.TWork(net)
            g = D.Grad(net)
            t = np.array([wk.eval(xn) for xn in x])
            for n in range(5):
                tw.grad_point(x[n], seed[n, 0], g)
            out.append((t, g.get()))
            et = max(rel_by_order(t[:, None, :], tr[:, :1, :], o.deg()).values())
            eg = rel(g.get(), gr)
            print(name, dims, "point kernels" if on else "layer kernels", "forward", et, "gradient", eg)
            assert et <= TOL64 and eg <= TOL64
            wk.free(); tw.free(); g.free()
    finally:
        lib().dnnf90_set_point_kernels(prev)
    (t1, g1), (t0, g0) = out
    print("point vs layer kernels: forward", np.abs(t1 - t0).max(), "gradient", rel(g1, g0))
    assert rel(g1, g0) <= 1e-13
    net.free()


def test_per_point_abi_and_adam(D, oracle_mod):
    """dnnf90_grad_point and dnnf90_adam_step (per-point reference ABI)."""
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS["kdv"]
    ts, net, w, x, o = _setup(D, oracle_mod, wl, 9)
    tw = D.TWork(net)
    g = D.Grad(net)
    seed = np.random.default_rng(3).standard_normal((9, 1, o.na))
    for n in range(9):
        tw.grad_point(x[n], seed[n, 0], g)
    gr = o.grad(wl.dims, w, x, seed)
    assert rel(g.get(), gr) <= TOL64
    # Adam (lib/train_module.f90:656-683) against the same update in numpy
    eta, b1, b2, eps, nb = 1e-3, 0.9, 0.999, 1e-8, 9
    m = np.zeros_like(w)
    v = np.zeros_like(w)
    ww = w.copy()
    gv = g.get()
    for it in (1, 2):
        D.adam_step(net, g, eta, b1, b2, eps, nb, it)
        q = gv / nb
        m = b1 * m + (1 - b1) * q
        v = b2 * v + (1 - b2) * q ** 2
        ww = ww - eta * (m / (1 - b1 ** it)) / (np.sqrt(v / (1 - b2 ** it)) + eps)
    assert np.abs(net.weights - ww).max() <= 1e-15


def test_adam_vs_compiled_reference(D, oracle_mod):
    """dnnf90_adam_step on the GPU against the reference's own opt_adam_step
    (lib/train_module.f90:656-683, compiled from its sources into
    oracle/_ref and driven by ref_driver.f90's refx_adam) on the same
    weights and the same fixed gradient, over seven steps from istep 1 and
    from istep 40: the same bits."""
    if not oracle_mod.RefLib.available():
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    dims = [3, 20, 12, 2]
    ts = D.TableSet(3, 2)
    net = D.Net(ts, dims)
    rng = np.random.default_rng(21)
    w = rng.normal(0, 0.4, net.nw)
    gv = rng.normal(0, 2.0, net.nw)
    for istep0 in (1, 40):
        net.weights = w
        g = D.Grad(net)
        g.set(gv)
        for it in range(istep0, istep0 + 7):
            D.adam_step(net, g, 2e-3, 0.9, 0.999, 1e-8, 13, it)
        wg = net.weights
        wr = oracle_mod.ref_adam(dims, w, gv, 2e-3, 0.9, 0.999, 1e-8, 13, istep0, 7)
        print("Adam from istep", istep0, "max |gpu - reference|", np.abs(wg - wr).max())
        assert np.array_equal(wg, wr)
        g.free()
    net.free()


@pytest.mark.parametrize("prec,tol", [(1, 1.5e-6), (2, 2e-4), (3, 1.5e-6)])
@pytest.mark.parametrize("name,npt", [("dense4_k3", 197), ("closure4_k7", 130), ("gelu3_k4", 70)])
def test_mixed_precision(D, oracle_mod, codegen, name, npt, prec, tol):
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[name]
    ts, net, w, x, o = _setup(D, oracle_mod, wl, npt)
    bw = D.BatchWork(net, npt, prec)
    t = bw.forward(x)
    tr = o.eval(wl.dims, w, x)
    err = rel_by_order(t, tr, o.deg())
    g = D.Grad(net)
    seed = np.random.default_rng(7).standard_normal((npt, wl.dims[-1], o.na))
    bw.backward(seed, g)
    ge = rel(g.get(), o.grad(wl.dims, w, x, seed))
    print(name, "prec", prec, "forward err by order", err, "grad err", ge)
    assert max(err.values()) <= tol and ge <= tol


def test_rows_fp64(D, oracle_mod):
    """Gauss-Newton rows dR_n/dw (the NGD metric rows) against the oracle."""
    import torch
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS["ngd2_k4"]
    npt = 100
    ts, net, w, x, o = _setup(D, oracle_mod, wl, npt)
    terms = wl.terms(o.alpha())
    bw = D.BatchWork(net, npt, D.PREC_FP64)
    ld = 128
    rows_t = torch.zeros((net.nw, ld), dtype=torch.float32, device="cuda")
    bw.residual_rows(x, D.Residual(terms), rows_t, ld)
    torch.cuda.synchronize()
    rows = rows_t[:, :npt].double().cpu().numpy().T
    rr = o.pinn_rows(wl.dims, w, x, terms)
    e = rel(rows, rr)
    print("rows rel err", e)
    assert e <= 1e-6


def _ngd_inputs(D, oracle_mod, dims, npt):
    import torch
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS["ngd2_k4"]
    ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
    net = D.Net(ts, dims)
    rng = np.random.default_rng(5)
    w = np.concatenate([np.column_stack([np.zeros(dims[l]), rng.normal(0, np.sqrt(2.0 / (dims[l - 1] + dims[l])),
                                         (dims[l], dims[l - 1]))]).reshape(-1) for l in range(1, len(dims))])
    net.weights = w
    x = rng.uniform(-1, 1, (npt, wl.d0))
    o = oracle_mod.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
    terms = wl.terms(o.alpha())
    bw = D.BatchWork(net, npt, D.PREC_FP64)
    ld = ((npt + 63) // 64) * 64
    rows_t = torch.zeros((net.nw, ld), dtype=torch.float32, device="cuda")
    bw.residual_rows(x, D.Residual(terms), rows_t, ld)
    g = D.Grad(net)
    bw.residual_step(x, D.Residual(terms), 1.0, g)
    torch.cuda.synchronize()
    return net, bw, rows_t, ld, g.get(), o, w, x, terms


@pytest.mark.parametrize("dual", [False, True], ids=["primal", "dual"])
@pytest.mark.parametrize("mode,mu", [(1, 1e-2), (0, 1e-3)])
def test_ngd_solve_vs_oracle(D, oracle_mod, mode, mu, dual):
    """Gram metric (FP32 MFMA) + FP64 Cholesky, primal (nw x nw) or dual
    (nrow x nrow, app/optimizer_module.f90:408-446), against the oracle's
    primal Gaussian elimination (app/optimizer_module.f90:221-522) on the
    same FP32 rows.  The dual route is FP64 throughout (its Gram matrix of
    the FP32 rows is exact to FP64 rounding), so it must agree to 1e-14
    times the condition number; the primal route's metric comes from FP32
    MFMA, which bounds it at 1e-7 times the condition number.  Both use the
    same blocked FP64 Cholesky kernels."""
    dims = [2, 16, 16, 1]
    npt = 200
    net, bw, rows_t, ld, b, o, w, x, terms = _ngd_inputs(D, oracle_mod, dims, npt)
    u = bw.ngd_solve(rows_t, ld, npt, b, npt, mu, mode, dual=dual)
    J = rows_t[:, :npt].double().cpu().numpy().T
    ur = oracle_mod.Oracle.ngd_solve(J, b, npt, mu, mode)
    G = J.T @ J / npt
    c = mu if mode == 1 else mu * np.trace(G)
    cond = np.linalg.cond(G + c * np.eye(len(b)))
    e = rel(u, ur)
    print("ngd mode", mode, "rel err vs oracle", e, "cond", cond)
    assert e <= (1e-14 if dual else 1e-7) * cond
    # the oracle's rows (FP64) agree with the GPU's FP32 rows to FP32 rounding
    assert rel(J, o.pinn_rows(dims, w, x, terms)) <= 1e-6


def test_ngd_full_size_backward_error(D, oracle_mod):
    """The configured NGD step: 4096 rows of the [2,64,64,64,1] net (8577
    weights), ABS damping 1e-3; normwise backward error of the solve."""
    import torch
    dims = [2, 64, 64, 64, 1]
    npt = 4096
    net, bw, rows_t, ld, b, o, w, x, terms = _ngd_inputs(D, oracle_mod, dims, npt)
    assert net.nw == 8577
    u = bw.ngd_solve(rows_t, ld, npt, b, npt, 1e-3, 1)
    Jt = rows_t[:, :npt].double()
    ut = torch.from_numpy(u).cuda()
    r = (Jt @ (Jt.T @ ut)) / npt + 1e-3 * ut - torch.from_numpy(b).cuda() / npt
    gnorm = (Jt * Jt).sum().item() / npt + 1e-3   # Frobenius-type bound of the damped metric
    be = r.norm().item() / (gnorm * ut.norm().item() + np.linalg.norm(b) / npt)
    print("ngd full-size backward error", be)
    assert be <= 1e-7   # the FP32-MFMA metric's rounding (2.7e-8 measured)
    # the dual (Gram, 4096 x 4096) route solves the same system
    ud = bw.ngd_solve(rows_t, ld, npt, b, npt, 1e-3, 1, dual=True)
    udt = torch.from_numpy(ud).cuda()
    rd = (Jt @ (Jt.T @ udt)) / npt + 1e-3 * udt - torch.from_numpy(b).cuda() / npt
    bed = rd.norm().item() / (gnorm * udt.norm().item() + np.linalg.norm(b) / npt)
    print("ngd dual backward error", bed, "rel diff to primal", rel(ud, u))
    assert bed <= 1e-11   # FP64 throughout (2.9e-13 measured)


@pytest.mark.parametrize("prec,tol", [(0, TOL64), (3, 2e-6)], ids=["fp64", "bf16x6"])
def test_grad_batch_value_only(D, oracle_mod, prec, tol):
    """The value-only batched step (net_grad_batch, lib/train_module.f90:525-631)
    against the oracle's adjoint seeded on the value slot only: with T-bar
    zero on every derivative slot, every pair sum for b > 0 vanishes, so
    the derivative-carrying adjoint gives exactly the value-only gradient."""
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS["closure4_k7"]
    npt = 300
    ts, net, w, x, o = _setup(D, oracle_mod, wl, npt)
    y = np.random.default_rng(5).standard_normal(npt)
    coef = 0.37
    bw = D.BatchWork(net, npt, prec)
    g = D.Grad(net)
    bw.grad_batch(x, y, coef, g)
    gv = g.get()
    t = o.eval(wl.dims, w, x)
    seed = np.zeros((npt, 1, o.na))
    seed[:, 0, 0] = coef * (t[:, 0, 0] - y)
    gr = o.grad(wl.dims, w, x, seed)
    e = rel(gv, gr)
    print("value-only grad prec", prec, "rel err", e)
    assert e <= tol
    # the work space still serves the derivative-carrying path afterwards
    tt = bw.forward(x)
    assert max(rel_by_order(tt, t, o.deg()).values()) <= (TOL64 if prec == 0 else 2e-6)


def test_write_c_ref(D, oracle_mod, tmp_path):
    """c_ref.dat (tools/gen_c_ref.f90) of the KdV family through the GPU."""
    from dnnf90_amd.formats import c_ref_points, read_c_ref, write_c_ref
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS["kdv"]
    ts, net, w, x, o = _setup(D, oracle_mod, wl, 25)
    path = tmp_path / "c_ref.dat"
    write_c_ref(net, str(path))
    xr, tr = read_c_ref(str(path))
    assert np.array_equal(xr, c_ref_points())
    ref = o.eval(wl.dims, w, c_ref_points())[:, 0, :]
    e = rel(tr, ref)
    print("c_ref.dat values rel err", e)
    assert e <= TOL64


@pytest.mark.gpu
@pytest.mark.parametrize("w0", [0, 130])
def test_ngd_reports_first_bad_pivot(D, oracle_mod, w0):
    """A NaN in the rows of weight w0 makes row and column w0 of the metric
    NaN: the factorization reports column w0 + 1 (1-based, as the pivot
    check of the diagonal block finds it: w0 = 130 lies in the second
    panel) with -5, and the next solve on the same work space, with clean
    rows, gives the step of a fresh one."""
    dims = [2, 16, 16, 1]
    npt = 200
    net, bw, rows_t, ld, b, o, w, x, terms = _ngd_inputs(D, oracle_mod, dims, npt)
    u0 = bw.ngd_solve(rows_t, ld, npt, b, npt, 1e-3, 1)
    bad = rows_t.clone()
    bad[w0, 3] = float("nan")
    with pytest.raises(D.DNNF90Error) as e:
        bw.ngd_solve(bad, ld, npt, b, npt, 1e-3, 1)
    assert e.value.rc == -5
    assert f"column {w0 + 1}" in str(e.value), str(e.value)
    u1 = bw.ngd_solve(rows_t, ld, npt, b, npt, 1e-3, 1)
    assert rel(u1, u0) == 0.0
