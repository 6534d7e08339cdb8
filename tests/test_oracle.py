"""Pins of the oracle (CPU, no GPU).

Each check below is independent of the oracle's own formulas:
  * the reference's golden files (zwork/hod_4d_k3, hod_4d_k7, hod_4d_k7_active;
    made by an independent Taylor jet, tools/make_hod_golden.py);
  * the directional known-answer table (bench/cpp/README.md:101-109 on the
    weights and points of bench/julia/wx_ref.txt) for the adjoint, K = 1..7;
  * the closed form of a one-hidden-layer network for the forward;
  * symbolic sigma^(q) (sympy) for tanh, sin, erf and erf-GELU;
  * central differences of the seeded loss (tools/dbg_backward_fd.f90:10-66);
  * |A| against sum_p C(p+D0-1, p) and the slot order file;
  * bitwise agreement with the reference's own lib/ compiled by `make ref`
    (skipped where that library is absent).
"""
import math
import os

import numpy as np
import pytest

from helpers import GOLDEN, GOLDEN_CASES, check_weights, read_golden, rel


@pytest.mark.parametrize("case", list(GOLDEN_CASES))
def test_golden_files(oracle_mod, case):
    d0, k, seeds, dims = GOLDEN_CASES[case]
    o = oracle_mod.Oracle(d0, k, seeds)
    x, gold = read_golden(case)
    assert o.na == len(gold)
    t = o.eval(dims, check_weights(dims), x[None, :])[0, 0]
    # reference tolerance 1e-10 (app/hod_check_module.f90:21-24)
    assert np.abs(t - gold).max() <= 1e-14


def test_slot_order_file(oracle_mod):
    rows = [list(map(int, l.split())) for l in open(os.path.join(GOLDEN, "hod_4d_k3", "hod_alpha_order.dat"))
            if l.strip() and not l.startswith("#")]
    o = oracle_mod.Oracle(4, 3)
    a = o.alpha()
    assert len(rows) == o.na
    for r, (ia, dg, *al) in zip(a, rows):
        assert list(r) == al and sum(al) == dg


@pytest.mark.parametrize("d0,k", [(1, 5), (2, 3), (3, 4), (4, 7), (6, 3), (10, 2)])
def test_count_closed_form(oracle_mod, d0, k):
    o = oracle_mod.Oracle(d0, k)
    assert o.na == sum(math.comb(p + d0 - 1, p) for p in range(k + 1))


def test_closure_counts(oracle_mod):
    # |A| of the reference logs (bench/kdv/a.log:4, zwork/hod_4d_k7_active/a.log:4)
    assert oracle_mod.Oracle(2, 3, [[0, 1], [1, 0], [3, 0]]).na == 5
    assert oracle_mod.Oracle(4, 7, [[7, 0, 0, 0], [0, 0, 0, 1], [2, 2, 1, 0]]).na == 24


def _wx_ref():
    tok = open(os.path.join(GOLDEN, "wx_ref", "wx_ref.txt")).read().split()
    npt, nw, d0 = int(tok[1]), int(tok[2]), int(tok[3])
    nums = np.array([float(v) for v in tok[4:]])
    return nums[:nw], nums[nw:nw + d0 * npt].reshape(npt, d0)


# bench/cpp/README.md:101-109: loss sum_n (d^K u / dx_1^K)^2 and dL/dw_1
WX_TABLE = {1: (7.539098907068e-05, -3.00304789e-04), 2: (5.862671447015e-08, -4.05874717e-07),
            3: (7.890390176833e-08, -5.18918334e-07), 4: (7.663173332728e-10, -6.19269886e-09),
            5: (1.158632876420e-09, -7.04080029e-09), 6: (5.408322663523e-11, -3.42215171e-10),
            7: (8.175306274572e-11, -3.10379965e-10)}


def wx_weights():
    """wx_ref.txt stores each neuron as [w_in..., bias] (bias last,
    bench/julia/taylordiff_bench.jl:35-48); the ABI wants bias first."""
    wr, x = _wx_ref()
    dims = [4, 8, 8, 1]
    out, p = [], 0
    for l in range(1, 4):
        for j in range(dims[l]):
            blk = wr[p:p + dims[l - 1] + 1]
            out += [blk[-1]] + list(blk[:-1])
            p += dims[l - 1] + 1
    return dims, np.array(out), x


@pytest.mark.parametrize("K", range(1, 8))
def test_directional_known_answers(oracle_mod, K):
    dims, w, x = wx_weights()
    o = oracle_mod.Oracle(4, K, [[K, 0, 0, 0]])
    t = o.eval(dims, w, x)[:, 0, :]
    loss = float((t[:, K] ** 2).sum())
    seed = np.zeros((x.shape[0], 1, o.na))
    seed[:, 0, K] = 2.0 * t[:, K]
    g = o.grad(dims, w, x, seed)
    lref, gref = WX_TABLE[K]
    assert abs(loss - lref) <= 1e-9 * abs(lref)
    # w_1 is the first input weight of the first hidden neuron: ABI index 1
    assert abs(g[1] - gref) <= 1e-5 * abs(gref)


def test_one_hidden_layer_closed_form(oracle_mod):
    """u = b + sum_j v_j tanh(s_j): d^a u = sum_j v_j tanh^(|a|)(s_j) prod_i W_ji^a_i."""
    import sympy as sp
    rng = np.random.default_rng(1)
    d0, k, n = 3, 5, 7
    o = oracle_mod.Oracle(d0, k)
    W = rng.normal(size=(n, d0))
    b = rng.normal(size=n)
    v = rng.normal(size=n)
    c = rng.normal()
    w = np.concatenate([np.column_stack([b, W]).reshape(-1), [c], v])
    x = rng.uniform(-1, 1, d0)
    s = sp.Symbol("s")
    dt = [sp.lambdify(s, sp.diff(sp.tanh(s), s, q)) for q in range(k + 1)]
    t = o.eval([d0, n, 1], w, x[None])[0, 0]
    sj = b + W @ x
    exp = []
    for a in o.alpha():
        q = int(sum(a))
        val = sum(v[j] * dt[q](sj[j]) * np.prod(W[j] ** a) for j in range(n))
        exp.append(val + (c if q == 0 else 0.0))
    assert rel(t, np.array(exp)) <= 1e-13


@pytest.mark.parametrize("act", [0, 1, 2, 5])
def test_sigma_derivatives_symbolic(oracle_mod, act):
    import sympy as sp
    s = sp.Symbol("s")
    f = {0: sp.tanh(s), 1: sp.sin(s), 2: sp.erf(s), 5: s * (1 + sp.erf(s / sp.sqrt(2))) / 2}[act]
    o = oracle_mod.Oracle(1, 7, act=act)
    tpc = o.tpc()
    for a in (-2.3, -0.4, 0.0, 0.7, 1.9):
        d = o.act_derivs(a, 8)
        ref = np.array([float(sp.N(sp.diff(f, s, q).subs(s, a), 30)) for q in range(9)])
        scale = np.maximum(np.abs(ref), 1.0)
        if act == 0:
            # Horner on P_q(t) (the reference's form) loses sum_i |c_i| |t|^i
            # relative to |P_q(t)| as |t| -> 1: that is the bound it must meet
            t = abs(math.tanh(a))
            cond = np.array([sum(abs(tpc[q, i]) * t ** i for i in range(q + 2)) for q in range(9)])
            scale = np.maximum(scale, cond)
        assert np.all(np.abs(d - ref) <= 1e-13 * scale), (act, a, d - ref)


def test_adjoint_central_differences(oracle_mod):
    """The recipe of tools/dbg_backward_fd.f90:10-66: D0=2, K=3 dense, net
    [2,5,1], w = 0.3 sin(17l+7j+3i+0.5), seed_i = cos(2.7 i), h = 1e-6."""
    dims = [2, 5, 1]
    o = oracle_mod.Oracle(2, 3)
    w = []
    for l in range(1, 3):
        for j in range(1, dims[l] + 1):
            for i in range(0, dims[l - 1] + 1):
                w.append(0.3 * math.sin(17.0 * (l + 1) + 7.0 * j + 3.0 * i + 0.5))
    w = np.array(w)
    seed = np.cos(2.7 * np.arange(1, o.na + 1))
    x = np.array([[0.37, -0.61]])
    g = o.grad(dims, w, x, seed[None, None, :])
    h, emax = 1e-6, 0.0
    for p in range(len(w)):
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        num = float((seed * (o.eval(dims, wp, x)[0, 0] - o.eval(dims, wm, x)[0, 0])).sum() / (2 * h))
        emax = max(emax, abs(num - g[p]) / max(abs(num), 1e-12))
    assert emax <= 1e-6   # the reference measured 2.2e-10


def test_pinn_seed_central_differences(oracle_mod):
    """The collocation seed with a product term (u*u_x) against FD of the loss."""
    o = oracle_mod.Oracle(2, 3, [[0, 1], [1, 0], [3, 0]])
    dims = [2, 6, 6, 1]
    rng = np.random.default_rng(4)
    w = rng.normal(0, 0.5, Oracle_nw := sum(dims[l] * (dims[l - 1] + 1) for l in range(1, 4)))
    x = rng.uniform(-1, 1, (5, 2))
    al = [tuple(a) for a in o.alpha()]
    terms = [(1.0, al.index((0, 1)), False), (6.0, al.index((1, 0)), True), (1.0, al.index((3, 0)), False)]
    g, rsq, _ = o.pinn_step(dims, w, x, terms, 1.0)
    h = 1e-6
    for p in range(0, Oracle_nw, 7):
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        lp = o.pinn_step(dims, wp, x, terms, 1.0)[1]
        lm = o.pinn_step(dims, wm, x, terms, 1.0)[1]
        num = 0.5 * (lp - lm) / (2 * h)      # L = (1/2) sum R^2 for fac = 1
        assert abs(num - g[p]) <= 1e-6 * max(abs(num), np.abs(g).max())


def test_ngd_solve_against_cholesky(oracle_mod):
    import scipy.linalg as sl
    rng = np.random.default_rng(2)
    J = rng.normal(size=(30, 20))
    b = rng.normal(size=20)
    for mode, mu in ((0, 1e-3), (1, 1e-2)):
        u = oracle_mod.Oracle.ngd_solve(J, b, 25, mu, mode)
        G = J.T @ J / 25
        c = mu if mode == 1 else mu * np.trace(G)
        ref = sl.cho_solve(sl.cho_factor(G + c * np.eye(20)), b / 25)
        assert rel(u, ref) <= 1e-10


REF_CASES = [(2, 3, [[0, 1], [1, 0], [3, 0]], [2, 32, 32, 32, 1], 0),
             (4, 3, None, [4, 16, 16, 1], 0),
             (4, 7, [[7, 0, 0, 0], [1, 0, 0, 0], [0, 0, 0, 1], [3, 2, 2, 0]], [4, 8, 8, 1], 0),
             (3, 4, None, [3, 12, 9, 1], 2),
             (2, 4, [[4, 0], [0, 2], [1, 0]], [2, 10, 10, 2], 1)]


@pytest.mark.parametrize("d0,k,seeds,dims,act", REF_CASES)
def test_bitwise_vs_compiled_reference(oracle_mod, d0, k, seeds, dims, act):
    if not oracle_mod.RefLib.available():
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    rng = np.random.default_rng(d0 * 10 + k)
    o = oracle_mod.Oracle(d0, k, seeds, act)
    w = rng.standard_normal(o.nweights(dims)) * 0.4
    r = oracle_mod.RefNet(d0, k, seeds, act, dims, w)
    x = rng.uniform(-1, 1, (6, d0))
    seed = rng.standard_normal((6, dims[-1], o.na))
    assert np.array_equal(o.eval(dims, w, x, 1), r.eval(x))
    assert np.array_equal(o.grad(dims, w, x, seed, 1), r.grad(x, seed))


def test_vectors_from_reference(oracle_mod):
    """The oracle against the committed vectors the compiled reference wrote
    at the measured workloads' shapes (tests/golden/gen_ref_vectors.py)."""
    from dnnf90_amd.workloads import WORKLOADS
    path = os.path.join(GOLDEN, "ref_vectors.npz")
    if not os.path.exists(path):
        pytest.skip("ref_vectors.npz not generated")
    z = np.load(path)
    for name, wl in WORKLOADS.items():
        if f"{name}_npt" not in z:
            continue   # erf-GELU: no reference counterpart
        w, x = wl.inputs(int(z[f"{name}_npt"]))
        o = oracle_mod.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
        t = o.eval(wl.dims, w, x)
        assert rel(t, z[f"{name}_T"]) <= 1e-15, name
        idx = z[f"{name}_gidx"]
        g = o.grad(wl.dims, w, x, z[f"{name}_seed"])
        assert rel(g[idx], z[f"{name}_g"]) <= 1e-14, name


# ---- Gauss-Newton rows: pins of Oracle.pinn_rows / Oracle.grad_rows ----
# The reference forms row n as the gradient of R_n: set_pinn_seed(t, 1, 1)
# then lnet_seed_row, i.e. net_backward_point on a zeroed nabla
# (app/sgd_batch_module.f90:740-745, app/lib_net_module.f90:314-330).

def _row_case(oracle_mod):
    o = oracle_mod.Oracle(2, 4, [[4, 0], [0, 2], [1, 0]])
    al = [tuple(a) for a in o.alpha()]
    # linear terms and two product terms, one of them on slot 0 itself (u*u)
    terms = [(1.0, al.index((4, 0)), False), (-0.7, al.index((0, 2)), False),
             (2.5, al.index((1, 0)), True), (0.3, 0, True)]
    dims = [2, 7, 6, 1]
    rng = np.random.default_rng(11)
    w = rng.normal(0, 0.6, o.nweights(dims))
    x = rng.uniform(-1, 1, (9, 2))
    return o, terms, dims, w, x


def _residual(tv, terms):
    """R = sum_k c_k [T_0 *] T_{a_k} (app/pinn_module.f90:32-47), tv[slot]."""
    return sum(c * (tv[0] * tv[s] if nl else tv[s]) for c, s, nl in terms)


def _dR_dT(tv, terms, na):
    """dR/dT written out: c T_a gives c at a; c T_0 T_a gives c T_0 at a and
    c T_a at 0 (2 c T_0 at 0 when a = 0)."""
    d = np.zeros(na)
    for c, s, nl in terms:
        if nl:
            d[s] += c * tv[0]
            d[0] += c * tv[s]
        else:
            d[s] += c
    return d


def test_pinn_rows_are_residual_jacobian(oracle_mod):
    """Row n is dR_n/dw: central differences of R_n computed here from the
    forward's carried derivatives."""
    o, terms, dims, w, x = _row_case(oracle_mod)
    rows = o.pinn_rows(dims, w, x, terms)
    h = 1e-6
    for p in range(len(w)):
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        tp = o.eval(dims, wp, x)[:, 0, :]
        tm = o.eval(dims, wm, x)[:, 0, :]
        num = np.array([(_residual(a, terms) - _residual(b, terms)) / (2 * h) for a, b in zip(tp, tm)])
        assert np.abs(num - rows[:, p]).max() <= 1e-7 * max(np.abs(rows).max(), 1.0), p


def test_pinn_rows_weighted_sum_is_step_gradient(oracle_mod):
    """sum_n R_n row_n is the gradient of (1/2) sum_n R_n^2, i.e. the
    collocation step at fac = 1."""
    o, terms, dims, w, x = _row_case(oracle_mod)
    rows = o.pinn_rows(dims, w, x, terms)
    g, rsq, tout = o.pinn_step(dims, w, x, terms, 1.0, want_t=True)
    R = np.array([_residual(tv, terms) for tv in tout[:, 0, :]])
    assert abs(float((R * R).sum()) - rsq) <= 1e-14 * rsq
    assert rel(R @ rows, g) <= 1e-13


def test_pinn_rows_one_point_seeded_gradient(oracle_mod):
    """Row n equals the one-point gradient under the seed dR/dT written out
    here, on the oracle and, where built, on the compiled reference's
    net_backward_point (what lnet_seed_row calls)."""
    o, terms, dims, w, x = _row_case(oracle_mod)
    rows = o.pinn_rows(dims, w, x, terms)
    t = o.eval(dims, w, x)[:, 0, :]
    ref = None
    if oracle_mod.RefLib.available():
        ref = oracle_mod.RefNet(2, 4, [[4, 0], [0, 2], [1, 0]], 0, dims, w)
    for n in range(x.shape[0]):
        seed = _dR_dT(t[n], terms, o.na)[None, None, :]
        gn = o.grad(dims, w, x[n:n + 1], seed, 1)
        assert np.array_equal(gn, rows[n]), n
        if ref is not None:
            assert np.array_equal(ref.grad(x[n:n + 1], seed), rows[n]), n


def test_grad_rows_sum_and_single_points(oracle_mod):
    """grad_rows: the rows sum to the batch gradient, and each row is the
    gradient of its point alone."""
    o, terms, dims, w, x = _row_case(oracle_mod)
    seed = np.random.default_rng(12).standard_normal((x.shape[0], 1, o.na))
    rows = o.grad_rows(dims, w, x, seed)
    assert rel(rows.sum(axis=0), o.grad(dims, w, x, seed)) <= 1e-14
    for n in range(x.shape[0]):
        assert np.array_equal(rows[n], o.grad(dims, w, x[n:n + 1], seed[n:n + 1], 1)), n


@pytest.mark.parametrize("name", ["kdv", "closure4_k7"])
def test_pinn_step_vs_compiled_reference(oracle_mod, name):
    """The collocation step (the benchmark's unit of work) on the reference's
    own modules, point by point, against the oracle on one thread: the same
    bits for the gradient and sum R^2."""
    if not oracle_mod.RefLib.available():
        pytest.skip("oracle/_ref not built (needs the reference tree)")
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[name]
    w, x = wl.inputs(40)
    o = oracle_mod.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
    terms = wl.terms(o.alpha())
    g, rsq, _ = o.pinn_step(wl.dims, w, x, terms, 2.0 / 40, nthreads=1)
    r = oracle_mod.RefNet(wl.d0, wl.k, wl.seeds, wl.act, wl.dims, w)
    gr, rsqr, secs = r.pinn_step(x, terms, 2.0 / 40)
    assert np.array_equal(g, gr) and rsq == rsqr and secs >= 0.0
