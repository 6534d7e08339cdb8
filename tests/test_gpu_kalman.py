"""The extended Kalman filter (lib/kalman_module.f90) and the committee
(lib/committee_module.f90:79-111) on the GPU against the oracle.

The oracle's filter is pinned by tests/test_oracle_kf.py (ridge closed form,
Sherman-Morrison, IEKF = plain update, bit-for-bit agreement with the
reference's own kalman_module).  Here the device filter -- covariance in
HBM, one fused pass per observation, observations processed in order on
the device -- runs the same observation sequences and must give the same
weights and covariance to 1e-12 (relative to the largest entry).  The
decoupled filter sums P j in the reference's order, so its agreement is
closer still; the dense filter sums it over 64-wide tiles.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEEDS = [[0, 1], [1, 0], [3, 0]]
TERMS = [(1.0, 1, 0), (3.0, 2, 1), (1.0, 3, 0)]   # u_t + 3 u u_x + u_xxx (bench/opt_kalman)


@pytest.fixture(scope="module")
def D():
    import dnnf90_amd
    return dnnf90_amd


def _obs(o, dims, nobs, seed):
    rng = np.random.default_rng(seed)
    w = rng.uniform(-0.7, 0.7, o.nweights(dims))
    x = rng.uniform(-1, 1, (nobs, o.d0))
    sd = rng.normal(size=(nobs, dims[-1], o.na))
    val = rng.normal(size=nobs)
    rn = rng.uniform(0.5, 2.0, nobs)
    return w, x, sd, val, rn


def _gpu_seq(D, o, ts, dims, w, p0, lam_ini, lam0, dec, gate, q, mode, x, sd, val, rn, niter=1, terms=()):
    net = D.Net(ts, dims)
    net.weights = w
    bw = D.BatchWork(net, 4, D.PREC_FP64)
    kf = D.KalmanFilter(net, p0, lam_ini, lam0, dec, gate, q)
    try:
        if mode == 0:
            kf.update_obs(bw, x, sd, val)
        elif mode == 1:
            kf.update_resid(bw, x, sd, val, rn)
        elif mode == 2:
            for n in range(x.shape[0]):
                kf.iekf_begin()
                for it in range(niter):
                    T = bw.forward(x[n:n + 1])[0]
                    R = 0.0
                    for j in range(T.shape[0]):
                        for a in range(o.na):
                            R = R + float(sd[n, j, a]) * float(T[j, a])
                    kf.iekf_iter(bw, x[n:n + 1], sd[n:n + 1], R - float(val[n]), float(rn[n]), it == niter - 1)
        elif mode == 3:
            kf.update_residual(bw, x, D.Residual(terms), src=val)
        st = kf.state()
        return net.weights, kf.covariance(), st["lam"], st["ngated"], st
    finally:
        kf.free()
        bw.free()
        net.free()


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


CASES = [(mode, dec, gate, q) for mode in (0, 1, 2, 3) for dec in (False, True)
         for gate, q in ((0.0, 0.0), (0.2, 1e-3))]


@pytest.fixture(params=[1, 0], ids=["point_kernels", "layer_kernels"])
def point(request, D):
    """One observation's forward and adjoint as the one-point kernels (one
    launch each way), then as the layer kernels."""
    from dnnf90_amd._lib import lib
    prev = lib().dnnf90_set_point_kernels(request.param)
    yield request.param
    lib().dnnf90_set_point_kernels(prev)


@pytest.mark.parametrize("mode,dec,gate,q", CASES)
def test_kf_sequence_vs_oracle(D, oracle_mod, point, mode, dec, gate, q):
    ts = D.TableSet(2, 3, SEEDS)
    o = oracle_mod.Oracle(2, 3, SEEDS)
    dims = [2, 6, 5, 1] if mode in (0, 3) else [2, 6, 5, 2]
    w, x, sd, val, rn = _obs(o, dims, 12, seed=10 + mode)
    args = (10.0, 0.98, 0.9987, dec, gate, q, mode, x, sd, val, rn, 3 if mode == 2 else 1,
            TERMS if mode == 3 else ())
    g = _gpu_seq(D, o, ts, dims, w, *args)
    r = o.kf_seq(dims, w, *args)
    ew, eP = _rel(g[0], r[0]), _rel(g[1], r[1])
    print(f"kf mode {mode} decoupled {dec} gate {gate} q {q}: weights {ew:.2e} P {eP:.2e} "
          f"ngated {g[3]} / {r[3]}")
    assert ew <= 1e-12 and eP <= 1e-12
    assert g[2] == r[2] and g[3] == r[3]


@pytest.mark.parametrize("dec", [False, True])
def test_kf_many_tiles_vs_oracle(D, oracle_mod, dec):
    """nw = 1249 (20 tiles a side, 210 lower tiles): the tile sums, the
    part reduction and the padding of the last tile row."""
    ts = D.TableSet(4, 2, None)
    o = oracle_mod.Oracle(4, 2, None)
    dims = [4, 32, 32, 1]
    w, x, sd, val, rn = _obs(o, dims, 16, seed=21)
    args = (5.0, 0.99, 0.999, dec, 0.5, 1e-4, 0, x, sd, val, rn)
    g = _gpu_seq(D, o, ts, dims, w, *args)
    r = o.kf_seq(dims, w, *args)
    ew, eP = _rel(g[0], r[0]), _rel(g[1], r[1])
    print(f"kf nw {o.nweights(dims)} decoupled {dec}: weights {ew:.2e} P {eP:.2e} ngated {g[3]} / {r[3]}; "
          f"{g[4]['bytes'] / 2**20:.1f} MiB on the device")
    assert ew <= 1e-12 and eP <= 1e-12 and g[3] == r[3]


def test_kf_linear_net_is_ridge(D):
    """tools/example_libverify.f90:216-280 on the device: D0 = 3, 60 samples,
    p0 = 50, lambda = 1; the filter must give the ridge solution (1e-9)."""
    ts = D.TableSet(3, 1, None)
    NS, p0 = 60, 50.0
    n = np.arange(1, NS + 1)[:, None]
    x = np.sin(1.7 * n + 0.9 * np.arange(1, 4)[None, :])
    y = 0.7 - 1.3 * x[:, 0] + 0.4 * x[:, 1] + 2.1 * x[:, 2]
    net = D.Net(ts, [3, 1])
    net.weights = np.zeros(4)
    bw = D.BatchWork(net, 1, D.PREC_FP64)
    kf = D.KalmanFilter(net, p0, 1.0, 1.0)
    seed = np.zeros((NS, 1, ts.na))
    seed[:, 0, 0] = 1.0                     # kf_update with islot = 1
    kf.update_obs(bw, x, seed, y)
    xe = np.hstack([np.ones((NS, 1)), x])
    b = np.linalg.solve(xe.T @ xe + np.eye(4) / p0, xe.T @ y)
    e = _rel(net.weights, b)
    print("kf vs ridge:", e)
    assert e <= 1e-9
    kf.free(); bw.free(); net.free()


def test_kf_update_grad_vs_oracle(D, oracle_mod):
    """kf_update_grad from a gradient accumulator, two updates in a row."""
    ts = D.TableSet(2, 3, SEEDS)
    o = oracle_mod.Oracle(2, 3, SEEDS)
    dims = [2, 8, 8, 1]
    w, x, sd, val, rn = _obs(o, dims, 2, seed=31)
    net = D.Net(ts, dims)
    net.weights = w
    bw = D.BatchWork(net, 4, D.PREC_FP64)
    kf = D.KalmanFilter(net, 2.0, 0.97, 0.999, False, 0.0, 1e-3)
    g = D.Grad(net)
    L = oracle_mod.Oracle.lib()
    h = L.ot_kf_create(len(dims), oracle_mod._ia(dims).ctypes.data_as(oracle_mod._pi), 2.0, 0.97, 0.999, 0)
    L.ot_kf_set(h, 0.0, 1e-3)
    wr = w.copy()
    for n in range(2):
        g.zero()
        bw.forward(x[n:n + 1])
        bw.backward(sd[n:n + 1], g)
        kf.update_grad(g, float(val[n]), 0.8)
        gr = o.grad(dims, wr, x[n:n + 1], sd[n:n + 1])
        L.ot_kf_update_grad(h, oracle_mod._dp(wr), oracle_mod._dp(gr), float(val[n]), 0.8)
    P = np.zeros(L.ot_kf_cov_len(h))
    L.ot_kf_state(h, oracle_mod._dp(P), None, None)
    L.ot_kf_free(h)
    ew, eP = _rel(net.weights, wr), _rel(kf.covariance(), P)
    print("kf_update_grad: weights", ew, "P", eP)
    assert ew <= 1e-12 and eP <= 1e-12
    g.free(); kf.free(); bw.free(); net.free()


def test_kf_broken_filter_reports(D):
    """A NaN observation gives a NaN denominator: -5, and the filter then
    refuses further updates (the reference stops, kalman_module.f90:399-405)."""
    ts = D.TableSet(2, 3, SEEDS)
    net = D.Net(ts, [2, 4, 1])
    net.weights = np.full(net.nw, 0.1)
    bw = D.BatchWork(net, 1, D.PREC_FP64)
    kf = D.KalmanFilter(net, 1.0, 1.0, 1.0)
    seed = np.zeros((1, 1, ts.na))
    seed[0, 0, 0] = 1.0
    with pytest.raises(D.DNNF90Error) as e:
        kf.update_obs(bw, np.array([[np.nan, 0.0]]), seed, np.array([1.0]))
    assert e.value.rc == -5
    with pytest.raises(D.DNNF90Error) as e:
        kf.update_obs(bw, np.array([[0.1, 0.2]]), seed, np.array([1.0]))
    assert e.value.rc == -5
    assert kf.state()["bad"] == 1
    kf.free(); bw.free(); net.free()


@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (3, 2e-6)], ids=["fp64", "bf16x6"])
def test_committee_vs_oracle(D, oracle_mod, prec, tol):
    ts = D.TableSet(2, 3, SEEDS)
    o = oracle_mod.Oracle(2, 3, SEEDS)
    dims = [2, 16, 16, 1]
    rng = np.random.default_rng(7)
    W = rng.uniform(-0.6, 0.6, (5, o.nweights(dims)))
    x = rng.uniform(-1, 1, (100, 2))
    nets = []
    for k in range(5):
        nk = D.Net(ts, dims)
        nk.weights = W[k]
        nets.append(nk)
    bw = D.BatchWork(nets[0], 128, prec)
    cm = D.Committee(nets)
    m, s = cm.eval(bw, x)
    rm, rs = o.comm_eval(dims, W, x)
    em, es = _rel(m, rm), float(np.abs(s - rs).max() / np.abs(rm).max())
    print(f"committee prec {prec}: mean {em:.2e} std {es:.2e}")
    assert em <= tol and es <= tol
    if prec == 0:
        # identical members: exactly zero spread; one member: zero
        m2, s2 = D.Committee([nets[0]] * 3).eval(bw, x)
        assert not s2.any()
        m1, s1 = D.Committee(nets[:1]).eval(bw, x)
        assert np.array_equal(m1, m2) and not s1.any()
    bw.free()
    for nk in nets:
        nk.free()
