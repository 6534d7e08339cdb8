"""Pins of the oracle's extended Kalman filter and committee (CPU, no GPU).

The filter restatement (lib/kalman_module.f90:107-659) is checked against
  * the ridge normal-equation solution on a net linear in its weights, the
    reference's own self-check (tools/example_libverify.f90:216-280: D0 = 3,
    60 samples, p0 = 50, lambda = 1, tolerance 1e-9);
  * the Sherman-Morrison identity (lambda P1)^-1 = P0^-1 + j j^T / lambda
    after one update;
  * the reference's own statement that one IEKF iteration reproduces the
    plain update (:523-524), and that a net with one neuron has one block, so
    the decoupled filter equals the dense one;
  * the reference's kalman_module itself, compiled by `make ref`, through
    ref_driver.f90's refx_kf_seq, bit for bit (skipped where absent).
The committee (lib/committee_module.f90:79-111) is checked against numpy's
two-pass mean and ddof=1 standard deviation, its exact zero spread for
identical members, and the compiled reference's comm_eval.
"""
import numpy as np
import pytest

SEEDS = [[0, 1], [1, 0], [3, 0]]
TERMS = [(1.0, 1, 0), (3.0, 2, 1), (1.0, 3, 0)]   # u_t + 3 u u_x + u_xxx on this family


def _obs(o, dims, nobs, seed=3):
    rng = np.random.default_rng(seed)
    w = rng.uniform(-0.7, 0.7, o.nweights(dims))
    x = rng.uniform(-1, 1, (nobs, o.d0))
    sd = rng.normal(size=(nobs, dims[-1], o.na))
    val = rng.normal(size=nobs)
    rn = rng.uniform(0.5, 2.0, nobs)
    return w, x, sd, val, rn


def test_kf_linear_net_is_ridge(oracle_mod):
    o = oracle_mod.Oracle(3, 1, None)
    NS, p0 = 60, 50.0
    n = np.arange(1, NS + 1)[:, None]
    i = np.arange(1, 4)[None, :]
    x = np.sin(1.7 * n + 0.9 * i)
    y = 0.7 - 1.3 * x[:, 0] + 0.4 * x[:, 1] + 2.1 * x[:, 2]
    seed = np.zeros((NS, 1, o.na))
    seed[:, 0, 0] = 1.0
    w, P, lam, ng = o.kf_seq([3, 1], np.zeros(4), p0, 1.0, 1.0, False, 0.0, 0.0, 0, x, seed, y)
    xe = np.hstack([np.ones((NS, 1)), x])
    b = np.linalg.solve(xe.T @ xe + np.eye(4) / p0, xe.T @ y)
    assert np.abs(w - b).max() / np.abs(b).max() <= 1e-9
    # the covariance is the inverse of the ridge matrix
    Pm = P.reshape(4, 4).T
    assert np.abs(Pm @ (xe.T @ xe + np.eye(4) / p0) - np.eye(4)).max() <= 1e-9
    assert lam == 1.0 and ng == 0


def test_kf_sherman_morrison(oracle_mod):
    o = oracle_mod.Oracle(2, 3, SEEDS)
    dims = [2, 4, 3, 1]
    w, x, sd, val, rn = _obs(o, dims, 1)
    nw = o.nweights(dims)
    lam_ini, p0 = 0.97, 3.0
    w1, P, lam, _ = o.kf_seq(dims, w, p0, lam_ini, 0.99, False, 0.0, 0.0, 0, x, sd, val)
    j = o.grad(dims, w, x, sd)
    P1 = P.reshape(nw, nw).T
    lhs = np.linalg.inv(lam_ini * P1)
    rhs = np.eye(nw) / p0 + np.outer(j, j) / lam_ini
    assert np.abs(lhs - rhs).max() / np.abs(rhs).max() <= 1e-10
    # the step is the gain times the innovation: dw = P0 j xi / (lam + j.P0 j)
    T = o.eval(dims, w, x)[0]
    xi = val[0] - float(sd[0, 0] @ T[0])
    dw = p0 * j * xi / (lam_ini + p0 * (j @ j))
    assert np.abs((w1 - w) - dw).max() <= 1e-12 * max(1.0, np.abs(dw).max())
    assert lam == 0.99 * lam_ini + 1.0 - 0.99


def test_kf_iekf_one_iteration_is_plain(oracle_mod):
    """One IEKF iteration per observation equals the plain multi-residual
    update fed the residual the IEKF evaluates (kalman_module.f90:523-524)."""
    o = oracle_mod.Oracle(2, 3, SEEDS)
    dims = [2, 5, 4, 2]
    w, x, sd, val, rn = _obs(o, dims, 6)
    a = o.kf_seq(dims, w, 5.0, 0.98, 0.9987, False, 0.0, 0.0, 2, x, sd, val, rn, niter=1)
    # replay observation by observation through the plain update, forming
    # R = seed.T - val at the weights the previous step left
    wc, resid = w.copy(), []
    for n in range(6):
        T = o.eval(dims, wc, x[n:n + 1])[0]
        R = 0.0
        for jj in range(dims[-1]):
            for s in range(o.na):
                R = R + sd[n, jj, s] * T[jj, s]
        resid.append(R - val[n])
        wc = o.kf_seq(dims, w, 5.0, 0.98, 0.9987, False, 0.0, 0.0, 1, x[:n + 1], sd[:n + 1],
                      np.array(resid), rn[:n + 1])[0]
    b = o.kf_seq(dims, w, 5.0, 0.98, 0.9987, False, 0.0, 0.0, 1, x, sd, np.array(resid), rn)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_kf_single_block_decoupled_is_dense(oracle_mod):
    o = oracle_mod.Oracle(3, 2, None)
    dims = [3, 1]
    w, x, sd, val, rn = _obs(o, dims, 20)
    for gate, q in ((0.0, 0.0), (0.3, 1e-3)):
        a = o.kf_seq(dims, w, 7.0, 0.98, 0.999, False, gate, q, 0, x, sd, val)
        b = o.kf_seq(dims, w, 7.0, 0.98, 0.999, True, gate, q, 0, x, sd, val)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[3] == b[3]


KF_CASES = [(mode, dec, gate, q) for mode in (0, 1, 2, 3) for dec in (False, True)
            for gate, q in ((0.0, 0.0), (0.2, 1e-3))]


@pytest.mark.parametrize("mode,dec,gate,q", KF_CASES)
def test_kf_bitwise_vs_compiled_reference(oracle_mod, mode, dec, gate, q):
    if not oracle_mod.RefLib.available():
        pytest.skip("reference library not built")
    o = oracle_mod.Oracle(2, 3, SEEDS)
    dims = [2, 6, 5, 1] if mode in (0, 3) else [2, 6, 5, 2]
    w, x, sd, val, rn = _obs(o, dims, 12, seed=10 + mode)
    args = (10.0, 0.98, 0.9987, dec, gate, q, mode, x, sd, val, rn, 3 if mode == 2 else 1,
            TERMS if mode == 3 else ())
    a = o.kf_seq(dims, w, *args)
    b = oracle_mod.RefNet(2, 3, SEEDS, 0, dims, w).kf_seq(*args)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])
    assert a[2] == b[2] and a[3] == b[3]
    if gate > 0 and mode != 2:     # the gate trips on these observations (IEKF never gates)
        assert a[3] > 0


def test_committee(oracle_mod):
    o = oracle_mod.Oracle(2, 3, SEEDS)
    dims = [2, 8, 8, 1]
    rng = np.random.default_rng(5)
    W = rng.uniform(-0.6, 0.6, (5, o.nweights(dims)))
    x = rng.uniform(-1, 1, (7, 2))
    m, sd = o.comm_eval(dims, W, x)
    T = np.stack([o.eval(dims, W[k], x) for k in range(5)])
    assert np.abs(m - T.mean(0)).max() <= 1e-14 * np.abs(T).max()
    assert np.abs(sd - T.std(0, ddof=1)).max() <= 1e-12 * np.abs(T).max()
    # identical members: exactly zero spread; one member: zero by definition
    m2, sd2 = o.comm_eval(dims, np.stack([W[0]] * 4), x)
    assert np.array_equal(m2, T[0]) and not sd2.any()
    m1, sd1 = o.comm_eval(dims, W[:1], x)
    assert np.array_equal(m1, T[0]) and not sd1.any()
    if oracle_mod.RefLib.available():
        rm, rs = oracle_mod.RefNet(2, 3, SEEDS, 0, dims, W[0]).comm_eval(W, x)
        assert np.array_equal(m[:, 0, :], rm) and np.array_equal(sd[:, 0, :], rs)
