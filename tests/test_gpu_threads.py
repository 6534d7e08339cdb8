"""The reference's threading contract on the GPU library: evaluation and
gradient calls are re-entrant with one work handle and one accumulator per
host thread (lib/c_api_module.f90:21-23), and per-thread gradients summed in
a fixed order equal the serial sum (tools/example_train.f90:128-170, which
checks 1e-12).  ctypes releases the interpreter lock around each foreign
call, so the host threads below run inside the library at the same time.
"""
import threading

import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def D():
    import dnnf90_amd
    return dnnf90_amd


def _run_threads(fns):
    errs = []

    def wrap(f):
        try:
            f()
        except BaseException as e:   # surfaced after join
            errs.append(e)

    th = [threading.Thread(target=wrap, args=(f,)) for f in fns]
    for t in th:
        t.start()
    for t in th:
        t.join()
    if errs:
        raise errs[0]


def test_grad_point_four_threads(D, oracle_mod):
    """Four threads, each with its own tid and gid, run dnnf90_grad_point
    over a static share of the points; the fixed-order sum of the four
    accumulators against one serial accumulator and the oracle."""
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS["kdv"]
    npt, nth = 96, 4
    w, x = wl.inputs(npt)
    ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
    net = D.Net(ts, wl.dims)
    net.weights = w
    seed = np.random.default_rng(3).standard_normal((npt, ts.na))
    print("grad_point: kdv net, 96 points, 4 threads", flush=True)

    tws = [D.TWork(net) for _ in range(nth)]
    gs = [D.Grad(net) for _ in range(nth)]

    def worker(i):
        def f():
            for n in range(i, npt, nth):
                tws[i].grad_point(x[n], seed[n], gs[i])
        return f

    _run_threads([worker(i) for i in range(nth)])
    gsum = np.zeros(net.nw)
    for g in gs:
        gsum += g.get()

    tw, gser = D.TWork(net), D.Grad(net)
    for n in range(npt):
        tw.grad_point(x[n], seed[n], gser)
    gs1 = gser.get()
    o = oracle_mod.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
    gr = o.grad(wl.dims, w, x, seed[:, None, :])
    e_ser, e_or = rel(gsum, gs1), rel(gsum, gr)
    print("threaded vs serial", e_ser, "vs oracle", e_or)
    assert e_ser <= 1e-12 and e_or <= 1e-12
    for h in tws + gs + [tw, gser]:
        h.free()
    net.free()


def test_batched_work_spaces_concurrent(D, oracle_mod):
    """Host threads on networks of different widths and precisions at once
    (the launch-time LDS limits and occupancy caches are shared by every
    thread): each thread's collocation step, repeated, equals the same step
    run alone, bit for bit."""
    from dnnf90_amd.workloads import WORKLOADS
    cases = [("kdv", 0, 700), ("closure4_k7", 3, 300), ("ngd2_k4", 3, 500), ("closure4_k7", 0, 200)]
    setups = []
    for name, prec, npt in cases:
        wl = WORKLOADS[name]
        w, x = wl.inputs(npt)
        ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
        net = D.Net(ts, wl.dims)
        net.weights = w
        o = oracle_mod.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
        res = D.Residual(wl.terms(o.alpha()))
        setups.append((net, D.BatchWork(net, npt, prec), D.Grad(net), x, res, npt))
    print("concurrent batched steps:", cases, flush=True)

    def step(s):
        net, bw, g, x, res, npt = s
        g.zero()
        r = bw.residual_step(x, res, 2.0 / npt, g)
        return g.get(), r

    alone = [step(s) for s in setups]
    got = [None] * len(setups)

    def worker(i):
        def f():
            for _ in range(5):
                got[i] = step(setups[i])
        return f

    _run_threads([worker(i) for i in range(len(setups))])
    for (ga, ra), (gb, rb), c in zip(alone, got, cases):
        print(c, "max |diff|", float(np.abs(ga - gb).max()))
        assert np.array_equal(ga, gb) and ra == rb, c
    for net, bw, g, *_ in setups:
        bw.free()
        g.free()
        net.free()


def test_weight_update_waits_for_readers(D, oracle_mod):
    """A weight change while another work space of the same network still
    has kernels queued: the upload waits for them (the network's reader
    events), so the earlier forward sees the old weights and the later one
    the new."""
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS["closure4_k7"]
    npt = 2048
    w, x = wl.inputs(npt)
    ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
    net = D.Net(ts, wl.dims)
    net.weights = w
    o = oracle_mod.Oracle(wl.d0, wl.k, wl.seeds, wl.act)
    import torch
    xd = torch.from_numpy(x).cuda()
    b1, b2 = D.BatchWork(net, npt, 0), D.BatchWork(net, npt, 0)
    t1 = torch.empty((npt, 1, ts.na), dtype=torch.float64, device="cuda")
    b1.forward(xd, t1)                       # queued on b1's stream, not waited for
    w2 = w * 1.01
    net.weights = w2                         # host change; uploaded on the next call
    t2 = b2.forward(xd)
    b1.sync()
    b2.sync()
    sel = np.arange(0, npt, 64)
    r1 = o.eval(wl.dims, w, x[sel])
    r2 = o.eval(wl.dims, w2, x[sel])
    e1 = rel(t1.cpu().numpy()[sel], r1)
    e2 = rel(t2.cpu().numpy()[sel], r2)
    print("old-weight forward", e1, "new-weight forward", e2)
    assert e1 <= 1e-12 and e2 <= 1e-12
    b1.free()
    b2.free()
    net.free()
