"""dnnf90_grad_point_noalloc: the one-point gradient with its planes in a
fixed device array instead of a work space, against dnnf90_grad_point (the
existing entry point on the same generated kernels) on two calls in a row,
the second on a smaller network (fewer scratch doubles, the array keeping
the first call's contents), and the refusals (a network outside the
one-point kernels, a network never prepared on the device)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(D, name, npts):
    from dnnf90_amd.workloads import WORKLOADS
    wl = WORKLOADS[name]
    w, x = wl.inputs(npts)
    ts = D.TableSet(wl.d0, wl.k, wl.seeds, wl.act)
    net = D.Net(ts, wl.dims)
    net.weights = w
    seed = np.random.default_rng(9).standard_normal((npts, ts.na))
    return ts, net, x, seed


def _reference(D, net, x, seed):
    tw, g = D.TWork(net), D.Grad(net)
    for n in range(x.shape[0]):
        tw.grad_point(x[n], seed[n], g)
    out = g.get()
    tw.free()
    g.free()
    return out


def test_noalloc_matches_grad_point():
    import torch
    import dnnf90_amd as D
    from dnnf90_amd._lib import lib
    L = lib()
    for name, npts in (("kdv", 5), ("kf_kdv", 3)):      # [2,32,32,32,1], then [2,16,16,1]
        print(f"no-alloc one-point gradient: {name}, {npts} points", flush=True)
        ts, net, x, seed = _case(D, name, npts)
        gref = _reference(D, net, x, seed)             # also prepares the network
        xd = torch.from_numpy(x).cuda()
        sd = torch.from_numpy(seed).cuda()
        gd = torch.zeros(net.nw, dtype=torch.float64, device="cuda")
        for n in range(npts):
            rc = L.dnnf90_grad_point_noalloc(net.id, xd[n].data_ptr(), sd[n].data_ptr(), gd.data_ptr(), None)
            assert rc == 0, L.dnnf90_last_error()
        torch.cuda.synchronize()
        g = gd.cpu().numpy()
        err = float(np.abs(g - gref).max() / np.abs(gref).max())
        print(f"  max rel diff to dnnf90_grad_point {err:.2e}", flush=True)
        assert err <= 1e-15
        net.free()
        ts.free()


def test_noalloc_refusals():
    import torch
    import dnnf90_amd as D
    from dnnf90_amd._lib import lib
    L = lib()
    xd = torch.zeros(4, dtype=torch.float64, device="cuda")
    # closure4_k7: width x slots = 64 x 41 above the one-point kernels' 1536
    ts, net, x, seed = _case(D, "closure4_k7", 1)
    tw = D.TWork(net)                                  # prepared, so the size check is what refuses
    sd = torch.zeros(ts.na, dtype=torch.float64, device="cuda")
    gd = torch.zeros(net.nw, dtype=torch.float64, device="cuda")
    assert L.dnnf90_grad_point_noalloc(net.id, xd.data_ptr(), sd.data_ptr(), gd.data_ptr(), None) == -1
    msg = L.dnnf90_last_error().decode()
    print("refused:", msg, flush=True)
    assert "outside the one-point kernels" in msg
    tw.free()
    net.free()
    # a network that no work space has prepared on the device
    ts2 = D.TableSet(2, 3, [[0, 1], [1, 0], [3, 0]], 0)
    net2 = D.Net(ts2, [2, 8, 1])
    g2 = torch.zeros(net2.nw, dtype=torch.float64, device="cuda")
    s2 = torch.zeros(ts2.na, dtype=torch.float64, device="cuda")
    rc = L.dnnf90_grad_point_noalloc(net2.id, xd.data_ptr(), s2.data_ptr(), g2.data_ptr(), None)
    print("unprepared:", rc, L.dnnf90_last_error(), flush=True)
    assert rc == -1
    net2.free()
