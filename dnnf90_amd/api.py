"""Python mirror of the C ABI (include/dnnf90.h, include/dnnf90_amd.h).

Names and argument meaning follow the reference's C interface
(lib/dnnf90.h; lib/c_api_module.f90:100-528) and its ctypes example
(tools/example_py.py); errors become DNNF90Error carrying the C return code
and message.  Slots are 1-based in the C ABI (dnnf90_alpha) and 0-based here.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

from ._lib import lib

PREC_FP64, PREC_FP32, PREC_BF16X3, PREC_BF16X6 = 0, 1, 2, 3
DEVICE_PTRS = 1
NGD_DUAL = 2
ACT_TANH, ACT_SIN, ACT_ERF, ACT_J0, ACT_J1, ACT_GELU = 0, 1, 2, 3, 4, 5


def _free_on_del(self):
    """__del__ of every handle class: the library holds 64 handles of each
    kind, so the slot goes back when the Python object goes.  An object built
    on a network holds a reference to it, so a network outlives its
    dependents.  Nothing is released at interpreter exit."""
    try:
        if not sys.is_finalizing():
            self.free()
    except Exception:
        pass


class DNNF90Error(RuntimeError):
    def __init__(self, rc, where):
        msg = lib().dnnf90_last_error().decode(errors="replace")
        super().__init__(f"{where} returned {rc}: {msg}")
        self.rc = rc


def _chk(rc, where, handle=False):
    if rc < 0 or (handle and rc == 0):
        raise DNNF90Error(rc, where)
    return rc


def _host(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data


def _ptr(a, dtype):
    """(keepalive, address, is_device) for a numpy array or a torch tensor."""
    if hasattr(a, "data_ptr"):
        import torch
        want = torch.float64 if dtype == np.float64 else (torch.float32 if dtype == np.float32 else torch.int32)
        if a.dtype != want or not a.is_contiguous():
            raise TypeError(f"tensor must be contiguous {want}")
        return a, a.data_ptr(), a.is_cuda
    arr, p = _host(a, dtype)
    return arr, p, False


class TableSet:
    """An independent multi-index table family (dnnf90_tabset_dense/closure)."""

    def __init__(self, d0, k, seeds=None, act=ACT_TANH):
        L = lib()
        if seeds is None or len(seeds) == 0:
            self.id = _chk(L.dnnf90_tabset_dense(d0, k), "dnnf90_tabset_dense", True)
        else:
            s, p = _host(np.asarray(seeds).reshape(-1), np.int32)
            self.id = _chk(L.dnnf90_tabset_closure(d0, k, len(seeds), p), "dnnf90_tabset_closure", True)
        if act != ACT_TANH:
            _chk(L.dnnf90_tabset_activation(self.id, act), "dnnf90_tabset_activation")
        self.d0, self.k, self.act = d0, k, act
        self.na = _chk(L.dnnf90_tabset_nderiv(self.id), "dnnf90_tabset_nderiv")

    def write_order(self, path):
        """Write hod_alpha_order.dat (the reference's slot-order file)."""
        _chk(lib().dnnf90_tabset_write_order(self.id, os.fsencode(path)), "dnnf90_tabset_write_order")

    def free(self):
        """Release the handle (networks built from the set stay valid)."""
        if getattr(self, "id", 0) > 0:
            lib().dnnf90_tabset_free(self.id)
            self.id = 0

    __del__ = _free_on_del


def current_alpha():
    """Multi-indices of the current family, [nderiv][d0] (dnnf90_alpha)."""
    L = lib()
    n = L.dnnf90_nderiv()
    out = []
    a = (C.c_int * 64)()
    for ia in range(1, n + 1):
        _chk(L.dnnf90_alpha(ia, C.addressof(a)), "dnnf90_alpha")
        out.append(list(a))
    return out


class Net:
    """A network bound to a table family; weights live on the host and the GPU."""

    def __init__(self, tabset: TableSet, dims, nid=None):
        L = lib()
        self.tabset = tabset
        if nid is None:
            d, p = _host(dims, np.int32)
            nid = _chk(L.dnnf90_net_create_ts(tabset.id, len(dims), p), "dnnf90_net_create_ts", True)
        self.id = nid
        nl = L.dnnf90_net_nlayer(nid)
        dd = np.zeros(nl, np.int32)
        L.dnnf90_net_dims(nid, dd.ctypes.data)
        self.dims = [int(v) for v in dd]
        self.na = L.dnnf90_net_nderiv(nid)
        self.nw = L.dnnf90_net_nweights(nid)

    @classmethod
    def load(cls, path, tabset: TableSet):
        nid = _chk(lib().dnnf90_net_load_ts(tabset.id, str(path).encode()), "dnnf90_net_load_ts", True)
        return cls(tabset, None, nid)

    def save(self, path):
        _chk(lib().dnnf90_net_save(self.id, str(path).encode()), "dnnf90_net_save")

    def set_layer(self, l, wrow):
        a, p = _host(wrow, np.float64)
        _chk(lib().dnnf90_net_set_layer(self.id, l, p), "dnnf90_net_set_layer")

    def get_layer(self, l):
        n = self.dims[l - 1] * (self.dims[l - 2] + 1)
        out = np.zeros(n)
        _chk(lib().dnnf90_net_get_layer(self.id, l, out.ctypes.data), "dnnf90_net_get_layer")
        return out

    @property
    def weights(self):
        out = np.zeros(self.nw)
        _chk(lib().dnnf90_net_get_weights(self.id, out.ctypes.data), "dnnf90_net_get_weights")
        return out

    @weights.setter
    def weights(self, w):
        a, p = _host(w, np.float64)
        if a.size != self.nw:
            raise ValueError(f"expected {self.nw} weights")
        _chk(lib().dnnf90_net_set_weights(self.id, p), "dnnf90_net_set_weights")

    def free(self):
        if getattr(self, "id", 0):
            lib().dnnf90_net_free(self.id)
            self.id = 0

    __del__ = _free_on_del


class Work:
    """Per-point evaluation work space (dnnf90_work_create / dnnf90_eval)."""

    def __init__(self, net: Net):
        self.net = net
        self.id = _chk(lib().dnnf90_work_create(net.id), "dnnf90_work_create", True)

    def eval(self, x):
        xa, p = _host(x, np.float64)
        t = np.zeros(self.net.na)
        _chk(lib().dnnf90_eval(self.net.id, self.id, p, t.ctypes.data), "dnnf90_eval")
        return t

    def free(self):
        if getattr(self, "id", 0):
            lib().dnnf90_work_free(self.id)
            self.id = 0

    __del__ = _free_on_del


class Grad:
    """Gradient accumulator (dnnf90_grad_create); lives on the GPU."""

    def __init__(self, net: Net):
        self.net = net
        self.id = _chk(lib().dnnf90_grad_create(net.id), "dnnf90_grad_create", True)

    def zero(self):
        _chk(lib().dnnf90_grad_zero(self.id), "dnnf90_grad_zero")

    def get(self):
        out = np.zeros(self.net.nw)
        _chk(lib().dnnf90_grad_get(self.id, out.ctypes.data), "dnnf90_grad_get")
        return out

    def set(self, g):
        a, p = _host(g, np.float64)
        _chk(lib().dnnf90_grad_set(self.id, p), "dnnf90_grad_set")

    def free(self):
        if getattr(self, "id", 0):
            lib().dnnf90_grad_free(self.id)
            self.id = 0

    __del__ = _free_on_del


class TWork:
    """Per-point training work space (dnnf90_twork_create / dnnf90_grad_point)."""

    def __init__(self, net: Net):
        self.net = net
        self.id = _chk(lib().dnnf90_twork_create(net.id), "dnnf90_twork_create", True)

    def grad_point(self, x, seed, grad: Grad):
        xa, px = _host(x, np.float64)
        sa, ps = _host(seed, np.float64)
        _chk(lib().dnnf90_grad_point(self.net.id, self.id, px, ps, grad.id), "dnnf90_grad_point")

    def free(self):
        if getattr(self, "id", 0):
            lib().dnnf90_twork_free(self.id)
            self.id = 0

    __del__ = _free_on_del


def adam_step(net: Net, grad: Grad, eta, beta1, beta2, eps, nbatch, istep):
    _chk(lib().dnnf90_adam_step(net.id, grad.id, eta, beta1, beta2, eps, nbatch, istep),
         "dnnf90_adam_step")


class Residual:
    """R = sum_k coeff_k * [T_0 *] T_slot_k on output 0 (slots 0-based here)."""

    def __init__(self, terms):
        self.coeff = np.ascontiguousarray([t[0] for t in terms], np.float64)
        self.slot = np.ascontiguousarray([t[1] + 1 for t in terms], np.int32)
        self.nonlin = np.ascontiguousarray([1 if t[2] else 0 for t in terms], np.int32)
        self.nterm = len(terms)


class BatchWork:
    """Batched work space of up to maxpt points (dnnf90_bwork_create).

    budget: device bytes; when keeping a plane (S or h) of every hidden
    layer would pass it, the reverse pass recomputes that plane instead
    (dnnf90_bwork_create_budget).  0 keeps everything."""

    def __init__(self, net: Net, maxpt, prec=PREC_FP64, budget=0):
        self.net, self.maxpt, self.prec = net, maxpt, prec
        if budget and budget > 0:
            self.id = _chk(lib().dnnf90_bwork_create_budget(net.id, maxpt, prec, float(budget)),
                           "dnnf90_bwork_create_budget", True)
        else:
            self.id = _chk(lib().dnnf90_bwork_create(net.id, maxpt, prec), "dnnf90_bwork_create", True)
        self.npt = 0

    def info(self):
        """{recompute, bytes, bytes_stored, maxpt, prec} (dnnf90_bwork_info)."""
        v = np.zeros(5)
        _chk(lib().dnnf90_bwork_info(self.id, v.ctypes.data), "dnnf90_bwork_info")
        return {"recompute": bool(v[0]), "bytes": float(v[1]), "bytes_stored": float(v[2]),
                "maxpt": int(v[3]), "prec": int(v[4])}

    def set_stream(self, stream_handle):
        _chk(lib().dnnf90_bwork_set_stream(self.id, stream_handle), "dnnf90_bwork_set_stream")

    def sync(self):
        _chk(lib().dnnf90_bwork_sync(self.id), "dnnf90_bwork_sync")

    KERNEL_CATEGORIES = ("linear", "bell_fwd", "bell_bwd", "grad", "linear_t", "other", "grad_reduce",
                         "kf_pass", "kf_gain")

    def profile(self, on=True, only=None):
        """Bracket every kernel category with HIP events (only: one category name)."""
        v = (2 + self.KERNEL_CATEGORIES.index(only) if only else 1) if on else 0
        _chk(lib().dnnf90_bwork_profile(self.id, v), "dnnf90_bwork_profile")

    def kernel_ms(self):
        """{category: (total ms, launches)} since the last call (resets)."""
        ms = np.zeros(len(self.KERNEL_CATEGORIES))
        cnt = np.zeros(len(self.KERNEL_CATEGORIES), np.int64)
        _chk(lib().dnnf90_bwork_kernel_ms(self.id, ms.ctypes.data, cnt.ctypes.data), "dnnf90_bwork_kernel_ms")
        return {c: (float(ms[i]), int(cnt[i])) for i, c in enumerate(self.KERNEL_CATEGORIES)}

    def forward(self, x, tout=None, want=True):
        """x [npt][d0]; returns T [npt][nout][na] (numpy, or the given tout).
        With device tensors the kernels run on the work space's stream: the
        result is complete after sync(), or in stream order on the stream
        given to set_stream()."""
        xk, px, dev = _ptr(x, np.float64)
        npt = int(x.shape[0])
        if want and tout is None:
            if dev:
                import torch
                tout = torch.empty((npt, self.net.dims[-1], self.net.na), dtype=torch.float64, device=x.device)
            else:
                tout = np.zeros((npt, self.net.dims[-1], self.net.na))
        pt = 0
        if tout is not None:
            tk, pt, tdev = _ptr(tout, np.float64)
            if tdev != dev:
                raise TypeError("x and tout must both be host or both be device")
        _chk(lib().dnnf90_forward_batch(self.net.id, self.id, npt, px, pt or None,
                                        DEVICE_PTRS if dev else 0), "dnnf90_forward_batch")
        self.npt = npt
        return tout

    def backward(self, seed, grad: Grad):
        """Adjoint of seed [npt][nout][na] after a forward of npt points on
        this work space; the gradient is added to grad."""
        shape = tuple(int(v) for v in seed.shape)
        if shape != (self.npt, self.net.dims[-1], self.net.na):
            raise ValueError(f"seed of shape {shape} after a forward of {self.npt} points")
        sk, ps, dev = _ptr(seed, np.float64)
        _chk(lib().dnnf90_backward_batch(self.net.id, self.id, ps, grad.id, DEVICE_PTRS if dev else 0),
             "dnnf90_backward_batch")

    def grad_batch(self, x, y, coef, grad: Grad):
        """Value-only batched step (net_grad_batch): gradient of
        sum_n coef/2 (u_n - y_n)^2 added to grad; single-output networks."""
        xk, px, xdev = _ptr(x, np.float64)
        yk, py, ydev = _ptr(y, np.float64)
        if xdev != ydev:
            raise TypeError("x and y must both be host or both be device")
        _chk(lib().dnnf90_grad_batch(self.net.id, self.id, int(x.shape[0]), px, py, float(coef), grad.id,
                                     DEVICE_PTRS if xdev else 0), "dnnf90_grad_batch")
        self.npt = 0   # the planes now hold the one-slot layout: a backward needs a new forward

    def residual_step(self, x, res: Residual, fac, grad: Grad, want_rsq=True):
        """Fused forward + collocation seed + adjoint; returns sum R^2."""
        xk, px, dev = _ptr(x, np.float64)
        rsq = C.c_double(0.0)
        L = lib()
        args = (res.coeff.ctypes.data, res.slot.ctypes.data, res.nonlin.ctypes.data)
        _chk(L.dnnf90_residual_step_batch(self.net.id, self.id, int(x.shape[0]), px, res.nterm, *args,
                                          float(fac), grad.id, C.addressof(rsq) if want_rsq else None,
                                          DEVICE_PTRS if dev else 0),
             "dnnf90_residual_step_batch")
        self.npt = int(x.shape[0])
        return rsq.value if want_rsq else None

    def step_output(self):
        """The output row [npt][nout][na] the last residual_step formed
        (dnnf90_bwork_step_output), as forward() returns it."""
        out = np.zeros((self.npt, self.net.dims[-1], self.net.na))
        _chk(lib().dnnf90_bwork_step_output(self.id, out.ctypes.data, 0), "dnnf90_bwork_step_output")
        return out

    def residual_rows(self, x, res: Residual, rows_t, ld):
        """Gauss-Newton rows dR/dw into the device FP32 buffer rows_t[nw][ld];
        ld >= npt rounded up to 64 (padding rows are written as zeros)."""
        xk, px, dev = _ptr(x, np.float64)
        rk, pr, rdev = _ptr(rows_t, np.float32)
        if not rdev:
            raise TypeError("rows_t must be a device tensor")
        args = (res.coeff.ctypes.data, res.slot.ctypes.data, res.nonlin.ctypes.data)
        _chk(lib().dnnf90_residual_rows_batch(self.net.id, self.id, int(x.shape[0]), px, res.nterm, *args,
                                              pr, int(ld), DEVICE_PTRS if dev else 0),
             "dnnf90_residual_rows_batch")
        self.npt = int(x.shape[0])

    def ngd_solve(self, rows_t, ld, nrow, b, nb, mu, mode=1, dual=False):
        """u = (R R^T/nb + cI)^{-1} (b/nb); c = mu (mode 1) or mu tr(G) (mode 0).
        dual: the same step through the nrow x nrow Gram matrix."""
        rk, pr, rdev = _ptr(rows_t, np.float32)
        if not rdev:
            raise TypeError("rows_t must be a device tensor")
        bk, pb, bdev = _ptr(b, np.float64)
        if bdev:
            import torch
            u = torch.empty_like(b)
            pu = u.data_ptr()
        else:
            u = np.zeros(self.net.nw)
            pu = u.ctypes.data
        _chk(lib().dnnf90_ngd_solve(self.net.id, self.id, int(nrow), pr, int(ld), pb, int(nb), float(mu),
                                    int(mode), pu, (DEVICE_PTRS if bdev else 0) | (NGD_DUAL if dual else 0)),
             "dnnf90_ngd_solve")
        return u

    def free(self):
        if getattr(self, "id", 0):
            lib().dnnf90_bwork_free(self.id)
            self.id = 0

    __del__ = _free_on_del


class KalmanFilter:
    """Extended Kalman filter over a network's weights (lib/kalman_module.f90),
    covariance on the device.  Method names follow the reference's kf_*
    routines; observations are processed in order within one call."""

    def __init__(self, net: Net, p0, lam_ini, lam0, decoupled=False, gate=0.0, q=0.0):
        self.net = net
        self.id = _chk(lib().dnnf90_kf_create(net.id, float(p0), float(lam_ini), float(lam0),
                                              1 if decoupled else 0), "dnnf90_kf_create", True)
        self.decoupled = bool(decoupled)
        if gate or q:
            self.set(gate, q)

    def set(self, gate=0.0, q=0.0):
        _chk(lib().dnnf90_kf_set(self.id, float(gate), float(q)), "dnnf90_kf_set")

    def update_grad(self, grad: Grad, xi, scale=1.0):
        """kf_update_grad: the row held by grad, residual xi."""
        _chk(lib().dnnf90_kf_update_grad(self.net.id, self.id, grad.id, float(xi), float(scale)),
             "dnnf90_kf_update_grad")

    def update_obs(self, bw: "BatchWork", x, seed, target, pred=None):
        """kf_update_obs over the observations x[n], seed[n][nout][na]."""
        xk, px, dev = _ptr(x, np.float64)
        sk, ps, sdev = _ptr(seed, np.float64)
        t, pt = _host(np.atleast_1d(target), np.float64)
        pr = _host(np.atleast_1d(pred), np.float64) if pred is not None else (None, None)
        _chk(lib().dnnf90_kf_update_obs(self.net.id, self.id, bw.id, int(x.shape[0]), px, ps, pt, pr[1],
                                        DEVICE_PTRS if dev else 0), "dnnf90_kf_update_obs")

    def update_resid(self, bw: "BatchWork", x, seed, resid, rnoise=None):
        """kf_update_resid_multi (kf_update_resid when rnoise is None)."""
        xk, px, dev = _ptr(x, np.float64)
        sk, ps, sdev = _ptr(seed, np.float64)
        r, pr = _host(np.atleast_1d(resid), np.float64)
        rn = _host(np.atleast_1d(rnoise), np.float64) if rnoise is not None else (None, None)
        _chk(lib().dnnf90_kf_update_resid(self.net.id, self.id, bw.id, int(x.shape[0]), px, ps, pr, rn[1],
                                          DEVICE_PTRS if dev else 0), "dnnf90_kf_update_resid")

    def update_residual(self, bw: "BatchWork", x, res: Residual, src=None):
        """The trainer's collocation observation: R = src + sum_k c_k [T_0 *] T_slot
        at the running weights, seed dR/dT, then kf_update_resid."""
        xk, px, dev = _ptr(x, np.float64)
        s = _host(np.atleast_1d(src), np.float64) if src is not None else (None, None)
        if dev and src is not None:
            raise TypeError("src must be on the device when x is")
        _chk(lib().dnnf90_kf_update_residual(self.net.id, self.id, bw.id, int(x.shape[0]), px, res.nterm,
                                             res.coeff.ctypes.data, res.slot.ctypes.data, res.nonlin.ctypes.data,
                                             s[1], DEVICE_PTRS if dev else 0), "dnnf90_kf_update_residual")

    def iekf_begin(self):
        _chk(lib().dnnf90_kf_iekf_begin(self.net.id, self.id), "dnnf90_kf_iekf_begin")

    def iekf_iter(self, bw: "BatchWork", x, seed, resid, rnoise, final):
        xk, px, dev = _ptr(x, np.float64)
        sk, ps, sdev = _ptr(seed, np.float64)
        _chk(lib().dnnf90_kf_iekf_iter(self.net.id, self.id, bw.id, px, ps, float(resid), float(rnoise),
                                       1 if final else 0, DEVICE_PTRS if dev else 0), "dnnf90_kf_iekf_iter")

    def state(self):
        info = np.zeros(10)
        _chk(lib().dnnf90_kf_get(self.id, None, info.ctypes.data), "dnnf90_kf_get")
        keys = ("lam", "lam0", "gate", "q", "ngated", "nw", "nb", "maxb", "bad", "bytes")
        return {k: (int(v) if k in ("ngated", "nw", "nb", "maxb", "bad") else float(v)) for k, v in zip(keys, info)}

    def covariance(self):
        """P(nw, nw) or Pb(maxb, maxb, nb), flattened column-major as the reference holds it."""
        st = self.state()
        n = st["maxb"] * st["maxb"] * st["nb"] if self.decoupled else st["nw"] * st["nw"]
        P = np.zeros(n)
        _chk(lib().dnnf90_kf_get(self.id, P.ctypes.data, None), "dnnf90_kf_get")
        return P

    def free(self):
        if getattr(self, "id", 0):
            lib().dnnf90_kf_free(self.id)
            self.id = 0

    __del__ = _free_on_del


class Committee:
    """Members of one shape (lib/committee_module.f90); eval returns the mean
    and the sample standard deviation of every slot (comm_eval)."""

    def __init__(self, nets):
        self.nets = list(nets)
        self.ids = np.ascontiguousarray([n.id for n in self.nets], np.int32)

    def eval(self, bw: "BatchWork", x):
        xk, px, dev = _ptr(x, np.float64)
        n0 = self.nets[0]
        shp = (int(x.shape[0]), n0.dims[-1], n0.na)
        if dev:
            import torch
            m = torch.empty(shp, dtype=torch.float64, device=x.device)
            s = torch.empty(shp, dtype=torch.float64, device=x.device)
            pm, ps = m.data_ptr(), s.data_ptr()
        else:
            m, s = np.zeros(shp), np.zeros(shp)
            pm, ps = m.ctypes.data, s.ctypes.data
        _chk(lib().dnnf90_comm_eval_batch(len(self.nets), self.ids.ctypes.data, bw.id, shp[0], px, pm, ps,
                                          DEVICE_PTRS if dev else 0), "dnnf90_comm_eval_batch")
        return m, s
