"""ctypes wrappers over the oracle (TEST INFRASTRUCTURE ONLY).

`Oracle` drives liboracle.so, the C FP64 restatement of the reference's
table build, forward, adjoint, collocation seed, NGD solve, committee and
extended Kalman filter (lib/multi_index_bell_module.f90:128-351,
lib/train_module.f90:241-450, app/pinn_module.f90:32-65,
app/optimizer_module.f90:221-522, lib/committee_module.f90:79-111,
lib/kalman_module.f90:107-659).

`RefLib` drives _ref/libdnnf90_ref.so, the reference's own lib/ compiled
from its sources by `make ref`, through its dnnf90.h C ABI.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may
import this package; the product (dnnf90_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libdnnf90_ref.so")

_i = C.c_int
_d = C.c_double
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)

ACT_TANH, ACT_SIN, ACT_ERF, ACT_GELU = 0, 1, 2, 5


def build(ref: bool | None = None) -> None:
    """Compile liboracle.so (and the reference library when its tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if ref is None:
        ref = os.path.isdir("/root/reference/lib")
    if ref:
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True,
                       stdout=subprocess.DEVNULL)
        if os.path.isdir("/root/reference/app"):
            subprocess.run(["make", "-s", "-C", HERE, "trainer"], check=True,
                           stdout=subprocess.DEVNULL)


def _ip(a):
    return a.ctypes.data_as(_pi)


def _dp(a):
    return a.ctypes.data_as(_pd)


def _ia(x):
    return np.ascontiguousarray(x, dtype=np.int32)


def _da(x):
    return np.ascontiguousarray(x, dtype=np.float64)


class Oracle:
    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            if not os.path.exists(LIB):
                build(ref=False)
            L = C.CDLL(LIB)
            for name, argt in {
                "ot_tables_build": [_i, _i, _i, _pi],
                "ot_tables_free": [_i],
                "ot_set_activation": [_i, _i],
                "ot_na": [_i], "ot_nfq": [_i], "ot_npair": [_i],
                "ot_alpha": [_i, _pi], "ot_deg": [_i, _pi], "ot_ind_e1": [_i, _pi],
                "ot_pairs": [_i, _pi, _pi, _pi, _pi, _pd],
                "ot_fq": [_i, _pi, _pi, _pi, _pi, _pd],
                "ot_tpc": [_i, _pd],
                "ot_nweights": [_i, _pi],
                "ot_eval": [_i, _i, _pi, _pd, _i, _pd, _pd, _i],
                "ot_grad": [_i, _i, _pi, _pd, _i, _pd, _pd, _pd, _i],
                "ot_grad_rows": [_i, _i, _pi, _pd, _i, _pd, _pd, _pd, _i],
                "ot_pinn_step": [_i, _i, _pi, _pd, _i, _pd, _i, _pd, _pi, _pi, _d,
                                 _pd, _pd, _pd, _i],
                "ot_pinn_rows": [_i, _i, _pi, _pd, _i, _pd, _i, _pd, _pi, _pi, _pd, _i],
                "ot_ngd_solve": [_i, _i, _pd, _pd, _i, _d, _i, _pd],
                "ot_comm_eval": [_i, _i, _pi, _i, _pd, _i, _pd, _pd, _pd],
                "ot_kf_create": [_i, _pi, _d, _d, _d, _i],
                "ot_kf_set": [_i, _d, _d],
                "ot_kf_free": [_i],
                "ot_kf_update_grad": [_i, _pd, _pd, _d, _d],
                "ot_kf_iekf_begin": [_i, _pd],
                "ot_kf_iekf_iter": [_i, _pd, _pd, _d, _d, _i],
                "ot_kf_state": [_i, _pd, _pd, _pi],
            }.items():
                f = getattr(L, name)
                f.argtypes = argt
                f.restype = _i
            L.ot_act_derivs.argtypes = [_i, _d, _i, _pd]
            L.ot_act_derivs.restype = None
            L.ot_kf_cov_len.argtypes = [_i]
            L.ot_kf_cov_len.restype = C.c_long
            cls._lib = L
        return cls._lib

    def __init__(self, d0, k, seeds=None, act=ACT_TANH):
        L = self.lib()
        seeds = [] if seeds is None else seeds
        s = _ia(np.asarray(seeds).reshape(-1)) if len(seeds) else np.zeros(1, np.int32)
        h = L.ot_tables_build(d0, k, len(seeds), _ip(s))
        if h < 0:
            raise ValueError("oracle table build refused")
        self.h, self.d0, self.k = h, d0, k
        if act != ACT_TANH:
            if L.ot_set_activation(h, act) != 0:
                raise ValueError("bad activation")
        self.act = act
        self.na = L.ot_na(h)

    def __del__(self):
        try:
            self.lib().ot_tables_free(self.h)
        except Exception:
            pass

    # ---- tables ----
    def alpha(self):
        a = np.zeros((self.na, self.d0), np.int32)
        self.lib().ot_alpha(self.h, _ip(a))
        return a

    def deg(self):
        a = np.zeros(self.na, np.int32)
        self.lib().ot_deg(self.h, _ip(a))
        return a

    def ind_e1(self):
        a = np.zeros(self.d0, np.int32)
        self.lib().ot_ind_e1(self.h, _ip(a))
        return a

    def pairs(self):
        L = self.lib()
        n = L.ot_npair(self.h)
        st, nm = np.zeros(self.na, np.int32), np.zeros(self.na, np.int32)
        ia, idd, b = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
        L.ot_pairs(self.h, _ip(st), _ip(nm), _ip(ia), _ip(idd), _dp(b))
        return st, nm, ia, idd, b

    def fq(self):
        L = self.lib()
        n = L.ot_nfq(self.h)
        st = np.zeros((self.na, self.k + 1), np.int32)
        nm = np.zeros((self.na, self.k + 1), np.int32)
        ib, idd, c = np.zeros(n, np.int32), np.zeros(n, np.int32), np.zeros(n)
        L.ot_fq(self.h, _ip(st), _ip(nm), _ip(ib), _ip(idd), _dp(c))
        return st, nm, ib, idd, c

    def tpc(self):
        t = np.zeros((self.k + 2, self.k + 3))
        self.lib().ot_tpc(self.h, _dp(t))
        return t

    def act_derivs(self, a, nq):
        d = np.zeros(nq + 1)
        self.lib().ot_act_derivs(self.h, float(a), nq, _dp(d))
        return d

    # ---- network ----
    @staticmethod
    def nweights(dims):
        return int(sum(dims[l] * (dims[l - 1] + 1) for l in range(1, len(dims))))

    def eval(self, dims, w, x, nthreads=0):
        dims, w, x = _ia(dims), _da(w), _da(x)
        npt = x.shape[0]
        out = np.zeros((npt, int(dims[-1]), self.na))
        rc = self.lib().ot_eval(self.h, len(dims), _ip(dims), _dp(w), npt, _dp(x), _dp(out), nthreads)
        assert rc == 0
        return out

    def grad(self, dims, w, x, seed, nthreads=0):
        dims, w, x, seed = _ia(dims), _da(w), _da(x), _da(seed)
        g = np.zeros(self.nweights(dims))
        rc = self.lib().ot_grad(self.h, len(dims), _ip(dims), _dp(w), x.shape[0], _dp(x),
                                _dp(seed), _dp(g), nthreads)
        assert rc == 0
        return g

    def grad_rows(self, dims, w, x, seed, nthreads=0):
        dims, w, x, seed = _ia(dims), _da(w), _da(x), _da(seed)
        r = np.zeros((x.shape[0], self.nweights(dims)))
        rc = self.lib().ot_grad_rows(self.h, len(dims), _ip(dims), _dp(w), x.shape[0], _dp(x),
                                     _dp(seed), _dp(r), nthreads)
        assert rc == 0
        return r

    def pinn_step(self, dims, w, x, terms, fac, nthreads=0, want_t=False):
        """terms: list of (coeff, slot, nonlin).  Returns (grad, sum R^2, T)."""
        dims, w, x = _ia(dims), _da(w), _da(x)
        co = _da([t[0] for t in terms])
        sl = _ia([t[1] for t in terms])
        nl = _ia([1 if t[2] else 0 for t in terms])
        g = np.zeros(self.nweights(dims))
        rsq = C.c_double(0.0)
        tout = np.zeros((x.shape[0], int(dims[-1]), self.na)) if want_t else None
        rc = self.lib().ot_pinn_step(self.h, len(dims), _ip(dims), _dp(w), x.shape[0], _dp(x),
                                     len(terms), _dp(co), _ip(sl), _ip(nl), float(fac), _dp(g),
                                     C.byref(rsq), _dp(tout) if want_t else None, nthreads)
        assert rc == 0
        return g, rsq.value, tout

    def pinn_rows(self, dims, w, x, terms, nthreads=0):
        dims, w, x = _ia(dims), _da(w), _da(x)
        co = _da([t[0] for t in terms])
        sl = _ia([t[1] for t in terms])
        nl = _ia([1 if t[2] else 0 for t in terms])
        r = np.zeros((x.shape[0], self.nweights(dims)))
        rc = self.lib().ot_pinn_rows(self.h, len(dims), _ip(dims), _dp(w), x.shape[0], _dp(x),
                                     len(terms), _dp(co), _ip(sl), _ip(nl), _dp(r), nthreads)
        assert rc == 0
        return r

    @classmethod
    def ngd_solve(cls, rows, b, nb, mu, mode):
        rows, b = _da(rows), _da(b)
        u = np.zeros(rows.shape[1])
        rc = cls.lib().ot_ngd_solve(rows.shape[0], rows.shape[1], _dp(rows), _dp(b), nb,
                                    float(mu), mode, _dp(u))
        if rc != 0:
            raise ArithmeticError("metric numerically singular")
        return u


    def comm_eval(self, dims, W, x):
        """Committee mean and sample std of every slot, W[nmem][nw]."""
        dims, W, x = _ia(dims), _da(W), _da(x)
        shp = (x.shape[0], int(dims[-1]), self.na)
        m, sd = np.zeros(shp), np.zeros(shp)
        rc = self.lib().ot_comm_eval(self.h, len(dims), _ip(dims), W.shape[0], _dp(W), x.shape[0],
                                     _dp(x), _dp(m), _dp(sd))
        assert rc == 0
        return m, sd

    def kf_seq(self, dims, w, p0, lam_ini, lam0, decoupled, gate, q, mode, x, seed, val,
               rnoise=None, niter=1, terms=()):
        """A sequence of filter updates, observation by observation, in the
        four forms ref_driver.f90's refx_kf_seq drives on the reference:
        0 kf_update_obs (target val), 1 kf_update_resid_multi (resid val,
        rnoise), 2 IEKF on R = seed.T - val, 3 the collocation residual
        R = val + sum_k c_k [T_0 *] T_slot.  Returns (w, P, lam, ngated) with
        P as the filter holds it (column-major P or the Pb blocks)."""
        L = self.lib()
        dims, w = _ia(dims), _da(w).copy()
        x, seed, val = _da(x), _da(seed), _da(val)
        nobs = x.shape[0]
        rnoise = np.ones(nobs) if rnoise is None else _da(rnoise)
        h = L.ot_kf_create(len(dims), _ip(dims), p0, lam_ini, lam0, 1 if decoupled else 0)
        assert h > 0
        try:
            L.ot_kf_set(h, gate, q)
            for n in range(nobs):
                xn = x[n:n + 1]
                sn = seed[n:n + 1]
                if mode == 0:
                    T = self.eval(dims, w, xn)[0]
                    pred = 0.0
                    for a in range(self.na):
                        pred = pred + float(sn[0, 0, a]) * float(T[0, a])
                    g = self.grad(dims, w, xn, sn, nthreads=1)
                    rc = L.ot_kf_update_grad(h, _dp(w), _dp(g), float(val[n]) - pred, 1.0)
                elif mode == 1:
                    g = self.grad(dims, w, xn, sn, nthreads=1)
                    rc = L.ot_kf_update_grad(h, _dp(w), _dp(g), -float(val[n]),
                                             1.0 / np.sqrt(float(rnoise[n])))
                elif mode == 2:
                    L.ot_kf_iekf_begin(h, _dp(w))
                    for it in range(niter):
                        T = self.eval(dims, w, xn)[0]
                        R = 0.0
                        for j in range(T.shape[0]):
                            for a in range(self.na):
                                R = R + float(sn[0, j, a]) * float(T[j, a])
                        R = R - float(val[n])
                        g = self.grad(dims, w, xn, sn, nthreads=1)
                        rc = L.ot_kf_iekf_iter(h, _dp(w), _dp(g), R, float(rnoise[n]),
                                               1 if it == niter - 1 else 0)
                        if rc:
                            break
                elif mode == 3:
                    T = self.eval(dims, w, xn)[0]
                    R = float(val[n])
                    sv = np.zeros((1, T.shape[0], self.na))
                    for c, sl, nl in terms:
                        if nl:
                            R = R + c * float(T[0, 0]) * float(T[0, sl])
                            sv[0, 0, sl] += c * float(T[0, 0])
                            sv[0, 0, 0] += c * float(T[0, sl])
                        else:
                            R = R + c * float(T[0, sl])
                            sv[0, 0, sl] += c
                    g = self.grad(dims, w, xn, sv, nthreads=1)
                    rc = L.ot_kf_update_grad(h, _dp(w), _dp(g), 0.0 - R, 1.0)
                else:
                    raise ValueError("mode")
                if rc:
                    raise ArithmeticError(f"filter update refused ({rc})")
            P = np.zeros(L.ot_kf_cov_len(h))
            lam = C.c_double(0.0)
            ng = C.c_int(0)
            L.ot_kf_state(h, _dp(P), C.byref(lam), C.byref(ng))
            return w, P, lam.value, ng.value
        finally:
            L.ot_kf_free(h)


class RefLib:
    """The reference's lib/ (compiled from its own sources) through dnnf90.h."""

    _lib = None

    @classmethod
    def available(cls):
        return os.path.exists(REFLIB)

    @classmethod
    def lib(cls):
        if cls._lib is None:
            L = C.CDLL(REFLIB)
            sig = {
                "dnnf90_tables_init_dense": [_i, _i],
                "dnnf90_tables_init_closure": [_i, _i, _i, _pi],
                "dnnf90_nderiv": [],
                "dnnf90_alpha": [_i, _pi],
                "dnnf90_tabset_dense": [_i, _i],
                "dnnf90_tabset_closure": [_i, _i, _i, _pi],
                "dnnf90_tabset_nderiv": [_i],
                "dnnf90_tabset_activation": [_i, _i],
                "dnnf90_net_create_ts": [_i, _i, _pi],
                "dnnf90_net_load": [C.c_char_p],
                "dnnf90_net_load_ts": [_i, C.c_char_p],
                "dnnf90_net_create": [_i, _pi],
                "dnnf90_net_set_layer": [_i, _i, _pd],
                "dnnf90_net_free": [_i],
                "dnnf90_work_create": [_i],
                "dnnf90_work_free": [_i],
                "dnnf90_eval": [_i, _i, _pd, _pd],
                "dnnf90_twork_create": [_i],
                "dnnf90_twork_free": [_i],
                "dnnf90_grad_create": [_i],
                "dnnf90_grad_free": [_i],
                "dnnf90_grad_zero": [_i],
                "dnnf90_grad_point": [_i, _i, _pd, _pd, _i],
                "dnnf90_adam_step": [_i, _i, _d, _d, _d, _d, _i, _i],
            }
            for name, argt in sig.items():
                f = getattr(L, name)
                f.argtypes = argt
                f.restype = _i
            cls._lib = L
        return cls._lib


def ref_adam(dims, w, g, eta, beta1, beta2, eps, nbatch, istep0, nsteps):
    """nsteps of the reference's opt_adam_step with the fixed gradient g,
    through ref_driver.f90's refx_adam; returns the weights."""
    L = RefLib.lib()
    if not hasattr(L, "_refx_adam_ready"):
        L.refx_adam.argtypes = [_i, _pi, _pd, _pd, _d, _d, _d, _d, _i, _i, _i]
        L.refx_adam.restype = _i
        L._refx_adam_ready = True
    dims, w, g = _ia(dims), _da(np.array(w, copy=True)), _da(g)
    rc = L.refx_adam(len(dims), _ip(dims), _dp(w), _dp(g), eta, beta1, beta2, eps, nbatch, istep0, nsteps)
    assert rc == 0
    return w


class RefNet:
    """One network on the compiled reference, driven through ref_driver.f90."""

    def __init__(self, d0, k, seeds, act, dims, w):
        L = RefLib.lib()
        if not hasattr(L, "_refx_ready"):
            L.refx_eval.argtypes = [_i, _i, _i, _pi, _i, _i, _pi, _pd, _i, _pd, _pd]
            L.refx_grad.argtypes = [_i, _i, _i, _pi, _i, _i, _pi, _pd, _i, _pd, _pd, _pd]
            L.refx_eval.restype = L.refx_grad.restype = _i
            L._refx_ready = True
        self.L = L
        self.d0, self.k, self.act = d0, k, act
        self.seeds = _ia(np.asarray(seeds).reshape(-1)) if seeds else np.zeros(1, np.int32)
        self.nseed = len(seeds) if seeds else 0
        self.dims = _ia(dims)
        self.w = _da(w)
        self.na = Oracle(d0, k, seeds).na   # the slot count does not depend on the activation

    def eval(self, x):
        x = _da(x)
        out = np.zeros((x.shape[0], int(self.dims[-1]), self.na))
        rc = self.L.refx_eval(self.d0, self.k, self.nseed, _ip(self.seeds), self.act,
                              len(self.dims), _ip(self.dims), _dp(self.w), x.shape[0],
                              _dp(x), _dp(out))
        assert rc == 0
        return out

    def _kf_lib(self):
        L = self.L
        if not hasattr(L, "_refx_kf_ready"):
            L.refx_kf_seq.argtypes = [_i, _i, _i, _pi, _i, _i, _pi, _pd, _d, _d, _d, _i, _d, _d,
                                      _i, _i, _i, _pd, _pd, _pd, _pd, _i, _pd, _pi, _pi, _pd, _pd,
                                      _pd, _pi]
            L.refx_kf_seq.restype = _i
            L.refx_comm_eval.argtypes = [_i, _i, _i, _pi, _i, _i, _pi, _i, _pd, _i, _pd, _pd, _pd]
            L.refx_comm_eval.restype = _i
            L._refx_kf_ready = True
        return L

    def kf_seq(self, p0, lam_ini, lam0, decoupled, gate, q, mode, x, seed, val,
               rnoise=None, niter=1, terms=()):
        """refx_kf_seq: the same sequence as Oracle.kf_seq on the reference's
        own kalman_module.  Returns (w, P, lam, ngated)."""
        L = self._kf_lib()
        x, seed, val = _da(x), _da(seed), _da(val)
        nobs = x.shape[0]
        rnoise = np.ones(nobs) if rnoise is None else _da(rnoise)
        co = _da([t[0] for t in terms] or [0.0])
        sl = _ia([t[1] for t in terms] or [0])
        nl = _ia([1 if t[2] else 0 for t in terms] or [0])
        dims = [int(v) for v in self.dims]
        nw = Oracle.nweights(dims)
        if decoupled:
            maxb = max(dims[l - 1] + 1 for l in range(1, len(dims)))
            plen = maxb * maxb * sum(dims[1:])
        else:
            plen = nw * nw
        w, P = np.zeros(nw), np.zeros(plen)
        lam, ng = C.c_double(0.0), C.c_int(0)
        rc = L.refx_kf_seq(self.d0, self.k, self.nseed, _ip(self.seeds), self.act, len(dims),
                           _ip(self.dims), _dp(self.w), p0, lam_ini, lam0, 1 if decoupled else 0,
                           gate, q, mode, niter, nobs, _dp(x), _dp(seed), _dp(val), _dp(rnoise),
                           len(terms), _dp(co), _ip(sl), _ip(nl), _dp(w), _dp(P), C.byref(lam),
                           C.byref(ng))
        assert rc == 0
        return w, P, lam.value, ng.value

    def comm_eval(self, W, x):
        """refx_comm_eval: comm_eval of output 1 over members W[nmem][nw]."""
        L = self._kf_lib()
        W, x = _da(W), _da(x)
        m, sd = np.zeros((x.shape[0], self.na)), np.zeros((x.shape[0], self.na))
        rc = L.refx_comm_eval(self.d0, self.k, self.nseed, _ip(self.seeds), self.act,
                                   len(self.dims), _ip(self.dims), W.shape[0], _dp(W), x.shape[0],
                                   _dp(x), _dp(m), _dp(sd))
        assert rc == 0
        return m, sd

    def pinn_step(self, x, terms, fac):
        """refx_pinn_step: the collocation step on the reference's modules,
        point by point.  terms: (coeff, slot 0-based, nonlin).  Returns
        (grad, sum R^2, seconds of the point loop)."""
        L = self.L
        if not hasattr(L, "_refx_pinn_ready"):
            L.refx_pinn_step.argtypes = [_i, _i, _i, _pi, _i, _i, _pi, _pd, _i, _pd, _i, _pd, _pi, _pi, _d,
                                         _pd, C.POINTER(C.c_double), C.POINTER(C.c_double)]
            L.refx_pinn_step.restype = _i
            L._refx_pinn_ready = True
        x = _da(x)
        co = _da([t[0] for t in terms])
        sl = _ia([t[1] + 1 for t in terms])
        nl = _ia([1 if t[2] else 0 for t in terms])
        g = np.zeros(Oracle.nweights(self.dims))
        rsq, secs = C.c_double(0.0), C.c_double(0.0)
        rc = L.refx_pinn_step(self.d0, self.k, self.nseed, _ip(self.seeds), self.act, len(self.dims),
                              _ip(self.dims), _dp(self.w), x.shape[0], _dp(x), len(terms), _dp(co), _ip(sl),
                              _ip(nl), float(fac), _dp(g), C.byref(rsq), C.byref(secs))
        assert rc == 0
        return g, rsq.value, secs.value

    def grad(self, x, seed):
        x, seed = _da(x), _da(seed)
        g = np.zeros(Oracle.nweights(self.dims))
        rc = self.L.refx_grad(self.d0, self.k, self.nseed, _ip(self.seeds), self.act,
                              len(self.dims), _ip(self.dims), _dp(self.w), x.shape[0],
                              _dp(x), _dp(seed), _dp(g))
        assert rc == 0
        return g
