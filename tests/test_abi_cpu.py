"""The C-ABI library on the host (no GPU): it loads, exports every symbol
include/*.h declares, builds tables identical to the oracle's, and honours
the reference's argument checks and return codes
(lib/c_api_module.f90:72-96, :100-205, :476-528)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def L():
    from dnnf90_amd._lib import LIBPATH, lib
    if not os.path.exists(LIBPATH):
        from dnnf90_amd import build
        build.build()
    return lib()


def declared_symbols():
    names = []
    for h in ("dnnf90.h", "dnnf90_amd.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names += re.findall(r"\b(dnnf90_\w+)\s*\(", src)
    return sorted(set(names))


def test_exports_every_declared_symbol(L):
    names = declared_symbols()
    assert len(names) >= 24
    for n in names:
        assert hasattr(L, n), n
    ref = re.findall(r"\b(dnnf90_\w+)\s*\(", re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "dnnf90.h")).read(), flags=re.S))
    assert len(set(ref)) == 24   # the reference's lib/dnnf90.h:16-54


def _lists(L, tid):
    c = np.zeros(4, np.int32)
    assert L.dnnf90_tabset_counts(tid, c.ctypes.data) == 0
    na, P, F, k = (int(v) for v in c)
    arr = [np.zeros(na, np.int32), np.zeros(na, np.int32), np.zeros(P, np.int32), np.zeros(P, np.int32),
           np.zeros(P), np.zeros(na * (k + 1), np.int32), np.zeros(na * (k + 1), np.int32),
           np.zeros(F, np.int32), np.zeros(F, np.int32), np.zeros(F)]
    assert L.dnnf90_tabset_lists(tid, *[a.ctypes.data for a in arr]) == 0
    return arr


TAB_CASES = [(2, 3, [[0, 1], [1, 0], [3, 0]]), (4, 3, None), (4, 7, [[7, 0, 0, 0], [1, 0, 0, 0], [0, 0, 0, 1], [3, 2, 2, 0]]),
             (3, 4, None), (2, 4, [[4, 0], [0, 2], [1, 0]]), (4, 7, None), (6, 3, None), (1, 15, None),
             (5, 5, [[2, 1, 0, 1, 1], [0, 0, 5, 0, 0]])]


@pytest.mark.parametrize("d0,k,seeds", TAB_CASES)
def test_tables_bitwise_vs_oracle(L, oracle_mod, d0, k, seeds):
    import dnnf90_amd as D
    ts = D.TableSet(d0, k, seeds)
    o = oracle_mod.Oracle(d0, k, seeds)
    assert ts.na == o.na
    got = _lists(L, ts.id)
    st, nm, ia, idd, b = o.pairs()
    fst, fnm, fib, fid, fc = o.fq()
    for g, r in zip(got, [st, nm, ia, idd, b, fst.reshape(-1), fnm.reshape(-1), fib, fid, fc]):
        assert np.array_equal(g, r)
    al = np.array([a[:d0] for a in D.current_alpha()])
    assert np.array_equal(al, o.alpha())


def test_force_field_sized_set(L):
    import dnnf90_amd as D
    ts = D.TableSet(100, 2)
    assert ts.na == 1 + 100 + 5050


def test_argument_checks(L):
    s = np.array([0, 1, 1, 0, 3, 0], np.int32)
    assert L.dnnf90_tables_init_dense(0, 3) == -1
    assert L.dnnf90_tables_init_dense(2, 16) == -1
    assert L.dnnf90_tables_init_dense(2, 0) == -1
    bad = np.array([0, 1, -1, 0], np.int32)
    assert L.dnnf90_tables_init_closure(2, 3, 2, bad.ctypes.data) == -1
    big = np.array([0, 4], np.int32)
    assert L.dnnf90_tables_init_closure(2, 3, 1, big.ctypes.data) == -1
    assert L.dnnf90_tables_init_closure(2, 3, 3, s.ctypes.data) == 0
    assert L.dnnf90_nderiv() == 5
    a = np.zeros(2, np.int32)
    assert L.dnnf90_alpha(0, a.ctypes.data) == -1
    assert L.dnnf90_alpha(6, a.ctypes.data) == -1
    assert L.dnnf90_alpha(5, a.ctypes.data) == 0 and list(a) == [3, 0]
    tid = L.dnnf90_tabset_dense(2, 3)
    assert tid > 0
    assert L.dnnf90_tabset_activation(tid, 9) == -2
    assert L.dnnf90_tabset_activation(tid + 50, 0) == -1
    assert L.dnnf90_tabset_nderiv(tid) == 10
    dims = np.array([2, 5, 1], np.int32)
    assert L.dnnf90_net_create_ts(tid, 1, dims.ctypes.data) == -1
    wrongd0 = np.array([3, 5, 1], np.int32)
    assert L.dnnf90_net_create_ts(tid, 3, wrongd0.ctypes.data) == -1
    zero = np.array([2, 0, 1], np.int32)
    assert L.dnnf90_net_create_ts(tid, 3, zero.ctypes.data) == -1
    assert L.dnnf90_net_load(b"/nonexistent/nn_weight.dat") == -2
    nid = L.dnnf90_net_create_ts(tid, 3, dims.ctypes.data)
    assert nid > 0
    row = np.zeros(5 * 3)
    assert L.dnnf90_net_set_layer(nid, 1, row.ctypes.data) == -1
    assert L.dnnf90_net_set_layer(nid, 4, row.ctypes.data) == -1
    assert L.dnnf90_net_set_layer(nid, 2, row.ctypes.data) == 0
    assert L.dnnf90_net_free(nid) == 0
    assert L.dnnf90_net_free(nid) == -1
    assert L.dnnf90_net_set_layer(nid, 2, row.ctypes.data) == -1
    # no work space, no output array
    out = np.zeros(8)
    assert L.dnnf90_bwork_step_output(0, out.ctypes.data, 0) == -1
    assert L.dnnf90_bwork_step_output(63, out.ctypes.data, 0) == -1
    assert L.dnnf90_bwork_step_output(1, None, 0) == -1
    # the fused kernels' switch returns its previous setting
    prev = L.dnnf90_set_fused_layers(0)
    assert L.dnnf90_set_fused_layers(prev) == 0


def test_weight_file_round_trip(L, tmp_path):
    """nn_weight.dat written by dnnf90_net_save is read back exactly, and the
    reference's own weight files load with their shapes."""
    import dnnf90_amd as D
    ts = D.TableSet(2, 3, [[0, 1], [1, 0], [3, 0]])
    net = D.Net(ts, [2, 7, 130, 1])   # a row longer than 100 values wraps
    w = np.random.default_rng(0).normal(size=net.nw)
    net.weights = w
    p = tmp_path / "nn_weight.dat"
    net.save(p)
    back = D.Net.load(p, ts)
    assert back.dims == [2, 7, 130, 1]
    assert np.array_equal(back.weights, w)
    assert np.array_equal(back.get_layer(3), w[7 * 3:7 * 3 + 130 * 8])
    # a file trained with another activation is refused against a tanh set
    erf = D.TableSet(2, 3, [[0, 1], [1, 0], [3, 0]], act=D.ACT_ERF)
    net2 = D.Net(erf, [2, 4, 1])
    q = tmp_path / "erf.dat"
    net2.save(q)
    with pytest.raises(D.DNNF90Error) as e:
        D.Net.load(q, ts)
    assert e.value.rc == -2
    assert D.Net.load(q, erf).dims == [2, 4, 1]


def test_tabset_free_returns_handles(L):
    """The pool holds 64 table sets (lib/c_api_module.f90:38); releasing them
    lets a host create more, and a network built from a released set keeps its
    tables."""
    import dnnf90_amd as D
    ts = D.TableSet(2, 3, [[0, 1], [1, 0], [3, 0]])
    dims = np.array([2, 8, 1], np.int32)
    nid = L.dnnf90_net_create_ts(ts.id, 3, dims.ctypes.data)
    assert nid > 0
    na = ts.na
    ts.free()
    assert ts.id == 0
    assert L.dnnf90_tabset_free(0) == -1
    assert L.dnnf90_net_nderiv(nid) == na
    for _ in range(3 * 64):
        t = D.TableSet(4, 3)
        assert t.id > 0
        t.free()
    assert L.dnnf90_net_free(nid) == 0


def test_filter_committee_value_argument_checks(L):
    """The extension entries refuse bad arguments before any device work
    (kf_init's own checks, lib/kalman_module.f90:118-128)."""
    import dnnf90_amd as D
    ts = D.TableSet(2, 3, [[0, 1], [1, 0], [3, 0]])
    dims = np.array([2, 4, 1], np.int32)
    nid = L.dnnf90_net_create_ts(ts.id, 3, dims.ctypes.data)
    assert nid > 0
    for p0, lam, lam0 in ((0.0, 1.0, 1.0), (-1.0, 1.0, 1.0), (float("nan"), 1.0, 1.0), (1.0, 0.0, 1.0),
                          (1.0, 1.5, 1.0), (1.0, 1.0, 0.0), (1.0, 1.0, 1.01)):
        assert L.dnnf90_kf_create(nid, p0, lam, lam0, 0) == -1, (p0, lam, lam0)
    assert L.dnnf90_kf_create(nid + 50, 1.0, 1.0, 1.0, 0) == -1
    assert L.dnnf90_kf_set(60, 0.0, 0.0) == -1
    assert L.dnnf90_kf_free(60) == -1
    assert L.dnnf90_kf_get(60, None, None) == -1
    assert L.dnnf90_kf_update_grad(nid, 60, 60, 0.0, 1.0) == -1
    assert L.dnnf90_kf_iekf_begin(nid, 60) == -1
    x = np.zeros(2)
    assert L.dnnf90_kf_update_obs(nid, 60, 60, 1, x.ctypes.data, x.ctypes.data, x.ctypes.data, None, 0) == -1
    ids = np.array([nid], np.int32)
    out = np.zeros(16)
    assert L.dnnf90_comm_eval_batch(0, ids.ctypes.data, 60, 1, x.ctypes.data, out.ctypes.data,
                                    out.ctypes.data, 0) == -1
    assert L.dnnf90_comm_eval_batch(1, ids.ctypes.data, 60, 1, x.ctypes.data, out.ctypes.data,
                                    out.ctypes.data, 0) == -1
    assert L.dnnf90_grad_batch(nid, 60, 1, x.ctypes.data, x.ctypes.data, 1.0, 60, 0) == -1
    assert L.dnnf90_tabset_write_order(60, b"/dev/null") == -1
    assert L.dnnf90_net_free(nid) == 0
    ts.free()


def test_switches_return_previous_setting(L):
    """dnnf90_set_codegen / dnnf90_set_store_s / dnnf90_set_split_kernels
    return the setting they replace (no device work); a bad work-space
    handle is refused by the kernel-timing entries."""
    prev = L.dnnf90_set_store_s(0)
    try:
        assert L.dnnf90_set_store_s(1) == 0
        assert L.dnnf90_set_store_s(1) == 1
    finally:
        L.dnnf90_set_store_s(prev)
    prev = L.dnnf90_set_codegen(0)
    try:
        assert L.dnnf90_set_codegen(1) == 0
    finally:
        L.dnnf90_set_codegen(prev)
    prev = L.dnnf90_set_split_kernels(0)
    try:
        assert L.dnnf90_set_split_kernels(1) == 0
        assert L.dnnf90_set_split_kernels(1) == 1
    finally:
        L.dnnf90_set_split_kernels(prev)
    assert L.dnnf90_bwork_profile(63, 1) == -1
    ms = np.zeros(16)
    cnt = np.zeros(16, np.int64)
    assert L.dnnf90_bwork_kernel_ms(63, ms.ctypes.data, cnt.ctypes.data) == -1


def test_fortran_modules_built():
    """The reference-named Fortran modules compile (amdflang) into the
    archive hosts link, with their .mod files; where the reference tree is
    present its host programs link against them unchanged."""
    import shutil
    from dnnf90_amd import build as b
    if not (shutil.which("amdflang") or os.path.exists("/opt/rocm/bin/amdflang")):
        pytest.skip("amdflang not installed")
    if not os.path.exists(b.FORTRAN_LIB):
        b.build_fortran()
    assert os.path.exists(b.FORTRAN_LIB)
    bdir = os.path.dirname(b.FORTRAN_LIB)
    for m in ("multi_index_bell_module", "net_module", "train_module", "dnnf90_capi"):
        assert os.path.exists(os.path.join(bdir, m + ".mod")), m
    if os.path.isdir(os.path.join(b.REFERENCE, "tools")):
        for t in b.REF_TOOLS:
            assert os.path.exists(os.path.join(b.REF_TOOL_DIR, t)), t
