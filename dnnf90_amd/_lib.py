"""ctypes binding of libdnnf90_amd.so (the gfx950 library behind dnnf90.h).

The library is built in-tree by `python -m dnnf90_amd.build`.  There is no
fallback: if it is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libdnnf90_amd.so")

_i, _l, _d, _vp = C.c_int, C.c_long, C.c_double, C.c_void_p
_pi = C.POINTER(C.c_int)
_pd = C.POINTER(C.c_double)
_cs = C.c_char_p

# name -> (restype, argtypes); pointers are passed as c_void_p so that both
# numpy host buffers and device addresses go through unchanged
SIGNATURES = {
    # include/dnnf90.h (the reference ABI, lib/dnnf90.h:16-54)
    "dnnf90_tables_init_dense": (_i, [_i, _i]),
    "dnnf90_tables_init_closure": (_i, [_i, _i, _i, _vp]),
    "dnnf90_nderiv": (_i, []),
    "dnnf90_alpha": (_i, [_i, _vp]),
    "dnnf90_tabset_dense": (_i, [_i, _i]),
    "dnnf90_tabset_closure": (_i, [_i, _i, _i, _vp]),
    "dnnf90_tabset_nderiv": (_i, [_i]),
    "dnnf90_net_create_ts": (_i, [_i, _i, _vp]),
    "dnnf90_net_load": (_i, [_cs]),
    "dnnf90_tabset_activation": (_i, [_i, _i]),
    "dnnf90_net_load_ts": (_i, [_i, _cs]),
    "dnnf90_net_create": (_i, [_i, _vp]),
    "dnnf90_net_set_layer": (_i, [_i, _i, _vp]),
    "dnnf90_net_free": (_i, [_i]),
    "dnnf90_work_create": (_i, [_i]),
    "dnnf90_work_free": (_i, [_i]),
    "dnnf90_eval": (_i, [_i, _i, _vp, _vp]),
    "dnnf90_twork_create": (_i, [_i]),
    "dnnf90_twork_free": (_i, [_i]),
    "dnnf90_grad_create": (_i, [_i]),
    "dnnf90_grad_free": (_i, [_i]),
    "dnnf90_grad_zero": (_i, [_i]),
    "dnnf90_grad_point": (_i, [_i, _i, _vp, _vp, _i]),
    "dnnf90_adam_step": (_i, [_i, _i, _d, _d, _d, _d, _i, _i]),
    # include/dnnf90_amd.h (batched and device-pointer extensions)
    "dnnf90_last_error": (_cs, []),
    "dnnf90_tabset_counts": (_i, [_i, _vp]),
    "dnnf90_tabset_lists": (_i, [_i] + [_vp] * 10),
    "dnnf90_tabset_write_order": (_i, [_i, _cs]),
    "dnnf90_tabset_free": (_i, [_i]),
    "dnnf90_kf_create": (_i, [_i, _d, _d, _d, _i]),
    "dnnf90_kf_set": (_i, [_i, _d, _d]),
    "dnnf90_kf_free": (_i, [_i]),
    "dnnf90_kf_update_grad": (_i, [_i, _i, _i, _d, _d]),
    "dnnf90_kf_update_obs": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _i]),
    "dnnf90_kf_update_resid": (_i, [_i, _i, _i, _i, _vp, _vp, _vp, _vp, _i]),
    "dnnf90_kf_update_residual": (_i, [_i, _i, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _i]),
    "dnnf90_kf_iekf_begin": (_i, [_i, _i]),
    "dnnf90_kf_iekf_iter": (_i, [_i, _i, _i, _vp, _vp, _d, _d, _i, _i]),
    "dnnf90_kf_get": (_i, [_i, _vp, _vp]),
    "dnnf90_comm_eval_batch": (_i, [_i, _vp, _i, _i, _vp, _vp, _vp, _i]),
    "dnnf90_set_codegen": (_i, [_i]),
    "dnnf90_set_store_s": (_i, [_i]),
    "dnnf90_set_point_kernels": (_i, [_i]),
    "dnnf90_set_split_kernels": (_i, [_i]),
    "dnnf90_set_edge_kernels": (_i, [_i]),
    "dnnf90_net_kernel_info": (_i, [_i, _i, _vp]),
    "dnnf90_net_step_info": (_i, [_i, _i, _vp]),
    "dnnf90_set_fused_grad": (_i, [_i]),
    "dnnf90_set_fused_layers": (_i, [_i]),
    "dnnf90_net_fused_info": (_i, [_i, _vp]),
    "dnnf90_set_poison": (_i, [_i]),
    "dnnf90_net_nweights": (_i, [_i]),
    "dnnf90_net_nlayer": (_i, [_i]),
    "dnnf90_net_dims": (_i, [_i, _vp]),
    "dnnf90_net_nderiv": (_i, [_i]),
    "dnnf90_net_get_layer": (_i, [_i, _i, _vp]),
    "dnnf90_net_get_weights": (_i, [_i, _vp]),
    "dnnf90_net_set_weights": (_i, [_i, _vp]),
    "dnnf90_net_save": (_i, [_i, _cs]),
    "dnnf90_grad_get": (_i, [_i, _vp]),
    "dnnf90_grad_set": (_i, [_i, _vp]),
    "dnnf90_bwork_create": (_i, [_i, _i, _i]),
    "dnnf90_bwork_create_budget": (_i, [_i, _i, _i, _d]),
    "dnnf90_bwork_info": (_i, [_i, _vp]),
    "dnnf90_bwork_step_output": (_i, [_i, _vp, _i]),
    "dnnf90_bwork_free": (_i, [_i]),
    "dnnf90_bwork_set_stream": (_i, [_i, _vp]),
    "dnnf90_bwork_sync": (_i, [_i]),
    "dnnf90_bwork_profile": (_i, [_i, _i]),
    "dnnf90_bwork_kernel_ms": (_i, [_i, _vp, _vp]),
    "dnnf90_forward_batch": (_i, [_i, _i, _i, _vp, _vp, _i]),
    "dnnf90_backward_batch": (_i, [_i, _i, _vp, _i, _i]),
    "dnnf90_grad_batch": (_i, [_i, _i, _i, _vp, _vp, _d, _i, _i]),
    "dnnf90_residual_step_batch": (_i, [_i, _i, _i, _vp, _i, _vp, _vp, _vp, _d, _i, _vp, _i]),
    "dnnf90_residual_rows_batch": (_i, [_i, _i, _i, _vp, _i, _vp, _vp, _vp, _vp, _l, _i]),
    "dnnf90_ngd_solve": (_i, [_i, _i, _i, _vp, _l, _vp, _i, _d, _i, _vp, _i]),
    "dnnf90_grad_point_noalloc": (_i, [_i, _vp, _vp, _vp, _vp]),
}

REFERENCE_ABI = [n for n in SIGNATURES if not n.startswith("dnnf90_last_error")][:24]

_LIB = None


def lib():
    """The loaded library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIBPATH):
            raise RuntimeError(
                f"{LIBPATH} is missing: build it with `python -m dnnf90_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _LIB = L
    return _LIB
