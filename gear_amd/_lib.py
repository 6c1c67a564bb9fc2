"""ctypes binding of libgearhip.so (the C ABI declared in include/gearhip.h).

The library is built in-tree by gear_amd.build.  Loading fails loudly when it
is missing: there is no fallback path.
"""
import ctypes
import os

import torch

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("GEAR_AMD_LIB") or os.path.join(HERE, "libgearhip.so")

GH_OK, GH_EINVAL, GH_EUNSUPPORTED, GH_ELAUNCH = 0, 1, 2, 3
_ERRORS = {GH_EINVAL: "invalid argument", GH_EUNSUPPORTED: "unsupported size",
           GH_ELAUNCH: "kernel launch failed"}

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64

# name -> argtypes (restype int unless listed in _RESTYPES)
SIGNATURES = {
    "gh_version": [],
    "gh_qcqp_pack_bytes": [_I, _I],
    "gh_qcqp_pack_f64": [_I, _I, _P, _P, _P],
    "gh_qcqp_crown_f64": [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P],
    "gh_qcqp_pack_f32_bytes": [_I, _I],
    "gh_qcqp_pack_f32": [_I, _I, _P, _P, _P],
    "gh_qcqp_crown_f32s": [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P],
    "gh_branch_widest_f64": [_I, _I, _P, _P, _P, _P],
    "gh_split_f64": [_I, _I, _P, _P, _P, _P, _P, _P],
    "gh_filter_workspace_bytes": [_I],
    "gh_filter_append_f64": [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _L, _L, _P, _P, _P, _P],
    "gh_min_strided_f64": [_L, _L, _P, _P, _P],
    "gh_qcqp_large_plan_bytes": [_I, _I],
    "gh_qcqp_large_prepare": [_I, _I, _P, _P, _P],
    "gh_qcqp_large_workspace_bytes": [_I, _I, _I],
    "gh_qcqp_crown_f32s_large": [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, ctypes.c_size_t, _P],
    "gh_alm_plan_bytes": [_I, _I],
    "gh_alm_prepare": [_I, _I, _P, _P, _P],
    "gh_alm_trace_width": [_I],
    "gh_dag_plan_work": [_P, _P, _P],
    "gh_alm_pgd_f64": [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "gh_lp_dual_bound_f64": [_I, _I, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P],
    "gh_child_keys_f64": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P],
    "gh_branch_split_keys_f64": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "gh_constrained_concretize_f64": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _I, _P, _P, _P, _P],
    "gh_constrained_bound_f64": [_I, _I, _I, _I, _P, _P, _P, ctypes.c_int64, _P, ctypes.c_int64,
                                 _P, _P, _I, _P, _P, _P, _P],
    "gh_clip_domains_f64": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "gh_filter_payload_f64": [_I, _I, _I, _P, _P, _P, ctypes.c_int64, ctypes.c_int64, _P],
    "gh_frontier_keys_f64": [ctypes.c_int64, _I, _P, _P, _P, _P, _P],
    "gh_min_rows_f64": [ctypes.c_int64, ctypes.c_int64, _P, _P, _P, _P],
    "gh_gather_rows_f64": [ctypes.c_int64, _P, _I, _P, _P, _I, _P, _P, _I, _P, _P, _P],
    "gh_fill_order_i64": [_I, _P, ctypes.c_int64, ctypes.c_int64, _P, _P],
    "gh_merge_desc_f64": [ctypes.c_int64, _P, _P, ctypes.c_int64, _P, _P, _P, _P, _P],
    "gh_split_append_workspace_bytes": [_I],
    "gh_branch_sb_f64": [_I, _I, _I, _P, _P, _P, _P, _P, _P, ctypes.c_double, _P,
                         ctypes.c_double, _P, _P, _P],
    "gh_split_append_idx_f64": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _L,
                                _L, _P, _L, _P, _P, _P, _P],
    "gh_split_append_f64_noalloc": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _L, _L,
                                    _P, _L, _P, _P, _P],
    "gh_split_append_noalloc_max_parents": [],
    "gh_frontier_sort_f64": [_L, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    "gh_split_append_f64": [_I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _L, _L, _P,
                            _L, _P, _P, _P, _P],
    "gh_qcqp_crown_rows_f64": [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P],
    "gh_qcqp_crown_rows_f32s": [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I, _P],
    "gh_dag_plan_create": [_I, _P, ctypes.POINTER(_P)],
    "gh_dag_plan_destroy": [_P],
    "gh_dag_plan_info": [_P, _P, _P, _P],
    "gh_dag_plan_stats": [_P, _P, _P, _P],
    "gh_dag_workspace_bytes": [_P, _I],
    "gh_dag_crown_f64": [_P, _P, _P, ctypes.c_size_t, _I, _I, _P],
    "gh_dag_plan_aslots": [_P, _P],
}
_RESTYPES = {"gh_filter_workspace_bytes": ctypes.c_size_t,
             "gh_split_append_workspace_bytes": ctypes.c_size_t, "gh_qcqp_pack_bytes": ctypes.c_size_t,
             "gh_qcqp_pack_f32_bytes": ctypes.c_size_t, "gh_dag_workspace_bytes": ctypes.c_size_t,
             "gh_qcqp_large_plan_bytes": ctypes.c_size_t,
             "gh_alm_plan_bytes": ctypes.c_size_t,
             "gh_qcqp_large_workspace_bytes": ctypes.c_size_t}

_lib = None


class GearHipError(RuntimeError):
    pass


def lib():
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise GearHipError(
                f"{LIB_PATH} is missing: build it with `python -m gear_amd.build`")
        handle = ctypes.CDLL(LIB_PATH)
        for name, args in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, ctypes.c_int)
        _lib = handle
    return _lib


def check(code, what):
    if code != GH_OK:
        raise GearHipError(f"{what}: {_ERRORS.get(code, code)}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise GearHipError("gearhip entry points take device tensors")
    return t.data_ptr()


def stream_ptr(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
