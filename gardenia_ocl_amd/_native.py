"""ctypes binding of libgardenia_hip.so (C ABI declared in include/*.h).

The library is built in-tree by ``__graft_entry__.build()`` (make -C
gardenia_ocl_amd/csrc).  There is no fallback: if the library is missing or a
call fails, a :class:`GardeniaError` is raised.
"""
from __future__ import annotations

import ctypes
import os
import threading

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libgardenia_hip.so")

GD_OK = 0
GD_ERR_INVALID = 1
GD_ERR_ALLOC = 2
GD_ERR_DEVICE = 3
GD_ERR_IO = 4
GD_ERR_NO_DEVICE = 5

_STATUS_NAMES = {
    GD_ERR_INVALID: "GD_ERR_INVALID",
    GD_ERR_ALLOC: "GD_ERR_ALLOC",
    GD_ERR_DEVICE: "GD_ERR_DEVICE",
    GD_ERR_IO: "GD_ERR_IO",
    GD_ERR_NO_DEVICE: "GD_ERR_NO_DEVICE",
}


class GardeniaError(RuntimeError):
    def __init__(self, status: int, func: str, msg: str):
        self.status = status
        super().__init__(f"{func} failed with {_STATUS_NAMES.get(status, status)}: {msg}")


class HostCsr(ctypes.Structure):
    _fields_ = [
        ("m", ctypes.c_int32),
        ("nnz", ctypes.c_int64),
        ("row_offsets", ctypes.POINTER(ctypes.c_int32)),
        ("column_indices", ctypes.POINTER(ctypes.c_int32)),
        ("weight", ctypes.POINTER(ctypes.c_float)),
        ("degree", ctypes.POINTER(ctypes.c_int32)),
    ]


class PrOptions(ctypes.Structure):
    _fields_ = [
        ("damping", ctypes.c_float),
        ("tol", ctypes.c_double),
        ("max_iter", ctypes.c_int),
        ("seed_vertex", ctypes.c_int),
        ("chunk", ctypes.c_int),
        ("profile", ctypes.c_int),
    ]


class PrStats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("final_error", ctypes.c_double),
        ("solve_ms", ctypes.c_float),
        ("pull_ms", ctypes.c_float),
        ("pull_launches", ctypes.c_int),
        ("near_tie", ctypes.c_int),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class PrPlanConfig(ctypes.Structure):
    _fields_ = [("relabel", ctypes.c_int), ("items_per_thread", ctypes.c_int),
                ("nontemporal", ctypes.c_int), ("persistent", ctypes.c_int),
                ("sort_tiles", ctypes.c_int), ("block", ctypes.c_int),
                ("hub_threshold", ctypes.c_int), ("stage_hubs", ctypes.c_int)]


class PrPlanInfo(ctypes.Structure):
    _fields_ = [
        ("n_tiles", ctypes.c_int64),
        ("n_hubs", ctypes.c_int64),
        ("n_hub_segments", ctypes.c_int64),
        ("max_row_len", ctypes.c_int32),
        ("relabel", ctypes.c_int32),
        ("items_per_thread", ctypes.c_int32),
        ("nontemporal", ctypes.c_int32),
        ("persistent", ctypes.c_int32),
        ("pull_grid", ctypes.c_int32),
        ("plan_ms", ctypes.c_float),
        ("sort_tiles", ctypes.c_int32),
        ("block", ctypes.c_int32),
        ("hub_threshold", ctypes.c_int32),
        ("staged_hubs", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


class MultiPrConfig(ctypes.Structure):
    _fields_ = [("use_graph", ctypes.c_int), ("chunk", ctypes.c_int)]


class SymgsPlanInfo(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int32), ("n_colors", ctypes.c_int32),
                ("offdiag_nnz", ctypes.c_int64), ("n_tiles", ctypes.c_int32),
                ("n_exact_rows", ctypes.c_int32), ("max_row_len", ctypes.c_int32),
                ("plan_ms", ctypes.c_double)]


class DeviceCsr(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int32), ("nnz", ctypes.c_int64), ("row_offsets", ctypes.c_void_p),
                ("column_indices", ctypes.c_void_p), ("weight", ctypes.c_void_p),
                ("degree", ctypes.c_void_p)]


class PprPlanConfig(ctypes.Structure):
    _fields_ = [("relabel", ctypes.c_int), ("k_chunk", ctypes.c_int),
                ("column_stop", ctypes.c_int), ("compact", ctypes.c_int),
                ("hub_cols", ctypes.c_int)]


class PprPlanInfo(ctypes.Structure):
    _fields_ = [
        ("n_tiles", ctypes.c_int64),
        ("n_hubs", ctypes.c_int64),
        ("k", ctypes.c_int32),
        ("k_chunk", ctypes.c_int32),
        ("relabel", ctypes.c_int32),
        ("grid", ctypes.c_int32),
        ("plan_ms", ctypes.c_float),
        ("hub_cols", ctypes.c_int32),
    ]

    def as_dict(self) -> dict:
        return {f: getattr(self, f) for f, _ in self._fields_}


_P = ctypes.c_void_p
_I32P = ctypes.POINTER(ctypes.c_int32)
_F32P = ctypes.POINTER(ctypes.c_float)
_F64P = ctypes.POINTER(ctypes.c_double)
_I64P = ctypes.POINTER(ctypes.c_int64)
_i = ctypes.c_int
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_d = ctypes.c_double

# name -> (restype, argtypes).  Every symbol declared in include/*.h.
SIGNATURES = {
    # gardenia_graph.h
    "gd_csr_free": (None, [ctypes.POINTER(HostCsr)]),
    "gd_csr_from_edges": (_i, [_i, _i64, _P, _P, _P, _i, _i, _i, _i, _i, _i, ctypes.POINTER(HostCsr)]),
    "gd_read_mtx": (_i, [ctypes.c_char_p, _i, _i, _i, _i, _i, ctypes.POINTER(HostCsr)]),
    "gd_read_gr": (_i, [ctypes.c_char_p, _i, _i, _i, _i, _i, ctypes.POINTER(HostCsr)]),
    "gd_read_dotgraph": (_i, [ctypes.c_char_p, _i, _i, _i, _i, _i, ctypes.POINTER(HostCsr)]),
    "gd_read_graph": (_i, [ctypes.c_char_p, _i, _i, _i, _i, _i, ctypes.POINTER(HostCsr)]),
    "gd_color_first_fit": (_i, [_i, _P, _P, _P, ctypes.POINTER(ctypes.c_int32)]),
    "gd_color_order": (_i, [_i, _P, ctypes.c_int32, _P, _P]),
    "gd_rmat_graph": (_i, [_i, _i, _d, _d, _d, _u64, ctypes.POINTER(HostCsr), ctypes.POINTER(HostCsr)]),
    "gd_rmat_edges": (_i, [_i, _i, _d, _d, _d, _u64, _i64, _i64, _P, _P]),
    "gd_er_matrix": (_i, [_i, _i, _u64, ctypes.POINTER(HostCsr)]),
    "gd_fill_uniform_pm1": (_i, [_u64, _i64, _P]),
    "gd_uniform_pm1": (ctypes.c_float, [_u64, _u64]),
    # gardenia_hip.h
    "gd_last_error": (ctypes.c_char_p, []),
    "gd_device_count": (_i, []),
    "gd_pr_default_options": (None, [ctypes.POINTER(PrOptions)]),
    "gd_pr_default_plan_config": (None, [ctypes.POINTER(PrPlanConfig)]),
    "gd_pr_plan_create_ex": (_i, [_i, _i64, _P, _P, _P, _P, ctypes.POINTER(PrPlanConfig),
                                  ctypes.POINTER(_P)]),
    "gd_pr_plan_create": (_i, [_i, _i64, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "gd_pr_plan_run": (_i, [_P, _P, ctypes.POINTER(PrOptions), ctypes.POINTER(PrStats), _P]),
    "gd_pr_plan_info": (_i, [_P, ctypes.POINTER(PrPlanInfo)]),
    "gd_pr_plan_destroy": (_i, [_P]),
    "gd_pr_solve": (_i, [_i, _i, _P, _P, _P, _P, _P, _P, ctypes.POINTER(PrOptions),
                         ctypes.POINTER(PrStats), _P]),
    "gd_pr_run_noalloc": (_i, [_i, _i64, _P, _P, _P, _P, ctypes.POINTER(PrOptions), _P,
                               ctypes.POINTER(PrStats), _P]),
    "gd_ppr_default_plan_config": (None, [ctypes.POINTER(PprPlanConfig)]),
    "gd_ppr_plan_create": (_i, [_i, _i64, _P, _P, _P, _i, _P, ctypes.POINTER(PprPlanConfig),
                                ctypes.POINTER(_P)]),
    "gd_ppr_plan_run": (_i, [_P, _P, _P, ctypes.POINTER(PrOptions), ctypes.POINTER(PrStats), _P,
                             _P]),
    "gd_csr_from_edges_device": (_i, [_i, _i64, _P, _P, _P, _i, _i, _i, _i, _i, _i, _P,
                                      ctypes.POINTER(DeviceCsr)]),
    "gd_device_csr_copy": (_i, [ctypes.POINTER(DeviceCsr), _P, _P, _P, _P, _P]),
    "gd_device_csr_free": (_i, [ctypes.POINTER(DeviceCsr)]),
    "gd_pr_verify": (_i, [_i, _i64, _P, _P, _P, _P, ctypes.c_float, _P, _F64P]),
    "gd_pr_push_solve": (_i, [_i, _i64, _P, _P, _P, ctypes.POINTER(PrOptions), _P,
                              ctypes.POINTER(PrStats), _P]),
    "gd_symgs_plan_create": (_i, [_i, _i64, _P, _P, _P, _P, _P, _i, _P, ctypes.POINTER(_P)]),
    "gd_symgs_plan_run": (_i, [_P, _P, _P, _i, _P]),
    "gd_symgs_plan_info": (_i, [_P, ctypes.POINTER(SymgsPlanInfo)]),
    "gd_symgs_plan_destroy": (_i, [_P]),
    "gd_symgs_solve": (_i, [_i, _i, _P, _P, _P, _P, _P, _P, _P, _i, ctypes.POINTER(ctypes.c_float)]),
    "gd_ppr_plan_info": (_i, [_P, ctypes.POINTER(PprPlanInfo)]),
    "gd_ppr_plan_column_iterations": (_i, [_P, _P]),
    "gd_ppr_plan_destroy": (_i, [_P]),
    "gd_multi_pr_default_config": (None, [ctypes.POINTER(MultiPrConfig)]),
    "gd_multi_pr_plan_create": (_i, [_i, _P, _P, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "gd_multi_pr_run": (_i, [_P, _P, ctypes.POINTER(PrOptions), ctypes.POINTER(MultiPrConfig),
                             _P, _P, ctypes.POINTER(ctypes.c_float)]),
    "gd_multi_pr_plan_destroy": (_i, [_P]),
    "gd_wpr_plan_create": (_i, [_i, _i64, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "gd_wpr_forward": (_i, [_P, _P, ctypes.POINTER(PrOptions), ctypes.POINTER(PrStats), _P]),
    "gd_wpr_backward": (_i, [_P, _P, _P, _P, _P, ctypes.POINTER(PrOptions),
                             ctypes.POINTER(PrStats)]),
    "gd_wpr_plan_destroy": (_i, [_P]),
    "gd_spmv_bwd_plan_create": (_i, [_i, _i, _i64, _P, _P, _P, ctypes.POINTER(_P)]),
    "gd_spmv_bwd_run": (_i, [_P, _P, _P, _P, _P, _P]),
    "gd_spmv_bwd_plan_destroy": (_i, [_P]),
    "gd_spmv_plan_create": (_i, [_i, _i, _i64, _P, _P, _P, _P, ctypes.POINTER(_P)]),
    "gd_spmv_plan_create_ex": (_i, [_i, _i, _i64, _P, _P, _P, _P, _i, ctypes.POINTER(_P)]),
    "gd_spmv_plan_mode": (_i, [_P]),
    "gd_spmv_plan_refresh_values": (_i, [_P]),
    "gd_spmv_plan_run": (_i, [_P, _P, _P]),
    "gd_spmv_plan_info": (_i, [_P, _I64P, _I64P, _I32P]),
    "gd_spmv_plan_destroy": (_i, [_P]),
    "gd_spmv_solve": (_i, [_i, _i, _P, _P, _P, _P, _P, _F32P]),
}

_lock = threading.Lock()
_lib = None


def lib() -> ctypes.CDLL:
    """Load the native library once; raise if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise GardeniaError(
                    GD_ERR_IO, "load",
                    f"{LIB_PATH} is missing; build it with `python -c 'import __graft_entry__ as g; g.build()'`",
                )
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _lib = handle
        return _lib


def check(status: int, func: str) -> None:
    if status != GD_OK:
        msg = lib().gd_last_error()
        raise GardeniaError(status, func, msg.decode() if msg else "")


def call(name: str, *args) -> int:
    """Call a status-returning entry point and raise on failure."""
    st = getattr(lib(), name)(*args)
    check(st, name)
    return st
