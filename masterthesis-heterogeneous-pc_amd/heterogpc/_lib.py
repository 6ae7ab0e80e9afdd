"""ctypes binding of libheterogpc_hip.so (include/heterogpc_hip.h).

The library is loaded from the package directory; there is no fallback: if it
is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "libheterogpc_hip.so")
HEADER = os.path.join(os.path.dirname(PKG_ROOT), "include", "heterogpc_hip.h")

HPC_OK = 0
HPC_EINVAL = -1
HPC_EHIP = -2
HPC_ENOMEM = -3
HPC_ESTATE = -4
HPC_ELEVEL = -5
HPC_CORR_CANON = 0
HPC_CORR_F32 = 1
HPC_TUNE_RANKS_PER_TASK = 1
HPC_TUNE_RECHECK_QUEUE = 2
HPC_TUNE_KERNEL_TIMING = 3
HPC_TUNE_LDS_LIMIT = 4
HPC_TUNE_WAIT_LIMIT = 5
HPC_TUNE_QUEUE = 6
HPC_TUNE_GRAPH = 7


class LevelStats(C.Structure):
    _fields_ = [
        ("level", C.c_int32),
        ("max_deg", C.c_int32),
        ("tests", C.c_int64),
        ("tests_max", C.c_int64),
        ("edges", C.c_int64),
        ("removed", C.c_int64),
        ("edges_left", C.c_int64),
        ("pinv_fallbacks", C.c_int64),
        ("ms_compact", C.c_double),
        ("ms_test", C.c_double),
        ("ms_commit", C.c_double),
        ("ms_wall", C.c_double),
        ("rechecks", C.c_int64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class RequestInfo(C.Structure):
    _fields_ = [
        ("p", C.c_int32),
        ("levels", C.c_int32),
        ("max_cond_size", C.c_int32),
        ("pad_", C.c_int32),
        ("tests", C.c_int64),
        ("queue_ms", C.c_double),
        ("latency_ms", C.c_double),
    ]


class ServerInfo(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("peak_active", C.c_int64),
        ("peak_bytes", C.c_uint64),
        ("budget", C.c_uint64),
        ("busy_ms", C.c_double),
        ("queued", C.c_int32),
        ("running", C.c_int32),
    ]


class TuneEntry(C.Structure):
    _fields_ = [
        ("level", C.c_int32),
        ("ranks_per_task", C.c_int32),
        ("tasks_per_block", C.c_int32),
        ("edges_per_block", C.c_int32),
        ("queue", C.c_int32),
        ("pad_", C.c_int32),
        ("ms_best", C.c_double),
        ("ms_default", C.c_double),
    ]


def tune_array(entries):
    """ctypes array of TuneEntry from dicts as autotune() returns them."""
    arr = (TuneEntry * max(len(entries), 1))()
    for k, e in enumerate(entries):
        for name, _ in TuneEntry._fields_:
            setattr(arr[k], name, e.get(name, 0))
    return arr, len(entries)


class HpcError(RuntimeError):
    pass


_vp, _i, _i64, _u64, _d = C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_double
_ps = C.POINTER(LevelStats)

SIGNATURES = {
    "hpc_last_error": (C.c_char_p, []),
    "hpc_device_count": (_i, [C.POINTER(C.c_int)]),
    "hpc_device_synchronize": (_i, [_i]),
    "hpc_fisher_r_crit": (_d, [_i, _d]),
    "hpc_sem_generate": (_i, [_i, _i, _d, _d, _d, _u64, _i, _vp, _vp]),
    "hpc_state_create": (_i, [_u64, _i, _d, _i, _i, C.POINTER(_vp)]),
    "hpc_state_destroy": (_i, [_vp]),
    "hpc_reset": (_i, [_vp]),
    "hpc_sync": (_i, [_vp]),
    "hpc_max_cond_size": (_i, [_vp]),
    "hpc_set_tuning": (_i, [_vp, _i, _i]),
    "hpc_autotune": (_i, [_vp, _i, C.POINTER(TuneEntry), _i, C.POINTER(C.c_int)]),
    "hpc_apply_tuning": (_i, [_vp, C.POINTER(TuneEntry), _i]),
    "hpc_pmcc": (_i, [_vp, _vp, _i]),
    "hpc_set_observations": (_i, [_vp, _vp]),
    "hpc_corr": (_i, [_vp, _i]),
    "hpc_set_corr": (_i, [_vp, _vp]),
    "hpc_corr_ms": (_i, [_vp, C.POINTER(_d), C.POINTER(_d)]),
    "hpc_compact": (_i, [_vp, C.POINTER(C.c_int)]),
    "hpc_test_level": (_i, [_vp, _i, _vp, _i, _ps]),
    "hpc_search_level": (_i, [_vp, _i, _vp, _i, _ps]),
    "hpc_commit_level": (_i, [_vp, _i, _ps]),
    "hpc_search_level_host": (_i, [_vp, _i, _vp, _i, _i, _ps]),
    "hpc_level": (_i, [_vp, _i, _ps]),
    "hpc_run": (_i, [_vp, _i, _ps, _i, C.POINTER(C.c_int)]),
    "hpc_get_cor": (_i, [_vp, _vp]),
    "hpc_get_adj": (_i, [_vp, _vp]),
    "hpc_get_compact": (_i, [_vp, _vp]),
    "hpc_get_sepsets": (_i, [_vp, _vp, _i]),
    "hpc_get_sepset_list": (_i, [_vp, _vp, _i64, C.POINTER(_i64)]),
    "hpc_get_first_pass": (_i, [_vp, _vp, _vp, C.POINTER(_i64)]),
    "hpc_footprint": (_i, [_u64, _i, _i, C.POINTER(_u64)]),
    "hpc_batch_create": (_i, [_i, C.POINTER(_vp)]),
    "hpc_batch_destroy": (_i, [_vp]),
    "hpc_batch_add": (_i, [_vp, _u64, _i, _d, _i, _vp, C.POINTER(C.c_int)]),
    "hpc_batch_remove": (_i, [_vp, _i]),
    "hpc_batch_step": (_i, [_vp, C.POINTER(C.c_int)]),
    "hpc_batch_run": (_i, [_vp, C.POINTER(C.c_int)]),
    "hpc_batch_status": (_i, [_vp, _i, C.POINTER(C.c_int), C.POINTER(C.c_int),
                              C.POINTER(C.c_int)]),
    "hpc_batch_get_stats": (_i, [_vp, _i, _ps, _i, C.POINTER(C.c_int)]),
    "hpc_batch_max_cond_size": (_i, [_vp, _i]),
    "hpc_batch_get_cor": (_i, [_vp, _i, _vp]),
    "hpc_batch_get_adj": (_i, [_vp, _i, _vp]),
    "hpc_batch_get_sepsets": (_i, [_vp, _i, _vp, _i]),
    "hpc_batch_get_sepset_list": (_i, [_vp, _i, _vp, _i64, C.POINTER(_i64)]),
    "hpc_batch_set_tuning": (_i, [_vp, _i, _i]),
    "hpc_batch_apply_tuning": (_i, [_vp, C.POINTER(TuneEntry), _i]),
    "hpc_batch_set_corr_mode": (_i, [_vp, _i, _i]),
    "hpc_server_set_corr_mode": (_i, [_vp, _i]),
    "hpc_server_apply_tuning": (_i, [_vp, C.POINTER(TuneEntry), _i]),
    "hpc_server_create": (_i, [_i, _u64, _i, C.POINTER(_vp)]),
    "hpc_server_destroy": (_i, [_vp]),
    "hpc_server_submit": (_i, [_vp, _u64, _i, _d, _i, _vp, C.POINTER(_i64)]),
    "hpc_server_submit_borrowed": (_i, [_vp, _u64, _i, _d, _i, _vp, C.POINTER(_i64)]),
    "hpc_server_wait": (_i, [_vp, _i64, _d, C.POINTER(C.c_int)]),
    "hpc_server_result": (_i, [_vp, _i64, _vp, _vp, _i, C.POINTER(RequestInfo)]),
    "hpc_server_sepset_list": (_i, [_vp, _i64, _vp, _i64, C.POINTER(_i64)]),
    "hpc_server_release": (_i, [_vp, _i64]),
    "hpc_server_stats": (_i, [_vp, C.POINTER(ServerInfo)]),
    "hpc_server_set_tuning": (_i, [_vp, _i, _i]),
    "hpc_pack_check": (_i, [_vp, _i]),
    "hpc_batch_inject_failure": (_i, [_vp, _i, _i]),
    "hpc_server_inject_failure": (_i, [_vp, _i64, _i]),
    "hpc_plan_tasks": (_i, [_u64, _i, _i, C.POINTER(_i64)]),
    "hpc_wait_word": (_i, [_vp, _u64, _d]),
    "hpc_corr_noalloc": (_i, [_i, _vp, _u64, _i, _vp, _vp]),
    "hpc_search_level_shared": (_i, [_vp, _i, _i, _ps, _ps]),
    "hpc_batch_autotune": (_i, [_vp, C.POINTER(TuneEntry), _i, C.POINTER(C.c_int)]),
    "hpc_graph_stats": (_i, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]),
}

_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HpcError(
            f"{path} is missing: build it with masterthesis-heterogeneous-pc_amd/build.py"
        )
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != HPC_OK:
        msg = load().hpc_last_error()
        raise HpcError(f"rc={rc}: {msg.decode() if msg else ''}")
