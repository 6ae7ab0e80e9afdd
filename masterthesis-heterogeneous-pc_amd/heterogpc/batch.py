"""numpy-facing wrappers of hpc_batch (many skeletons stepped together) and
hpc_server (continuous batching with HBM-footprint admission)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import LevelStats, RequestInfo, ServerInfo, check, load, tune_array
from .skeleton import _fetch_list


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def pack_check(ps) -> int:
    """0 when instances of these p can share one batch or server step, else
    the HPC_E* code (hpc_pack_check)."""
    arr = np.ascontiguousarray(ps, dtype=np.uint64)
    return load().hpc_pack_check(_ptr(arr) if len(arr) else None, len(arr))


def plan_tasks(deg: int, level: int, ranks_per_task: int = 512) -> int:
    """Search tasks of a row of `deg` neighbours at `level` (hpc_plan_tasks);
    raises HpcError when the level cannot be planned."""
    out = C.c_int64()
    check(load().hpc_plan_tasks(int(deg), int(level), int(ranks_per_task), C.byref(out)))
    return int(out.value)


def footprint(p: int, n: int, max_level: int) -> int:
    """Device bytes one instance of (p, n, max_level) needs at most."""
    out = C.c_uint64()
    check(load().hpc_footprint(int(p), int(n), int(max_level), C.byref(out)))
    return int(out.value)


class Batch:
    """Independent skeleton instances on one GPU, one level of each per step."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.hpc_batch_create(int(device), C.byref(h)))
        self.h = h
        self.shape: dict[int, tuple[int, int]] = {}

    def close(self) -> None:
        if self.h:
            check(self.lib.hpc_batch_destroy(self.h))
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, x: np.ndarray, alpha: float, max_level: int = 4, corr_mode: int = 0) -> int:
        """New instance on observations x ([p][n] float64); returns its id."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        p, n = x.shape
        out = C.c_int()
        check(self.lib.hpc_batch_add(self.h, p, n, float(alpha), int(max_level), _ptr(x),
                                     C.byref(out)))
        self.shape[out.value] = (p, n)
        if corr_mode:
            check(self.lib.hpc_batch_set_corr_mode(self.h, out.value, int(corr_mode)))
        return out.value

    def remove(self, iid: int) -> None:
        check(self.lib.hpc_batch_remove(self.h, int(iid)))
        self.shape.pop(iid, None)

    def step(self) -> int:
        left = C.c_int()
        check(self.lib.hpc_batch_step(self.h, C.byref(left)))
        return left.value

    def run(self) -> int:
        steps = C.c_int()
        check(self.lib.hpc_batch_run(self.h, C.byref(steps)))
        return steps.value

    def status(self, iid: int) -> tuple[bool, int, int]:
        f, lv, err = C.c_int(), C.c_int(), C.c_int()
        check(self.lib.hpc_batch_status(self.h, int(iid), C.byref(f), C.byref(lv), C.byref(err)))
        return bool(f.value), lv.value, err.value

    def stats(self, iid: int) -> list[dict]:
        buf = (LevelStats * 64)()
        n = C.c_int()
        check(self.lib.hpc_batch_get_stats(self.h, int(iid), buf, 64, C.byref(n)))
        return [buf[k].as_dict() for k in range(min(n.value, 64))]

    def max_cond_size(self, iid: int) -> int:
        return int(self.lib.hpc_batch_max_cond_size(self.h, int(iid)))

    def cor(self, iid: int) -> np.ndarray:
        p = self.shape[iid][0]
        out = np.empty((p, p), dtype=np.float64)
        check(self.lib.hpc_batch_get_cor(self.h, int(iid), _ptr(out)))
        return out

    def adj(self, iid: int) -> np.ndarray:
        p = self.shape[iid][0]
        out = np.empty((p, p), dtype=np.int32)
        check(self.lib.hpc_batch_get_adj(self.h, int(iid), _ptr(out)))
        return out

    def sepsets(self, iid: int, mc: int | None = None) -> np.ndarray:
        p = self.shape[iid][0]
        mc = self.max_cond_size(iid) if mc is None else int(mc)
        out = np.empty((p, p, mc), dtype=np.int32)
        check(self.lib.hpc_batch_get_sepsets(self.h, int(iid), _ptr(out), mc))
        return out

    def sepset_list(self, iid: int) -> np.ndarray:
        return _fetch_list(lambda out, cap, w: self.lib.hpc_batch_get_sepset_list(
            self.h, int(iid), out, cap, w))

    def set_ranks_per_task(self, value: int) -> None:
        check(self.lib.hpc_batch_set_tuning(self.h, 1, int(value)))

    def set_lds_limit(self, nbytes: int) -> None:
        """LDS bytes per workgroup for staging a row in the CI-test kernels
        (0 = the 64 KiB default); rows that need more take the gather
        kernels."""
        check(self.lib.hpc_batch_set_tuning(self.h, 4, int(nbytes)))

    def autotune(self) -> list[dict]:
        """Per-level launch shapes timed on the batch's own steps; every
        instance is left at its start."""
        from ._lib import TuneEntry

        buf = (TuneEntry * 32)()
        n = C.c_int()
        check(self.lib.hpc_batch_autotune(self.h, buf, 32, C.byref(n)))
        return [{name: getattr(buf[k], name) for name, _ in TuneEntry._fields_}
                for k in range(min(n.value, 32))]

    def inject_failure(self, iid: int, level: int) -> None:
        """Test hook: the host side of instance iid's `level` fails."""
        check(self.lib.hpc_batch_inject_failure(self.h, int(iid), int(level)))

    def apply_tuning(self, entries: list[dict]) -> None:
        """Launch shapes (as DeviceSkeleton.autotune() returns them)."""
        arr, n = tune_array(entries)
        check(self.lib.hpc_batch_apply_tuning(self.h, arr, n))


class Server:
    """Continuous-batching skeleton server on one GPU."""

    def __init__(self, device: int = 0, hbm_budget: int = 0, max_active: int = 0,
                 corr_mode: int = 0):
        self.lib = load()
        h = C.c_void_p()
        check(self.lib.hpc_server_create(int(device), int(hbm_budget), int(max_active),
                                         C.byref(h)))
        self.h = h
        if corr_mode:
            check(self.lib.hpc_server_set_corr_mode(self.h, int(corr_mode)))
        self.shape: dict[int, int] = {}
        self._borrowed: dict[int, np.ndarray] = {}

    def close(self) -> None:
        if self.h:
            check(self.lib.hpc_server_destroy(self.h))
            self.h = C.c_void_p()
            self._borrowed.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, x: np.ndarray, alpha: float, max_level: int = 4, copy: bool = False) -> int:
        """Queues a skeleton request on observations x ([p][n] float64).
        Without copy the server reads x in place; the wrapper keeps a
        reference to it until the request's result is released."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        p, n = x.shape
        req = C.c_int64()
        fn = self.lib.hpc_server_submit if copy else self.lib.hpc_server_submit_borrowed
        check(fn(self.h, p, n, float(alpha), int(max_level), _ptr(x), C.byref(req)))
        self.shape[req.value] = p
        if not copy:
            self._borrowed[req.value] = x
        return req.value

    def wait(self, req: int, timeout_s: float = -1.0) -> bool:
        done = C.c_int()
        check(self.lib.hpc_server_wait(self.h, int(req), float(timeout_s), C.byref(done)))
        return bool(done.value)

    def sepset_list(self, req: int) -> np.ndarray:
        return _fetch_list(lambda out, cap, w: self.lib.hpc_server_sepset_list(
            self.h, int(req), out, cap, w))

    def result(self, req: int, mc: int | None = None, release: bool = True,
               dense_sepsets: bool = True):
        """(adj [p][p], sepsets, info dict) of a finished request: sepsets in
        the reference layout [p][p][mc], or the compact list when
        dense_sepsets is False."""
        p = self.shape[req]
        info = RequestInfo()
        check(self.lib.hpc_server_result(self.h, int(req), None, None, 0, C.byref(info)))
        mc = info.max_cond_size if mc is None else int(mc)
        adj = np.empty((p, p), dtype=np.int32)
        if dense_sepsets:
            sep = np.empty((p, p, mc), dtype=np.int32)
            check(self.lib.hpc_server_result(self.h, int(req), _ptr(adj), _ptr(sep), mc,
                                             C.byref(info)))
        else:
            check(self.lib.hpc_server_result(self.h, int(req), _ptr(adj), None, 0,
                                             C.byref(info)))
            sep = self.sepset_list(req)
        if release:
            check(self.lib.hpc_server_release(self.h, int(req)))
            self.shape.pop(req, None)
            self._borrowed.pop(req, None)
        return adj, sep, {name: getattr(info, name) for name, _ in info._fields_}

    def set_tuning(self, param: int, value: int) -> None:
        """A knob (HPC_TUNE_*) for the server's later steps."""
        check(self.lib.hpc_server_set_tuning(self.h, int(param), int(value)))

    def inject_failure(self, req: int, level: int) -> None:
        """Test hook: the host side of request req's `level` fails (req may
        be an id not yet submitted; ids count up from 1)."""
        check(self.lib.hpc_server_inject_failure(self.h, int(req), int(level)))

    def apply_tuning(self, entries: list[dict]) -> None:
        """Launch shapes for the server's later steps."""
        arr, n = tune_array(entries)
        check(self.lib.hpc_server_apply_tuning(self.h, arr, n))

    def stats(self) -> dict:
        info = ServerInfo()
        check(self.lib.hpc_server_stats(self.h, C.byref(info)))
        return {name: getattr(info, name) for name, _ in info._fields_}
