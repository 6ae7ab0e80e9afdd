"""numpy-facing wrapper of one hpc_state (one skeleton instance on one GPU)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import HPC_CORR_CANON, LevelStats, check, load


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _fetch_list(call) -> np.ndarray:
    words = C.c_int64()
    check(call(None, 0, C.byref(words)))
    out = np.empty(max(words.value, 1), dtype=np.int32)
    check(call(_ptr(out), words.value, C.byref(words)))
    return out[: words.value]


def sepset_records(lst: np.ndarray) -> dict:
    """{(i, j): (level, members)} of a compact sepset list."""
    out, a = {}, 0
    while a < len(lst):
        i, j, lv = int(lst[a]), int(lst[a + 1]), int(lst[a + 2])
        out[(i, j)] = (lv, [int(v) for v in lst[a + 3: a + 3 + lv]])
        a += 3 + lv
    return out


def r_crit(n: int, alpha: float) -> float:
    """Largest |r| whose Fisher-z p-value (cpuInd.cpp:8-14) is >= alpha."""
    return load().hpc_fisher_r_crit(int(n), float(alpha))


def sem_generate(p: int, n: int, degree: float, wmin: float = 0.1,
                 wmax: float = 1.0, seed: int = 0, fp32: bool = True,
                 with_dag: bool = False):
    """Linear-Gaussian SEM sample, variable-major [p][n] float64."""
    x = np.empty((p, n), dtype=np.float64)
    dag = np.zeros((p, p), dtype=np.uint8) if with_dag else None
    check(load().hpc_sem_generate(p, n, float(degree), float(wmin), float(wmax),
                                  C.c_uint64(seed), 1 if fp32 else 0, _ptr(x),
                                  _ptr(dag) if dag is not None else None))
    return (x, dag) if with_dag else x


class DeviceSkeleton:
    """One skeleton instance resident on a GPU (MMState + executors)."""

    def __init__(self, p: int, n: int, alpha: float, max_level: int = 4,
                 device: int = 0):
        self.lib = load()
        self.p, self.n, self.alpha, self.max_level = int(p), int(n), float(alpha), int(max_level)
        h = C.c_void_p()
        check(self.lib.hpc_state_create(self.p, self.n, self.alpha, self.max_level,
                                        int(device), C.byref(h)))
        self.h = h

    def close(self) -> None:
        if self.h:
            check(self.lib.hpc_state_destroy(self.h))
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

    # ---- inputs ----------------------------------------------------------
    def pmcc(self, x: np.ndarray, mode: int = HPC_CORR_CANON) -> None:
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.shape == (self.p, self.n)
        check(self.lib.hpc_pmcc(self.h, _ptr(x), int(mode)))

    def set_observations(self, x: np.ndarray) -> None:
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert x.shape == (self.p, self.n)
        check(self.lib.hpc_set_observations(self.h, _ptr(x)))

    def corr(self, mode: int = HPC_CORR_CANON) -> None:
        check(self.lib.hpc_corr(self.h, int(mode)))

    def corr_ms(self) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        check(self.lib.hpc_corr_ms(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_corr(self, cor: np.ndarray) -> None:
        cor = np.ascontiguousarray(cor, dtype=np.float64)
        assert cor.shape == (self.p, self.p)
        check(self.lib.hpc_set_corr(self.h, _ptr(cor)))

    def set_ranks_per_task(self, value: int) -> None:
        check(self.lib.hpc_set_tuning(self.h, 1, int(value)))

    def autotune(self, max_level: int = 4) -> list[dict]:
        """Per-level launch shapes of the CI-test kernels, timed on this
        state's own levels; the state is left reset."""
        from ._lib import TuneEntry

        buf = (TuneEntry * 32)()
        n = C.c_int()
        check(self.lib.hpc_autotune(self.h, int(max_level), buf, 32, C.byref(n)))
        return [{name: getattr(buf[k], name) for name, _ in TuneEntry._fields_}
                for k in range(min(n.value, 32))]

    def apply_tuning(self, entries: list[dict]) -> None:
        """Launch shapes (as autotune() returns them) for later runs."""
        from ._lib import tune_array

        arr, n = tune_array(entries)
        check(self.lib.hpc_apply_tuning(self.h, arr, n))

    def set_recheck_queue(self, value: int) -> None:
        """Length of the HPC_CORR_F32 recheck queues (0 = full)."""
        check(self.lib.hpc_set_tuning(self.h, 2, int(value)))

    def set_lds_limit(self, nbytes: int) -> None:
        """LDS bytes per workgroup for staging a row in the CI-test kernels
        (0 = the 64 KiB default); rows that need more take the gather
        kernels."""
        check(self.lib.hpc_set_tuning(self.h, 4, int(nbytes)))

    def set_kernel_timing(self, on: bool) -> None:
        """Device-time events around each level's kernels in run() (the
        ms_compact / ms_test / ms_commit fields); off, the level loop runs
        without them and those fields are 0."""
        check(self.lib.hpc_set_tuning(self.h, 3, 1 if on else 0))

    # ---- levels ----------------------------------------------------------
    def reset(self) -> None:
        check(self.lib.hpc_reset(self.h))

    def sync(self) -> None:
        check(self.lib.hpc_sync(self.h))

    def compact(self) -> int:
        md = C.c_int()
        check(self.lib.hpc_compact(self.h, C.byref(md)))
        return md.value

    def level(self, level: int) -> dict:
        s = LevelStats()
        check(self.lib.hpc_level(self.h, int(level), C.byref(s)))
        return s.as_dict()

    def test_level(self, level: int, rows=None) -> dict:
        s = LevelStats()
        if rows is None:
            check(self.lib.hpc_test_level(self.h, int(level), None, 0, C.byref(s)))
        else:
            r = np.ascontiguousarray(rows, dtype=np.int32)
            check(self.lib.hpc_test_level(self.h, int(level), _ptr(r), len(r), C.byref(s)))
        return s.as_dict()

    def search_level(self, level: int, rows=None, host: bool = False, threads: int = 0) -> dict:
        """Search rows (all when None) on the device, or on the host when
        host=True; hpc_commit_level then removes the edges."""
        s = LevelStats()
        r = None if rows is None else np.ascontiguousarray(rows, dtype=np.int32)
        ptr, cnt = (None, 0) if r is None else (_ptr(r), len(r))
        if host:
            if r is None:
                r = np.arange(self.p, dtype=np.int32)
                ptr, cnt = _ptr(r), len(r)
            check(self.lib.hpc_search_level_host(self.h, int(level), ptr, cnt, int(threads),
                                                 C.byref(s)))
        else:
            check(self.lib.hpc_search_level(self.h, int(level), ptr, cnt, C.byref(s)))
        return s.as_dict()

    def search_level_shared(self, level: int, threads: int = 0) -> tuple[dict, dict]:
        """One queue of the level's tasks, the device taking blocks from the
        front and `threads` host threads from the back; returns (device,
        host) records (host `removed` = blocks the host finished).
        hpc_commit_level then removes the edges."""
        d, h = LevelStats(), LevelStats()
        check(self.lib.hpc_search_level_shared(self.h, int(level), int(threads), C.byref(d),
                                               C.byref(h)))
        return d.as_dict(), h.as_dict()

    def set_graph(self, on: bool) -> None:
        """hipGraph replay of the level loop for fresh, capped runs without
        kernel timing (default on)."""
        check(self.lib.hpc_set_tuning(self.h, 7, 1 if on else 0))

    def graph_stats(self) -> dict:
        r, c, m = C.c_int64(), C.c_int64(), C.c_int64()
        check(self.lib.hpc_graph_stats(self.h, C.byref(r), C.byref(c), C.byref(m)))
        return {"replays": r.value, "captures": c.value, "mismatches": m.value}

    def set_queue(self, on: bool) -> None:
        """Device task queue of the level >= 2 CI-test kernel (default on)."""
        check(self.lib.hpc_set_tuning(self.h, 6, 1 if on else 0))

    def commit_level(self, level: int) -> dict:
        s = LevelStats()
        check(self.lib.hpc_commit_level(self.h, int(level), C.byref(s)))
        return s.as_dict()

    def run(self, max_level: int = -2, cap: int = 64) -> list[dict]:
        arr = (LevelStats * cap)()
        ran = C.c_int()
        check(self.lib.hpc_run(self.h, int(max_level), arr, cap, C.byref(ran)))
        return [arr[k].as_dict() for k in range(min(ran.value, cap))]

    # ---- outputs ---------------------------------------------------------
    def cor(self) -> np.ndarray:
        out = np.empty((self.p, self.p), dtype=np.float64)
        check(self.lib.hpc_get_cor(self.h, _ptr(out)))
        return out

    def adj(self) -> np.ndarray:
        out = np.empty((self.p, self.p), dtype=np.int32)
        check(self.lib.hpc_get_adj(self.h, _ptr(out)))
        return out

    def adj_compact(self) -> np.ndarray:
        out = np.empty((self.p, self.p), dtype=np.int32)
        check(self.lib.hpc_get_compact(self.h, _ptr(out)))
        return out

    def max_cond_size(self) -> int:
        return self.lib.hpc_max_cond_size(self.h)

    def sepsets(self, mc: int | None = None) -> np.ndarray:
        mc = self.max_cond_size() if mc is None else int(mc)
        out = np.empty((self.p, self.p, mc), dtype=np.int32)
        check(self.lib.hpc_get_sepsets(self.h, _ptr(out), mc))
        return out

    def sepset_list(self) -> np.ndarray:
        """Compact sepsets: per edge removed at level l >= 1, the words
        i, j, l, members..., sorted by (i, j)."""
        return _fetch_list(lambda out, cap, w: self.lib.hpc_get_sepset_list(self.h, out, cap, w))

    def first_pass(self):
        ne = C.c_int64()
        check(self.lib.hpc_get_first_pass(self.h, None, None, C.byref(ne)))
        fp = np.empty(max(ne.value, 1), dtype=np.int64)
        rp = np.empty(self.p + 1, dtype=np.int32)
        check(self.lib.hpc_get_first_pass(self.h, _ptr(fp), _ptr(rp), C.byref(ne)))
        return fp[: ne.value], rp
