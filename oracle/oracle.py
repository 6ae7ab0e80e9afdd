"""ctypes binding of liboracle.so (oracle.h).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "liboracle.so")

MODE_CANON = 0
MODE_PINV = 1
FLAG_REF_EARLY_EXIT = 1

_vp, _i, _i64, _u64, _d = C.c_void_p, C.c_int, C.c_int64, C.c_uint64, C.c_double
_SIG = {
    "orc_pval": (_d, [_d, _i]),
    "orc_removes": (_i, [_d, _d]),
    "orc_pcor_l1": (_d, [_d, _d, _d]),
    "orc_pval_l1": (_d, [_d, _d, _d, _i]),
    "orc_pcor_canon": (_d, [_i, _vp, _vp]),
    "orc_pcor_pinv": (_d, [_i, _vp]),
    "orc_pval_ln": (_d, [_i, _vp, _i, _i]),
    "orc_binom": (_u64, [_i, _i]),
    "orc_unrank": (None, [_u64, _i, _i, _vp]),
    "orc_moments": (None, [_vp, _i, _i, _vp, _vp]),
    "orc_corr": (None, [_vp, _i, _i, _vp, _i]),
    "orc_new": (_vp, [_vp, _i, _i, _d, _i, _i, _i, _i]),
    "orc_free": (None, [_vp]),
    "orc_level": (_i, [_vp, _i, _vp]),
    "orc_run": (_i, [_vp, _i, _vp, _i]),
    "orc_get_adj": (None, [_vp, _vp]),
    "orc_get_compact": (None, [_vp, _vp]),
    "orc_get_sepsets": (None, [_vp, _vp]),
    "orc_num_ordered_edges": (_i64, [_vp]),
    "orc_get_first_pass": (None, [_vp, _vp, _vp]),
    "orc_mc": (_i, [_vp]),
    "orc_pval_ids": (_d, [_vp, _i, _i, _i, _i, _vp, _i, _i]),
    "orc_r_crit": (_d, [_i, _d]),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        for k, (res, args) in _SIG.items():
            f = getattr(L, k)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def pval(r: float, n: int) -> float:
    return lib().orc_pval(float(r), int(n))


def pval_l1(x1, x2, x3, n) -> float:
    return lib().orc_pval_l1(float(x1), float(x2), float(x3), int(n))


def pcor_l1(x1, x2, x3) -> float:
    return lib().orc_pcor_l1(float(x1), float(x2), float(x3))


def pcor_canon(M: np.ndarray):
    M = np.ascontiguousarray(M, dtype=np.float64)
    used = C.c_int()
    r = lib().orc_pcor_canon(M.shape[0], _p(M), C.addressof(used))
    return r, bool(used.value)


def pcor_pinv(M: np.ndarray) -> float:
    M = np.ascontiguousarray(M, dtype=np.float64)
    return lib().orc_pcor_pinv(M.shape[0], _p(M))


def pval_ln(M: np.ndarray, n: int, mode: int = MODE_CANON) -> float:
    M = np.ascontiguousarray(M, dtype=np.float64)
    return lib().orc_pval_ln(M.shape[0], _p(M), int(n), int(mode))


def binom(n: int, k: int) -> int:
    return int(lib().orc_binom(int(n), int(k)))


def unrank(rank: int, l: int, nn: int) -> list[int]:
    q = np.zeros(max(l, 1), dtype=np.int32)
    lib().orc_unrank(C.c_uint64(rank), int(l), int(nn), _p(q))
    return q[:l].tolist()


def moments(x: np.ndarray):
    x = np.ascontiguousarray(x, dtype=np.float64)
    p, n = x.shape
    mean = np.empty(p)
    sd = np.empty(p)
    lib().orc_moments(_p(x), p, n, _p(mean), _p(sd))
    return mean, sd


def corr(x: np.ndarray, threads: int = 0) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    p, n = x.shape
    cor = np.empty((p, p))
    lib().orc_corr(_p(x), p, n, _p(cor), int(threads))
    return cor


def r_crit(n: int, alpha: float) -> float:
    return lib().orc_r_crit(int(n), float(alpha))


def pval_ids(cor: np.ndarray, n: int, i: int, j: int, S, mode: int = MODE_CANON) -> float:
    cor = np.ascontiguousarray(cor, dtype=np.float64)
    s = np.ascontiguousarray(list(S) if len(S) else [0], dtype=np.int32)
    return lib().orc_pval_ids(_p(cor), cor.shape[0], int(n), int(i), int(j), _p(s), len(S), int(mode))


class Skeleton:
    """Oracle PC-stable run over a correlation matrix (reference CPU path)."""

    def __init__(self, cor: np.ndarray, n: int, alpha: float, max_level: int = 4,
                 mode: int = MODE_CANON, flags: int = 0, threads: int = 0):
        self._cor = np.ascontiguousarray(cor, dtype=np.float64)
        self.p = self._cor.shape[0]
        self.n = int(n)
        self.alpha = float(alpha)
        mc = max(max_level, 1) if max_level >= 0 else 4
        self.h = lib().orc_new(_p(self._cor), self.p, self.n, self.alpha, mc,
                               int(mode), int(flags), int(threads))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_free(self.h)
            self.h = None

    def level(self, level: int) -> dict:
        s = np.zeros(5, dtype=np.int64)
        none = lib().orc_level(self.h, int(level), _p(s))
        return {"tests": int(s[0]), "tests_max": int(s[1]), "edges": int(s[2]),
                "removed": int(s[3]), "max_deg": int(s[4]), "nothing": bool(none)}

    def run(self, max_level: int = 4) -> list[dict]:
        cap = 256
        s = np.zeros((cap, 5), dtype=np.int64)
        k = lib().orc_run(self.h, int(max_level), _p(s), cap)
        return [{"tests": int(a), "tests_max": int(b), "edges": int(c), "removed": int(d),
                 "max_deg": int(e)} for a, b, c, d, e in s[: min(k, cap)]]

    def adj(self) -> np.ndarray:
        out = np.empty((self.p, self.p), dtype=np.int32)
        lib().orc_get_adj(self.h, _p(out))
        return out

    def adj_compact(self) -> np.ndarray:
        out = np.empty((self.p, self.p), dtype=np.int32)
        lib().orc_get_compact(self.h, _p(out))
        return out

    def mc(self) -> int:
        return lib().orc_mc(self.h)

    def sepsets(self) -> np.ndarray:
        out = np.empty((self.p, self.p, self.mc()), dtype=np.int32)
        lib().orc_get_sepsets(self.h, _p(out))
        return out

    def first_pass(self):
        ne = lib().orc_num_ordered_edges(self.h)
        fp = np.empty(max(ne, 1), dtype=np.int64)
        rp = np.empty(self.p + 1, dtype=np.int32)
        lib().orc_get_first_pass(self.h, _p(fp), _p(rp))
        return fp[:ne], rp
