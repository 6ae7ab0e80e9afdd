"""Seeded inputs shared by the parity tests, smoke() and bench.py.

Each workload is a linear-Gaussian SEM drawn by hpc_sem_generate (host code
of the library; the reference ships no generator) or a reference data file.
"""
from __future__ import annotations

import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# name: (p, n, degree, wmin, wmax, fp32, alpha, max_level, seed)
SEM = {
    # reference CPU-runnable case: p=50, n=1000, edge prob 2/(p-1), FP64,
    # alpha=0.01, no level cap
    "c1": (50, 1000, 2.0, 0.1, 1.0, False, 0.01, -1, 101),
    # one-GPU plain case: p=1000, n=5000, degree 3, FP32 data, alpha 0.05
    "c2": (1000, 5000, 3.0, 0.1, 1.0, True, 0.05, -1, 102),
    # the headline plain run at p=2000 (c2 generator)
    "c2k": (2000, 5000, 3.0, 0.1, 1.0, True, 0.05, 4, 103),
    # precision form: p=4000, n=2000, degree 4, weights +-[0.05,1]
    "c3": (4000, 2000, 4.0, 0.05, 1.0, True, 0.01, -1, 104),
    # small cases for parity at every level
    "small": (300, 2000, 3.0, 0.1, 1.0, True, 0.05, -1, 7),
    "dense": (120, 300, 8.0, 0.1, 1.0, True, 0.05, -1, 11),
    "deep": (60, 20000, 6.0, 0.3, 1.0, False, 0.05, -1, 5),
}


def sem(name: str):
    """(x [p][n] float64, n, alpha, max_level) of a named SEM workload."""
    from heterogpc import sem_generate

    p, n, d, wmin, wmax, fp32, alpha, ml, seed = SEM[name]
    x = sem_generate(p, n, d, wmin, wmax, seed=seed, fp32=fp32)
    return x, n, alpha, ml


def latent(p: int = 150, block: int = 110, n: int = 1000, lo: float = 0.2, hi: float = 0.6,
           seed: int = 5) -> np.ndarray:
    """Dense neighbourhoods: the first `block` of p variables load on one
    latent common factor (loadings +-U[lo, hi]), which is not observed, so no
    conditioning set separates them and their level-2 rows hold about
    `block` neighbours (past what a workgroup stages in LDS); the other
    variables are independent N(0,1).  FP32-rounded, variable-major."""
    rng = np.random.default_rng(seed)
    f = rng.standard_normal(n)
    x = rng.standard_normal((p, n))
    lam = rng.uniform(lo, hi, block) * rng.choice([-1.0, 1.0], block)
    x[:block] += lam[:, None] * f[None, :]
    return np.ascontiguousarray(x.astype(np.float32).astype(np.float64))


def cooling_house() -> np.ndarray:
    """data/cooling_house.csv as variable-major [6][10000]."""
    X = np.loadtxt(os.path.join(GOLDEN, "cooling_house.csv"), delimiter=",")
    return np.ascontiguousarray(X.T)


def batch_mix(count: int = 256, seed: int = 2024):
    """Batched form: `count` independent SEM instances, p uniform in [20,400],
    n uniform in [200,5000], expected degree uniform in [2,6], alpha in
    {0.01,0.05}, FP32 data, levels until no edge has enough neighbours.
    Returns [(x [p][n], alpha, max_level)]."""
    from heterogpc import sem_generate

    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        p = int(rng.integers(20, 401))
        n = int(rng.integers(200, 5001))
        d = float(rng.uniform(2.0, 6.0))
        a = float(rng.choice([0.01, 0.05]))
        x = sem_generate(p, n, d, 0.1, 1.0, seed=seed * 1000 + k, fp32=True)
        out.append((x, a, -1))
    return out


def served_trace(count: int = 200, rate: float = 100.0, seed: int = 77, max_level: int = 4):
    """Served form: Poisson arrivals at `rate` requests/s of SEM skeleton
    requests, p log-uniform in [50,2000], n uniform in [500,10000], expected
    degree uniform in [2,5], alpha in {0.01,0.05}, FP32 data.
    Returns [(t_arrival_s, x [p][n], alpha, max_level)]."""
    from heterogpc import sem_generate

    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for k in range(count):
        t += float(rng.exponential(1.0 / rate))
        p = int(round(float(np.exp(rng.uniform(np.log(50.0), np.log(2000.0))))))
        n = int(rng.integers(500, 10001))
        d = float(rng.uniform(2.0, 5.0))
        a = float(rng.choice([0.01, 0.05]))
        x = sem_generate(p, n, d, 0.1, 1.0, seed=seed * 1000 + k, fp32=True)
        out.append((t, x, a, max_level))
    return out
