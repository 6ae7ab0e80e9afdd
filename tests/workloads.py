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


def cooling_house() -> np.ndarray:
    """data/cooling_house.csv as variable-major [6][10000]."""
    X = np.loadtxt(os.path.join(GOLDEN, "cooling_house.csv"), delimiter=",")
    return np.ascontiguousarray(X.T)
