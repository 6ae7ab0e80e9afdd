"""Writes tests/golden/vectors.npz: golden input/output vectors for the oracle.

Expected values come from scipy / numpy / itertools (independent of the
oracle and of the reference):
  pval      2 * norm.sf(z) with z = sqrt(n-3) * atanh(min(|r|, 0.9999999))
  pcor3     partial correlation of 0,1 given 2 from numpy.linalg.pinv of the
            3x3 submatrix
  pcorN     -P01 / sqrt(P00 P11) with P = numpy.linalg.pinv(M, rcond) and
            rcond = m * eps (Armadillo's default tolerance is
            max(m,n) * max(sv) * eps), for random SPD correlation matrices
            of size 4..12 and rank-deficient ones
  combos    itertools.combinations order for n <= 12, k <= 5

Run:  python tests/golden/make_golden.py
"""
import itertools
import os

import numpy as np
from scipy import stats

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "vectors.npz")
rng = np.random.default_rng(20241203)


def pval(r, n):
    r = 0.0 if np.isnan(r) else min(0.9999999, abs(r))
    z = np.sqrt(n - 3.0) * np.arctanh(r)
    return 2.0 * stats.norm.sf(z)


def pcor_pinv(M):
    m = M.shape[0]
    P = np.linalg.pinv(M, rcond=m * np.finfo(float).eps, hermitian=True)
    return -P[0, 1] / np.sqrt(P[0, 0] * P[1, 1])


def random_corr(m, cond_scale=1.0):
    A = rng.normal(size=(m + 3 + int(10 * rng.random()), m))
    A[:, 1:] += cond_scale * rng.normal(size=(A.shape[0], 1)) * rng.random()
    C = np.corrcoef(A.T)
    return C


def main():
    # p-values over a grid of r and n
    rs = np.concatenate([np.linspace(-1, 1, 41), rng.uniform(-0.2, 0.2, 60),
                         [0.0, 1e-9, 0.9999999, 0.99999999, -0.99999995, 0.5]])
    ns = np.array([4, 10, 50, 300, 1000, 2000, 5000, 10000], dtype=np.int64)
    pv_r, pv_n, pv_p = [], [], []
    for n in ns:
        for r in rs:
            pv_r.append(r)
            pv_n.append(n)
            pv_p.append(pval(r, n))
    # level-1 closed form vs pinv of the 3x3 submatrix
    p3_in, p3_out = [], []
    for _ in range(300):
        C = random_corr(3)
        p3_in.append([C[0, 1], C[0, 2], C[1, 2]])
        p3_out.append(pcor_pinv(C))
    # level >= 2: random SPD correlation submatrices
    pn_m, pn_M, pn_out = [], [], []
    for m in range(4, 13):
        for _ in range(40):
            C = random_corr(m, cond_scale=3.0 * rng.random())
            pn_m.append(m)
            buf = np.zeros((14, 14))
            buf[:m, :m] = C
            pn_M.append(buf)
            pn_out.append(pcor_pinv(C))
    # rank-deficient: a conditioning variable duplicated / a linear combination
    for m in range(4, 9):
        for kind in range(3):
            A = rng.normal(size=(40, m))
            if kind == 0:
                A[:, m - 1] = A[:, 2]
            elif kind == 1:
                A[:, m - 1] = A[:, 2] - 0.5 * A[:, 3]
            else:
                A[:, 0] = A[:, 2] + A[:, 3]
            C = np.corrcoef(A.T)
            pn_m.append(m)
            buf = np.zeros((14, 14))
            buf[:m, :m] = C
            pn_M.append(buf)
            pn_out.append(pcor_pinv(C))
    # combinations in lexicographic order
    cb_n, cb_k, cb_rank, cb_q = [], [], [], []
    for n in range(1, 13):
        for k in range(1, min(n, 5) + 1):
            for rank, comb in enumerate(itertools.combinations(range(n), k)):
                cb_n.append(n)
                cb_k.append(k)
                cb_rank.append(rank)
                q = np.full(5, -1, dtype=np.int32)
                q[:k] = comb
                cb_q.append(q)
    np.savez_compressed(
        OUT,
        pv_r=np.array(pv_r), pv_n=np.array(pv_n), pv_p=np.array(pv_p),
        p3_in=np.array(p3_in), p3_out=np.array(p3_out),
        pn_m=np.array(pn_m), pn_M=np.array(pn_M), pn_out=np.array(pn_out),
        cb_n=np.array(cb_n), cb_k=np.array(cb_k), cb_rank=np.array(cb_rank),
        cb_q=np.array(cb_q),
    )
    print(OUT)


if __name__ == "__main__":
    main()
