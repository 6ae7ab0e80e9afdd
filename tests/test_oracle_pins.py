"""Pins of the CPU oracle against facts independent of it and of the reference:
scipy / numpy golden vectors (tests/golden/make_golden.py), closed forms and
identities.  No GPU needed."""
import itertools
import os

import numpy as np
import pytest

GOLD = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vectors.npz"))


def test_pval_matches_scipy(orc):
    for r, n, p in zip(GOLD["pv_r"], GOLD["pv_n"], GOLD["pv_p"]):
        got = orc.pval(r, int(n))
        assert got == pytest.approx(p, rel=1e-11, abs=1e-300), (r, n, got, p)


def test_pval_special_cases(orc):
    assert orc.pval(float("nan"), 100) == 1.0          # NaN r counts as 0
    assert orc.pval(0.0, 100) == 1.0
    assert orc.pval(1.0, 100) == orc.pval(0.9999999, 100)  # clamp
    assert orc.pval(-0.3, 100) == orc.pval(0.3, 100)
    assert np.isnan(orc.pval_l1(float("nan"), 0.1, 0.2, 100))  # CPU NaN policy
    assert orc.lib().orc_removes(float("nan"), 0.05) == 0
    assert orc.lib().orc_removes(0.05, 0.05) == 1       # p >= alpha removes


def test_level1_closed_form_matches_pinv(orc):
    for (x1, x2, x3), want in zip(GOLD["p3_in"], GOLD["p3_out"]):
        assert orc.pcor_l1(x1, x2, x3) == pytest.approx(want, abs=1e-12)


def test_level_n_canonical_and_pinv_match_numpy(orc):
    nfb = 0
    for m, buf, want in zip(GOLD["pn_m"], GOLD["pn_M"], GOLD["pn_out"]):
        M = buf[:m, :m]
        r_c, used = orc.pcor_canon(M)
        r_p = orc.pcor_pinv(M)
        nfb += used
        if np.isnan(want):
            assert np.isnan(r_p)
            continue
        assert r_p == pytest.approx(want, abs=1e-8), (m, r_p, want)
        assert r_c == pytest.approx(want, abs=1e-8), (m, r_c, want, used)
    # the rank-deficient cases at the end take the pinv form
    assert nfb >= 10


def test_recursive_partial_correlation_identity(orc):
    """rho_ij.Sk = (rho_ij.S - rho_ik.S rho_jk.S) / sqrt((1-rho_ik.S^2)(1-rho_jk.S^2))."""
    rng = np.random.default_rng(5)
    for m in range(4, 10):
        for _ in range(20):
            A = rng.normal(size=(60, m))
            A[:, 1:] += rng.normal(size=(60, 1))
            C = np.corrcoef(A.T)
            # variables: 0 = i, 1 = j, last = k, 2..m-2 = S
            k = m - 1
            S = list(range(2, m - 1))

            def pc(a, b, cond):
                idx = [a, b] + cond
                return orc.pcor_canon(C[np.ix_(idx, idx)])[0]

            lhs = pc(0, 1, S + [k])
            rij, rik, rjk = pc(0, 1, S), pc(0, k, S), pc(1, k, S)
            rhs = (rij - rik * rjk) / np.sqrt((1 - rik ** 2) * (1 - rjk ** 2))
            assert lhs == pytest.approx(rhs, abs=1e-11)


def test_unrank_is_lexicographic(orc):
    for n, k, rank, q in zip(GOLD["cb_n"], GOLD["cb_k"], GOLD["cb_rank"], GOLD["cb_q"]):
        assert orc.unrank(int(rank), int(k), int(n)) == list(q[:k])


def test_unrank_brute_force_to_20(orc):
    for n in (13, 17, 20):
        for k in (1, 2, 3, 6):
            for rank, comb in enumerate(itertools.combinations(range(n), k)):
                if rank % 97 == 0 or rank < 50:
                    assert orc.unrank(rank, k, n) == list(comb)


def test_binom(orc):
    from math import comb
    for n in range(0, 70):
        for k in range(0, 12):
            assert orc.binom(n, k) == (comb(n, k) if k <= n else 0)
    assert orc.binom(200, 100) == 2 ** 64 - 1  # saturates


@pytest.mark.parametrize("n,alpha", [(1000, 0.01), (5000, 0.05), (2000, 0.01), (10000, 0.05),
                                      (200, 0.05), (300, 0.01), (20000, 0.05), (4, 0.05)])
def test_r_crit_is_the_exact_decision_boundary(orc, n, alpha):
    rc = orc.r_crit(n, alpha)
    assert orc.pval(rc, n) >= alpha
    assert orc.pval(np.nextafter(rc, 2.0), n) < alpha
    # monotone around the boundary: p >= alpha  <=>  r <= r_crit
    r = rc
    for _ in range(300):
        r = np.nextafter(r, 2.0)
        assert orc.pval(r, n) < alpha
    r = rc
    for _ in range(300):
        assert orc.pval(r, n) >= alpha
        r = np.nextafter(r, -1.0)
    rng = np.random.default_rng(n)
    for r in rng.uniform(0, 1, 2000):
        assert (orc.pval(r, n) >= alpha) == (min(r, 0.9999999) <= rc)


def test_r_crit_limits(orc):
    assert orc.r_crit(3, 0.05) == float("inf")   # sqrt(0) * z: every p = 1
    assert orc.r_crit(2, 0.05) == -1.0           # NaN p-values never remove


def test_moments_and_correlation_match_numpy(orc):
    rng = np.random.default_rng(9)
    x = rng.normal(size=(23, 517)) * rng.uniform(0.1, 30, size=(23, 1)) + rng.normal(size=(23, 1))
    x[3] += 2.0 * x[4]
    mean, sd = orc.moments(x)
    assert np.allclose(mean, x.mean(axis=1), rtol=1e-14, atol=1e-13)
    var = x.var(axis=1, ddof=1)
    sd_quirk = np.sqrt(var.astype(np.float32)).astype(np.float64)  # sqrtf of a float
    assert np.allclose(sd, sd_quirk, rtol=2e-7)
    assert np.array_equal(sd.astype(np.float32).astype(np.float64), sd)  # exactly a float
    cor = orc.corr(x)
    exact = np.corrcoef(x)
    # the reference divides by the float-rounded sd: r = r_exact * sd sd / (sdf sdf)
    s_exact = np.sqrt(var)
    scale = np.outer(s_exact, s_exact) / np.outer(sd, sd)
    assert np.allclose(cor, exact * scale, rtol=1e-12, atol=1e-14)
    assert np.allclose(cor, exact, rtol=3e-7, atol=1e-12)
    assert np.array_equal(cor, cor.T)


def test_correlation_nan_policy(orc):
    rng = np.random.default_rng(4)
    x = rng.normal(size=(6, 300))
    x[1, 5] = np.nan
    x[2, :] = 3.0   # zero variance
    cor = orc.corr(x)
    mean, sd = orc.moments(x)
    assert mean[1] == pytest.approx(np.nanmean(x[1]), rel=1e-13)
    ok = ~np.isnan(x[0]) & ~np.isnan(x[1])
    z0 = (x[0] - mean[0]) / sd[0]
    z1 = (x[1] - mean[1]) / sd[1]
    assert cor[0, 1] == pytest.approx(np.sum(z0[ok] * z1[ok]) / (300 - 1), rel=1e-12)
    assert np.isnan(cor[0, 2])  # sd = 0: (x - mean) / 0 -> NaN products
