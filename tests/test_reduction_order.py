"""The correlation kernels' recursive-halving reduction of many accumulators
adds the same pairs of partial sums as the reference's halving tree
(corHelper.cu:101-116: lane t += lane t + s for s = 32, ..., 1), so it gives
the same bits.

A numpy emulation of the device exchange (each stage: a lane keeps half of
its values and adds the partner lane t ^ s's partial sums of them) against
the tree, on values spread over ten decades so that any change in the order
of additions changes the rounding.
"""
import numpy as np


def _tree64(vals):
    v = vals.copy()
    for s in (32, 16, 8, 4, 2, 1):
        w = v.copy()
        for t in range(64 - s):
            w[t] = v[t] + v[t + s]
        v = w
    return v[0]


def _halve(V, H, S):
    """One stage: 2H values per lane in, H out; partner lane t ^ S."""
    out = np.empty((64, H))
    for t in range(64):
        p = t ^ S
        up, pup = (t & S) != 0, (p & S) != 0
        for a in range(H):
            keep = V[t][a + H] if up else V[t][a]
            recv = V[p][a] if pup else V[p][a + H]  # what the partner sends
            out[t][a] = keep + recv
    return out


def _values(rng, m):
    return rng.standard_normal((64, m)) * 10.0 ** rng.integers(-5, 5, size=(64, m))


def test_halving_64_accumulators_equals_tree():
    rng = np.random.default_rng(0)
    for _ in range(4):
        A = _values(rng, 64)
        V = A.copy()
        for H in (32, 16, 8, 4, 2, 1):
            V = _halve(V, H, H)
        for t in range(64):
            assert V[t][0].tobytes() == _tree64(A[:, t]).tobytes(), t


def test_halving_16_accumulators_equals_tree():
    rng = np.random.default_rng(1)
    for _ in range(4):
        A = _values(rng, 16)
        V = A.copy()
        for H, S in ((8, 32), (4, 16), (2, 8), (1, 4)):
            V = _halve(V, H, S)
        v = V[:, 0].copy()
        for s in (2, 1):
            v = np.array([v[t] + v[t ^ s] for t in range(64)])
        for t in range(64):
            assert v[t].tobytes() == _tree64(A[:, (t >> 2) & 15]).tobytes(), t


def test_a_changed_order_is_detected():
    """The values do expose the order: pairing (t, t + 1) first gives other bits
    for most accumulators."""
    rng = np.random.default_rng(2)
    A = _values(rng, 64)
    differ = 0
    for a in range(64):
        v = A[:, a].copy()
        for s in (1, 2, 4, 8, 16, 32):
            v = np.array([v[t] + v[t ^ s] for t in range(64)])
        differ += v[0].tobytes() != _tree64(A[:, a]).tobytes()
    assert differ > 32
