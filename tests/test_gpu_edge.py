"""Degenerate inputs through the C ABI against the oracle: the smallest p,
sample counts at which the reference's Fisher z factor sqrt(n - 3)
(gpuRowTests.cu:18, cpuRowTests.cpp:12) is 0 or NaN, a single sample (every
correlation NaN), and alpha at 0 and 1.

Bar: as the other parity tests -- correlation bit-identical, adjacency,
compact lists, first passing ranks, test counts and sepsets identical per
level; for hpc_run the per-level counts; for the precision form the exact
path's skeleton.
"""
import numpy as np
import pytest

from test_gpu_parity import _compare_run
from test_gpu_precision import _fast_vs_oracle

pytestmark = pytest.mark.gpu

# name: (p, n, alpha, max_level) -- the oracle's outcome in the comment
CASES = {
    "p2": (2, 50, 0.05, -1),     # one pair, kept; level 1 has nothing to test
    "p3": (3, 50, 0.05, -1),     # one edge removed at level 1
    "n4": (8, 4, 0.05, 3),       # sqrt(1): most edges removed at level 0
    "n3": (5, 3, 0.05, 3),       # sqrt(0): p-value 1, every edge removed
    "n2": (6, 2, 0.05, 3),       # sqrt(-1): NaN p-values, every edge kept
    "n1": (8, 1, 0.05, 2),       # one sample: every correlation NaN
    "a0": (12, 200, 0.0, -1),    # alpha 0: every edge removed at level 0
    "a1": (12, 200, 1.0, 3),     # alpha 1: only p-value 1 removes
}


def _data(hpc, name):
    p, n, alpha, ml = CASES[name]
    if n == 1:
        x = np.random.default_rng(1).normal(size=(p, n))
    else:
        x = hpc.sem_generate(p, n, 2.0, 0.1, 1.0, seed=7, fp32=False)
    return x, n, alpha, ml


@pytest.mark.parametrize("name", list(CASES))
def test_degenerate_levels(hpc, orc, name):
    x, n, alpha, ml = _data(hpc, name)
    _compare_run(hpc, orc, x, n, alpha, ml, label=f"edge/{name}")


@pytest.mark.parametrize("name", list(CASES))
def test_degenerate_run(hpc, orc, name):
    """hpc_run over the same inputs: per-level counts, adjacency, sepsets."""
    x, n, alpha, ml = _data(hpc, name)
    p = x.shape[0]
    print(f"[edge run] {name}: p={p} n={n} alpha={alpha} max_level={ml}", flush=True)
    ref = orc.Skeleton(orc.corr(x), n, alpha, max_level=ml)
    recs = ref.run(ml)
    if ml < 0:
        # the oracle records the level that finds no testable row; hpc_run
        # stops at its compaction without a record
        assert recs[-1]["tests"] == 0
        recs = recs[:-1]
    with hpc.DeviceSkeleton(p, n, alpha, max_level=ml) as dev:
        dev.pmcc(x)
        st = dev.run(ml)
        assert [s["level"] for s in st] == list(range(len(recs)))
        assert [s["tests"] for s in st] == [r["tests"] for r in recs]
        assert [s["removed"] for s in st] == [r["removed"] for r in recs]
        assert np.array_equal(dev.adj(), ref.adj())
        assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets())


@pytest.mark.parametrize("name", [k for k in CASES if CASES[k][1] > 1])
def test_degenerate_precision_form(hpc, orc, name):
    """The FP32-MFMA correlation with exact rechecks on the same inputs (one
    sample excluded: its correlations are all NaN on both paths)."""
    x, n, alpha, ml = _data(hpc, name)
    _fast_vs_oracle(hpc, orc, x, n, alpha, ml, f"edge/{name}")


def _inputs(hpc):
    import workloads

    out = []
    for name in CASES:
        x, n, alpha, ml = _data(hpc, name)
        out.append((f"edge/{name}", x, alpha, ml))
    x, n, alpha, ml = workloads.sem("small")
    out.append(("small", x, alpha, ml))
    return out


def test_degenerate_batch(hpc, orc):
    """All of them, with an ordinary instance, in one batch: each instance
    gives the oracle's result."""
    from test_gpu_batch import _check_instance

    inputs = _inputs(hpc)
    with hpc.Batch() as b:
        ids = [b.add(x, a, ml) for _, x, a, ml in inputs]
        b.run()
        for iid, (label, x, a, ml) in zip(ids, inputs):
            _check_instance(b, iid, label, orc, x, a, ml)


def test_degenerate_served(hpc, orc):
    """The same requests served alone and in a mix: the oracle's result."""
    from test_gpu_batch import _check_list, _oracle, _serve

    inputs = _inputs(hpc)
    alone, _ = _serve(hpc, inputs, one_at_a_time=True)
    mixed, _ = _serve(hpc, inputs)
    for label, x, a, ml in inputs:
        cor, ref, recs = _oracle(orc, x, a, ml)
        mc = ref.mc()
        so = ref.sepsets()
        for mode, res in (("alone", alone), ("mixed", mixed)):
            adj, sep, info, lst = res[label]
            assert np.array_equal(adj, ref.adj()), f"{label} ({mode}): adjacency differs"
            assert np.array_equal(sep[:, :, :mc], so), f"{label} ({mode}): sepsets"
            _check_list(lst, so, f"{label} ({mode})")
            assert info["tests"] == sum(o["tests"] for o in recs), (label, mode)
