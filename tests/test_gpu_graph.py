"""The one-instance level loop captured in a hipGraph (HPC_TUNE_GRAPH).

A fresh capped run without kernel timing records its launches; the next one
captures them and later ones replay the graph, checking every level's sizes
against the recorded run.  Bar: every run's adjacency, sepsets and per-level
test counts equal the oracle's, replayed or not; a replay on other data (a
new correlation matrix of the same size) is detected and recomputed."""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


def _oracle_run(orc, cor, n, alpha, ml):
    ref = orc.Skeleton(cor, n, alpha, max_level=ml)
    recs = ref.run(ml)
    if ml < 0:
        # the oracle records the level that finds no testable row; hpc_run
        # stops at its compaction without a record
        assert recs[-1]["tests"] == 0
        recs = recs[:-1]
    return ref.adj(), ref.sepsets(), [r["tests"] for r in recs], ref.mc()


def _check(dev, want, label):
    adj, sep, tests, mc = want
    assert np.array_equal(dev.adj(), adj), f"{label}: adjacency differs"
    assert np.array_equal(dev.sepsets(mc), sep), f"{label}: sepsets differ"


@pytest.mark.parametrize("name", ["small", "c2k"])
def test_graph_replays_equal_oracle(hpc, orc, name):
    x, n, alpha, _ = workloads.sem(name)
    ml = 4
    p = x.shape[0]
    want = _oracle_run(orc, orc.corr(x), n, alpha, ml)
    with hpc.DeviceSkeleton(p, n, alpha, max_level=ml) as dev:
        dev.set_kernel_timing(False)
        dev.pmcc(x)
        for k in range(5):
            print(f"[graph] {name}: run {k}", flush=True)
            dev.reset()
            st = dev.run(ml)
            assert [s["tests"] for s in st] == want[2], (name, k)
            _check(dev, want, f"{name} run {k}")
        g = dev.graph_stats()
        assert g == {"replays": 4, "captures": 1, "mismatches": 0}, g
        # the correlation recomputed between runs (as the benchmark's steps do)
        dev.reset()
        dev.corr()
        dev.run(ml)
        _check(dev, want, f"{name} after corr")
        assert dev.graph_stats()["replays"] == 5


def test_graph_replay_on_other_data_recomputes(hpc, orc):
    x, n, alpha, _ = workloads.sem("small")
    p = x.shape[0]
    ml = 3
    cor1 = orc.corr(x)
    rng = np.random.default_rng(9)
    x2 = x.copy()
    x2[:, :] = x2[:, rng.permutation(n)]
    x2[::3] = rng.standard_normal((len(range(0, p, 3)), n))
    cor2 = orc.corr(x2)
    want1 = _oracle_run(orc, cor1, n, alpha, ml)
    want2 = _oracle_run(orc, cor2, n, alpha, ml)
    with hpc.DeviceSkeleton(p, n, alpha, max_level=ml) as dev:
        dev.set_kernel_timing(False)
        dev.set_corr(cor1)
        for _ in range(3):
            dev.reset()
            dev.run(ml)
            _check(dev, want1, "first data")
        assert dev.graph_stats()["replays"] == 2
        dev.set_corr(cor2)
        dev.reset()
        dev.run(ml)
        _check(dev, want2, "other data, recomputed")
        g = dev.graph_stats()
        assert g["mismatches"] == 1 and g["captures"] == 2, g
        for _ in range(2):
            dev.reset()
            dev.run(ml)
            _check(dev, want2, "other data, replayed")
        assert dev.graph_stats()["replays"] == 4


def test_no_graph_with_timing_or_knob_off(hpc, orc):
    x, n, alpha, _ = workloads.sem("small")
    p = x.shape[0]
    want = _oracle_run(orc, orc.corr(x), n, alpha, 4)
    for timing, knob in ((True, True), (False, False)):
        with hpc.DeviceSkeleton(p, n, alpha, max_level=4) as dev:
            dev.set_kernel_timing(timing)
            dev.set_graph(knob)
            dev.pmcc(x)
            for _ in range(3):
                dev.reset()
                st = dev.run(4)
                _check(dev, want, f"timing={timing} graph={knob}")
                if timing:
                    assert all(s["ms_test"] > 0 for s in st)
            assert dev.graph_stats() == {"replays": 0, "captures": 0, "mismatches": 0}


def test_graph_then_tuning_change(hpc, orc):
    """New launch shapes void the graph: the next fresh run records again."""
    x, n, alpha, _ = workloads.sem("small")
    p = x.shape[0]
    want = _oracle_run(orc, orc.corr(x), n, alpha, 4)
    with hpc.DeviceSkeleton(p, n, alpha, max_level=4) as dev:
        dev.set_kernel_timing(False)
        dev.pmcc(x)
        for _ in range(2):
            dev.reset()
            dev.run(4)
        dev.apply_tuning([{"level": 2, "ranks_per_task": 128, "tasks_per_block": 4, "queue": 1}])
        for _ in range(3):
            dev.reset()
            dev.run(4)
            _check(dev, want, "after new shapes")
        g = dev.graph_stats()
        assert g["captures"] == 2 and g["replays"] == 3, g


@pytest.mark.parametrize("name", ["small", "c2"])
def test_graph_replays_uncapped(hpc, orc, name):
    """With no level cap the recorded run ends on the compaction that finds
    no row with enough neighbours; replays check that compaction's sizes too
    and equal the oracle, and other data is detected and recomputed."""
    x, n, alpha, _ = workloads.sem(name)
    p = x.shape[0]
    want = _oracle_run(orc, orc.corr(x), n, alpha, -1)
    with hpc.DeviceSkeleton(p, n, alpha, max_level=-1) as dev:
        dev.set_kernel_timing(False)
        dev.pmcc(x)
        for k in range(4):
            print(f"[graph] {name} uncapped: run {k}", flush=True)
            dev.reset()
            st = dev.run(-1)
            assert [s["tests"] for s in st] == want[2], (name, k)
            _check(dev, want, f"{name} uncapped run {k}")
        g = dev.graph_stats()
        assert g == {"replays": 3, "captures": 1, "mismatches": 0}, g
        # other data of the same shape: the replay's sizes differ
        x2 = x[:, np.random.default_rng(5).permutation(n)][:, : n - 7].copy()
        x2 = np.concatenate([x2, x[:, :7] * 0.5], axis=1)
        want2 = _oracle_run(orc, orc.corr(x2), n, alpha, -1)
        dev.pmcc(x2)
        dev.reset()
        st = dev.run(-1)
        assert [s["tests"] for s in st] == want2[2]
        _check(dev, want2, f"{name} uncapped other data")
