"""HIP path vs the CPU oracle through the C ABI, on the same seeded inputs.

Bar: correlation bit-identical (same summation order and operations on both
sides), adjacency and neighbour lists bit-identical after every level, first
passing rank of every ordered edge identical, canonical test counts equal,
sepsets identical and each one accepted by the reference's test.
"""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


def _compare_run(hpc, orc, x, n, alpha, max_level, cor=None, label=""):
    p = x.shape[0] if x is not None else cor.shape[0]
    print(f"[parity] {label}: p={p} n={n} alpha={alpha} max_level={max_level}", flush=True)
    dev = hpc.DeviceSkeleton(p, n, alpha, max_level=max_level)
    if cor is None:
        dev.pmcc(x)
        ref_cor = orc.corr(x)
        got = dev.cor()
        assert np.array_equal(got, ref_cor, equal_nan=True), \
            f"correlation differs: max |d| = {np.nanmax(np.abs(got - ref_cor))}"
    else:
        dev.set_corr(cor)
        ref_cor = cor
    ref = orc.Skeleton(ref_cor, n, alpha, max_level=max_level)
    level = 0
    while True:
        if max_level >= 0 and level > max_level:
            break
        d = dev.level(level)
        o = ref.level(level)
        if level >= 1:
            assert np.array_equal(dev.adj_compact(), ref.adj_compact()), f"compact differs at level {level}"
            if max_level < 0 and o["max_deg"] - 1 < level:
                break
            fp_d, rp_d = dev.first_pass()
            fp_o, rp_o = ref.first_pass()
            assert np.array_equal(rp_d, rp_o)
            assert np.array_equal(fp_d, fp_o), f"first passing ranks differ at level {level}"
        assert np.array_equal(dev.adj(), ref.adj()), f"adjacency differs after level {level}"
        assert d["tests"] == o["tests"], (level, d, o)
        assert d["tests_max"] == o["tests_max"], (level, d, o)
        assert d["removed"] == o["removed"], (level, d, o)
        level += 1
        if level > 12:
            break
    mc = ref.mc()
    sd = dev.sepsets(mc)
    so = ref.sepsets()
    assert np.array_equal(sd, so), "sepsets differ"
    # the compact list carries the same records
    from test_gpu_batch import _check_list

    _check_list(dev.sepset_list(), so, label)
    # every recorded sepset is one the reference's test accepts
    adj = ref.adj()
    checked = 0
    for i in range(p):
        for j in range(i + 1, p):
            if adj[i, j]:
                continue
            s = so[i, j]
            if s[0] == -2:
                continue
            S = [int(v) for v in s if v >= 0]
            pv = orc.pval_ids(ref_cor, n, i, j, S)
            pv2 = orc.pval_ids(ref_cor, n, j, i, S)
            assert pv >= alpha or pv2 >= alpha, (i, j, S, pv, pv2)
            checked += 1
            if checked > 3000:
                break
    dev.close()
    return level


@pytest.mark.parametrize("name", ["c1", "small", "dense", "deep"])
def test_sem_parity(hpc, orc, name):
    x, n, alpha, ml = workloads.sem(name)
    _compare_run(hpc, orc, x, n, alpha, ml, label=name)


def test_sem_parity_level_cap(hpc, orc):
    x, n, alpha, _ = workloads.sem("small")
    _compare_run(hpc, orc, x, n, alpha, 2, label="small/max_level=2")


def test_cooling_house_parity(hpc, orc):
    x = workloads.cooling_house()
    _compare_run(hpc, orc, x, x.shape[1], 0.05, 4, label="cooling_house")


def test_nan_and_constant_columns(hpc, orc):
    x, n, alpha, _ = workloads.sem("small")
    x = x[:80].copy()
    rng = np.random.default_rng(3)
    mask = rng.random(x.shape) < 0.01
    x[mask] = np.nan
    x[5] = 1.25          # zero variance
    x[9, :] = np.nan     # all NaN
    _compare_run(hpc, orc, x, n, alpha, 3, label="nan")


def test_duplicate_columns_pinv(hpc, orc):
    x, n, alpha, _ = workloads.sem("small")
    x = x[:60].copy()
    x[7] = x[3]
    x[11] = 2.0 * x[3] - 1.0
    x[20] = x[21] + x[22]
    _compare_run(hpc, orc, x, n, alpha, -1, label="collinear")


def test_given_correlation_matrix(hpc, orc):
    x, n, alpha, _ = workloads.sem("small")
    cor = np.corrcoef(x[:150])
    _compare_run(hpc, orc, None, n, alpha, -1, cor=cor, label="--corr")


def test_c2_full_parity(hpc, orc):
    x, n, alpha, ml = workloads.sem("c2")
    _compare_run(hpc, orc, x, n, alpha, ml, label="c2")


def test_host_share_matches_device(hpc, orc):
    """Rows split between the device and host threads (the heterogeneous
    executor) give the device-only first passes, adjacency and sepsets."""
    x, n, alpha, _ = workloads.sem("small")
    p = x.shape[0]
    full = hpc.DeviceSkeleton(p, n, alpha, max_level=5)
    mixed = hpc.DeviceSkeleton(p, n, alpha, max_level=5)
    full.pmcc(x)
    mixed.pmcc(x)
    full.level(0)
    mixed.level(0)
    for level in range(1, 6):
        full.level(level)
        mixed.compact()
        rows = np.arange(p)
        mixed.search_level(level, rows[rows % 3 != 0])
        mixed.search_level(level, rows[rows % 3 == 0], host=True, threads=4)
        mixed.commit_level(level)
        fa, _ = full.first_pass()
        fb, _ = mixed.first_pass()
        assert np.array_equal(fa, fb), level
        assert np.array_equal(full.adj(), mixed.adj()), level
    assert np.array_equal(full.sepsets(5), mixed.sepsets(5))


def test_row_subset_then_rest(hpc, orc):
    """hpc_test_level on two disjoint row sets gives the same first passes."""
    x, n, alpha, _ = workloads.sem("small")
    dev = hpc.DeviceSkeleton(x.shape[0], n, alpha, max_level=-1)
    dev.pmcc(x)
    dev.level(0)
    dev.compact()
    full = hpc.DeviceSkeleton(x.shape[0], n, alpha, max_level=-1)
    full.pmcc(x)
    full.level(0)
    full.level(1)
    rows = np.arange(0, x.shape[0], 2)
    dev.test_level(1, rows)
    fp_half, rp = dev.first_pass()
    fp_full, _ = full.first_pass()
    for i in range(x.shape[0]):
        seg = slice(rp[i], rp[i + 1])
        if i % 2 == 0:
            assert np.array_equal(fp_half[seg], fp_full[seg])
        else:
            assert np.all(fp_half[seg] == -1)


def test_host_share_alone_matches_device(hpc, orc):
    """Every row searched on the host (the CPU-only executor) at levels 1..5:
    the commit removes the same edges with the same sepsets as the device."""
    x, n, alpha, _ = workloads.sem("small")
    x = x[:150].copy()
    p = x.shape[0]
    full = hpc.DeviceSkeleton(p, n, alpha, max_level=5)
    host = hpc.DeviceSkeleton(p, n, alpha, max_level=5)
    full.pmcc(x)
    host.pmcc(x)
    full.level(0)
    host.level(0)
    rows = np.arange(p)
    for level in range(1, 6):
        print(f"[host-only] level {level}", flush=True)
        full.level(level)
        host.compact()
        host.search_level(level, rows, host=True, threads=4)
        host.commit_level(level)
        fa, _ = full.first_pass()
        fb, _ = host.first_pass()
        assert np.array_equal(fa, fb), level
        assert np.array_equal(full.adj(), host.adj()), level
    assert np.array_equal(full.sepsets(5), host.sepsets(5))


def test_autotune_picks_shapes_and_keeps_results(hpc, orc):
    """The autotuner times every launch shape per level on the state's own
    snapshots, leaves the state reset, and the tuned run equals the oracle."""
    x, n, alpha, _ = workloads.sem("c2")
    p = x.shape[0]
    with hpc.DeviceSkeleton(p, n, alpha, max_level=4) as dev:
        dev.pmcc(x)
        entries = dev.autotune(4)
        print(f"[autotune] {entries}", flush=True)
        assert [t["level"] for t in entries] == list(range(1, len(entries) + 1))
        assert entries[0]["edges_per_block"] in (8, 16, 32, 64, 128, 256)
        for t in entries[1:]:
            assert t["ranks_per_task"] in (256, 512, 1024, 2048, 4096)
            assert t["tasks_per_block"] in (4, 8, 16, 32, 64, 128)
            assert 0 < t["ms_best"] <= t["ms_default"]
        assert dev.adj().sum() == p * (p - 1)  # reset
        st = dev.run(4)
        ref = orc.Skeleton(orc.corr(x), n, alpha, max_level=4)
        recs = ref.run(4)
        assert np.array_equal(dev.adj(), ref.adj())
        assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets())
        assert [s["tests"] for s in st] == [r["tests"] for r in recs]


def test_applied_tuning_any_shape_keeps_results(hpc, orc):
    """Launch shapes handed over with hpc_apply_tuning -- the extremes of the
    accepted ranges included -- leave adjacency, sepsets and test counts
    unchanged; an invalid entry is rejected before anything is applied."""
    x, n, alpha, _ = workloads.sem("dense")
    p = x.shape[0]
    ref = orc.Skeleton(orc.corr(x), n, alpha, max_level=4)
    recs = ref.run(4)
    shapes = [
        [{"level": 1, "edges_per_block": 1}] +
        [{"level": l, "ranks_per_task": 64, "tasks_per_block": 1} for l in (2, 3, 4)],
        [{"level": 1, "edges_per_block": 4096}] +
        [{"level": l, "ranks_per_task": 1 << 20, "tasks_per_block": 1024} for l in (2, 3, 4)],
        [{"level": 3, "ranks_per_task": 192, "tasks_per_block": 3}],
    ]
    with hpc.DeviceSkeleton(p, n, alpha, max_level=4) as dev:
        dev.pmcc(x)
        for bad in ({"level": 0, "edges_per_block": 8},
                    {"level": 2, "ranks_per_task": 100, "tasks_per_block": 4},
                    {"level": 2, "ranks_per_task": 128, "tasks_per_block": 0},
                    {"level": 1, "edges_per_block": 5000}):
            with pytest.raises(hpc.HpcError):
                dev.apply_tuning([{"level": 1, "edges_per_block": 16}, bad])
        for k, entries in enumerate(shapes):
            print(f"[apply_tuning] shape set {k}: {entries}", flush=True)
            dev.apply_tuning(entries)
            dev.reset()
            st = dev.run(4)
            assert np.array_equal(dev.adj(), ref.adj()), k
            assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets()), k
            assert [s["tests"] for s in st] == [r["tests"] for r in recs], k


def test_run_caps_and_mixed_entry_points(hpc, orc):
    """hpc_run's one-wait level loop stops exactly at level caps 0, 1, 2 and
    4, and a run after levels taken one by one (hpc_level) gives the same
    skeleton as a fresh run."""
    x, n, alpha, _ = workloads.sem("small")
    p = x.shape[0]
    cor = orc.corr(x)
    with hpc.DeviceSkeleton(p, n, alpha, max_level=4) as dev:
        dev.pmcc(x)
        for cap in (0, 1, 2, 4):
            print(f"[run] cap {cap}", flush=True)
            ref = orc.Skeleton(cor, n, alpha, max_level=cap)
            recs = ref.run(cap)
            dev.reset()
            st = dev.run(cap)
            assert [s["level"] for s in st] == list(range(cap + 1)), (cap, st)
            assert np.array_equal(dev.adj(), ref.adj()), cap
            assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets()), cap
            assert [s["tests"] for s in st] == [r["tests"] for r in recs], cap
            assert [s["removed"] for s in st] == [r["removed"] for r in recs], cap
        # levels 0 and 1 one by one, then a full run from scratch
        dev.reset()
        dev.level(0)
        dev.level(1)
        dev.reset()
        st = dev.run(4)
        ref = orc.Skeleton(cor, n, alpha, max_level=4)
        recs = ref.run(4)
        assert np.array_equal(dev.adj(), ref.adj())
        assert [s["tests"] for s in st] == [r["tests"] for r in recs]


@pytest.mark.parametrize("name,cap", [("deep", -1), ("dense", -1), ("deep", 6)])
def test_run_unbounded_and_deep_cap(hpc, orc, name, cap):
    """hpc_run to the end (no level cap: the loop stops when no row has
    enough neighbours) and with a cap of 6 (the next level's search queued
    before each wait): the oracle's adjacency, sepsets and per-level counts."""
    x, n, alpha, _ = workloads.sem(name)
    p = x.shape[0]
    cor = orc.corr(x)
    ref = orc.Skeleton(cor, n, alpha, max_level=cap)
    recs = ref.run(cap)
    if cap < 0:
        # the oracle records the level that finds no testable row; hpc_run
        # stops at its compaction without a record
        assert recs[-1]["tests"] == 0 and recs[-1]["max_deg"] - 1 < len(recs) - 1
        recs = recs[:-1]
    print(f"[run] {name} cap={cap}: {len(recs)} levels", flush=True)
    with hpc.DeviceSkeleton(p, n, alpha, max_level=cap) as dev:
        dev.pmcc(x)
        st = dev.run(cap)
        assert [s["level"] for s in st] == list(range(len(recs)))
        assert [s["tests"] for s in st] == [r["tests"] for r in recs]
        assert [s["removed"] for s in st] == [r["removed"] for r in recs]
        assert np.array_equal(dev.adj(), ref.adj())
        assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets())


def test_kernel_timing_off_keeps_results(hpc, orc):
    """hpc_run without the per-level kernel-time events (HPC_TUNE_KERNEL_TIMING
    = 0): the oracle's adjacency, sepsets and per-level counts, and device
    times of 0; switched back on, the times return."""
    x, n, alpha, _ = workloads.sem("dense")
    p = x.shape[0]
    cor = orc.corr(x)
    ref = orc.Skeleton(cor, n, alpha, max_level=4)
    recs = ref.run(4)
    with hpc.DeviceSkeleton(p, n, alpha, max_level=4) as dev:
        dev.pmcc(x)
        for on in (False, True):
            print(f"[timing] on={on}", flush=True)
            dev.set_kernel_timing(on)
            dev.reset()
            st = dev.run(4)
            assert np.array_equal(dev.adj(), ref.adj()), on
            assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets()), on
            assert [s["tests"] for s in st] == [r["tests"] for r in recs], on
            assert [s["removed"] for s in st] == [r["removed"] for r in recs], on
            ms = sum(s["ms_test"] for s in st)
            assert (ms > 0) if on else (ms == 0), (on, ms)


@pytest.mark.parametrize("mode", ["exact", "f32"])
def test_benchmark_workload_parity(hpc, orc, mode):
    """The p=2000 benchmark workload (n=5000, degree 3, alpha 0.05, levels
    0..4) through hpc_run: the oracle's correlation (exact mode), adjacency,
    sepsets and per-level test counts, in the exact and the FP32-MFMA
    correlation modes."""
    x, n, alpha, ml = workloads.sem("c2k")
    p = x.shape[0]
    cor = orc.corr(x)
    ref = orc.Skeleton(cor, n, alpha, max_level=ml)
    recs = ref.run(ml)
    with hpc.DeviceSkeleton(p, n, alpha, max_level=ml) as dev:
        print(f"[c2k] mode {mode}", flush=True)
        dev.pmcc(x, mode=0 if mode == "exact" else 1)  # HPC_CORR_CANON / HPC_CORR_F32
        if mode == "exact":
            assert np.array_equal(dev.cor(), cor)
        st = dev.run(ml)
        assert np.array_equal(dev.adj(), ref.adj())
        assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets())
        assert [s["tests"] for s in st] == [r["tests"] for r in recs]
        assert [s["removed"] for s in st] == [r["removed"] for r in recs]
