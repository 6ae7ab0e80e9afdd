"""Precision form (HPC_CORR_F32): MFMA FP32 correlation with exact rechecks.

Bar: the fast correlation stays within its stated bound of the exact FP64
one (74 * 2^-24 * 1.01 absolute), and the skeleton computed from it -- the
adjacency after every level, every first passing rank, the canonical test
counts and the sepsets -- equals the exact path's (the oracle's) bit for bit,
including on data with many tests near the threshold and when the recheck
queue overflows and the level is redone on exact correlations.
"""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu

F32 = 1
EPS = 74.0 * 2.0 ** -24 * 1.01


def _near_threshold(hpc, p=400, n=2000, seed=104):
    """A small precision-form input: weights +-[0.05, 1] put many tests near
    the alpha = 0.01 threshold."""
    return hpc.sem_generate(p, n, 4.0, 0.05, 1.0, seed=seed, fp32=True), n, 0.01, -1


def _fast_vs_oracle(hpc, orc, x, n, alpha, ml, label, queue=0):
    p = x.shape[0]
    print(f"[f32] {label}: p={p} n={n} alpha={alpha} max_level={ml} queue={queue}", flush=True)
    dev = hpc.DeviceSkeleton(p, n, alpha, max_level=ml)
    if queue:
        dev.set_recheck_queue(queue)
    dev.pmcc(x, F32)
    fast = dev.cor()
    exact = orc.corr(x)
    off = ~np.eye(p, dtype=bool)
    err = np.abs(fast - exact)[off]
    assert np.nanmax(err) <= EPS, f"{label}: max |fast - exact| = {np.nanmax(err)}"
    ref = orc.Skeleton(exact, n, alpha, max_level=ml)
    rechecks = 0
    level = 0
    while True:
        if ml >= 0 and level > ml:
            break
        d = dev.level(level)
        o = ref.level(level)
        rechecks += d.get("rechecks", 0)
        if level >= 1:
            assert np.array_equal(dev.adj_compact(), ref.adj_compact()), (label, level)
            if ml < 0 and o["max_deg"] - 1 < level:
                break
            fp_d, _ = dev.first_pass()
            fp_o, _ = ref.first_pass()
            assert np.array_equal(fp_d, fp_o), f"{label}: first passing ranks differ at level {level}"
        assert np.array_equal(dev.adj(), ref.adj()), f"{label}: adjacency differs after level {level}"
        assert d["tests"] == o["tests"], (label, level, d, o)
        level += 1
        if level > 10:
            break
    mc = ref.mc()
    assert np.array_equal(dev.sepsets(mc), ref.sepsets()), f"{label}: sepsets differ"
    dev.close()
    return float(np.nanmax(err)), rechecks


@pytest.mark.parametrize("name", ["small", "dense", "deep", "c2"])
def test_f32_skeleton_equals_exact(hpc, orc, name):
    x, n, alpha, ml = workloads.sem(name)
    err, rc = _fast_vs_oracle(hpc, orc, x, n, alpha, ml, name)
    print(f"[f32] {name}: max |dcor| = {err:.3e}, rechecks = {rc}", flush=True)


def test_f32_near_threshold_rechecks(hpc, orc):
    x, n, alpha, ml = _near_threshold(hpc)
    err, rc = _fast_vs_oracle(hpc, orc, x, n, alpha, ml, "near-threshold")
    print(f"[f32] near-threshold: max |dcor| = {err:.3e}, rechecks = {rc}", flush=True)
    assert rc > 0  # the recheck path ran


def test_f32_queue_overflow_redoes_exactly(hpc, orc):
    x, n, alpha, ml = _near_threshold(hpc, p=200, seed=7)
    _fast_vs_oracle(hpc, orc, x, n, alpha, ml, "overflow", queue=1)


def test_f32_run_matches_exact_run(hpc):
    x, n, alpha, _ = workloads.sem("c2")
    p = x.shape[0]
    out = []
    for mode in (0, F32):
        with hpc.DeviceSkeleton(p, n, alpha, max_level=4) as dev:
            dev.pmcc(x, mode)
            st = dev.run(4)
            out.append((dev.adj(), dev.sepset_list(), [s["tests"] for s in st]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_f32_batch_and_server_mixed_modes(hpc, orc):
    items = []
    for name in ("small", "dense", "deep"):
        x, n, alpha, ml = workloads.sem(name)
        items.append((x, alpha, ml))
    x, n, alpha, ml = _near_threshold(hpc, p=250, seed=9)
    items.append((x, alpha, ml))
    xn = items[0][0][:70].copy()
    xn[np.random.default_rng(2).random(xn.shape) < 0.01] = np.nan  # NaN: exact kernels
    items.append((xn, 0.05, 3))
    want = []
    for x, a, ml in items:
        ref = orc.Skeleton(orc.corr(x), x.shape[1], a, max_level=ml)
        ref.run(ml)
        want.append((ref.adj(), ref.sepsets()))
    with hpc.Batch() as b:
        ids = [b.add(x, a, ml, corr_mode=F32 if k % 2 == 0 else 0)
               for k, (x, a, ml) in enumerate(items)]
        b.run()
        for iid, (adj, sep) in zip(ids, want):
            assert np.array_equal(b.adj(iid), adj)
            assert np.array_equal(b.sepsets(iid, sep.shape[2]), sep)
    with hpc.Server(corr_mode=F32) as s:
        reqs = [s.submit(x, a, ml) for x, a, ml in items]
        for r, (adj, sep) in zip(reqs, want):
            assert s.wait(r, 120.0)
            got_adj, got_sep, _ = s.result(r, mc=sep.shape[2])
            assert np.array_equal(got_adj, adj)
            assert np.array_equal(got_sep, sep)


@pytest.mark.parametrize("queue", [0, 1])
def test_f32_run_rechecks_and_overflow(hpc, orc, queue):
    """hpc_run in the FP32 mode on data with many tests near the threshold:
    rechecks run inside the one-wait level loop, and with a one-entry queue
    the voided level is redone on exact correlations; the skeleton and the
    per-level test counts are the oracle's either way."""
    x, n, alpha, _ = _near_threshold(hpc, p=200, seed=7)
    p = x.shape[0]
    exact = orc.corr(x)
    ref = orc.Skeleton(exact, n, alpha, max_level=4)
    recs = ref.run(4)
    with hpc.DeviceSkeleton(p, n, alpha, max_level=4) as dev:
        if queue:
            dev.set_recheck_queue(queue)
        print(f"[f32 run] queue={queue}", flush=True)
        dev.pmcc(x, F32)
        st = dev.run(4)
        assert np.array_equal(dev.adj(), ref.adj())
        assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets())
        assert [s["tests"] for s in st] == [r["tests"] for r in recs]
        assert [s["level"] for s in st] == list(range(len(recs)))
        if not queue:
            assert sum(s["rechecks"] for s in st) > 0
