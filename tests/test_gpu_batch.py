"""Batched and served skeletons through the C ABI (hpc_batch_*, hpc_server_*).

Bar: every instance of a mixed batch -- different p, n, alpha, level caps,
NaN cells, collinear columns -- gives the oracle's correlation, adjacency,
sepsets and per-level test counts bit for bit, and the same as a single
hpc_state run on its input; a served request gives the same result alone and
in a mix, whatever the admission budget.
"""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


def _mixed_inputs(hpc):
    """(label, x, alpha, max_level) of a heterogeneous set of instances."""
    out = []
    for name in ("c1", "small", "dense", "deep"):
        x, n, alpha, ml = workloads.sem(name)
        out.append((name, x, alpha, ml))
    x, n, alpha, _ = workloads.sem("small")
    out.append(("small/cap2", x[:200].copy(), alpha, 2))
    out.append(("small/cap0", x[:90].copy(), 0.01, 0))
    xn = x[:80].copy()
    rng = np.random.default_rng(3)
    xn[rng.random(xn.shape) < 0.01] = np.nan
    xn[5] = 1.25
    xn[9, :] = np.nan
    out.append(("nan", xn, alpha, 3))
    xc = x[:60].copy()
    xc[7] = xc[3]
    xc[11] = 2.0 * xc[3] - 1.0
    xc[20] = xc[21] + xc[22]
    out.append(("collinear", xc, alpha, -1))
    out.append(("cooling_house", workloads.cooling_house(), 0.05, 4))
    rng = np.random.default_rng(11)
    for k in range(5):
        p = int(rng.integers(20, 240))
        n = int(rng.integers(200, 3000))
        d = float(rng.uniform(2, 6))
        a = float(rng.choice([0.01, 0.05]))
        xs = hpc.sem_generate(p, n, d, 0.1, 1.0, seed=1000 + k, fp32=True)
        out.append((f"rand{k}", xs, a, int(rng.choice([-1, 3, 4]))))
    return out


def _oracle(orc, x, alpha, ml):
    n = x.shape[1]
    cor = orc.corr(x)
    ref = orc.Skeleton(cor, n, alpha, max_level=ml)
    recs = ref.run(ml)
    return cor, ref, recs


def _check_instance(b, iid, label, orc, x, alpha, ml, single=None, fast=False):
    """fast: the instance's correlation is the FP32-MFMA one (within its
    bound of the oracle's); everything else must still be identical."""
    print(f"[batch] {label}: p={x.shape[0]} n={x.shape[1]} alpha={alpha} max_level={ml}",
          flush=True)
    cor, ref, recs = _oracle(orc, x, alpha, ml)
    if fast:
        off = ~np.eye(x.shape[0], dtype=bool)
        assert np.nanmax(np.abs(b.cor(iid) - cor)[off]) <= 74.0 * 2.0 ** -24 * 1.01, label
    else:
        assert np.array_equal(b.cor(iid), cor, equal_nan=True), f"{label}: correlation differs"
    assert np.array_equal(b.adj(iid), ref.adj()), f"{label}: adjacency differs"
    mc = ref.mc()
    so = ref.sepsets()
    assert np.array_equal(b.sepsets(iid, mc), so), f"{label}: sepsets differ"
    _check_list(b.sepset_list(iid), so, label)
    finished, levels, err = b.status(iid)
    assert finished and err == 0, (label, finished, err)
    st = b.stats(iid)
    assert len(st) == levels
    # the oracle's unlimited run also records the final level with no tests
    assert len(recs) in (levels, levels + 1), (label, len(recs), levels)
    for lv, (d, o) in enumerate(zip(st, recs)):
        assert d["level"] == lv
        for key in ("tests", "tests_max", "removed"):
            assert d[key] == o[key], (label, lv, key, d, o)
    if single is not None:
        s_adj, s_sep, s_stats = single
        assert np.array_equal(b.adj(iid), s_adj), f"{label}: differs from the single run"
        assert np.array_equal(b.sepsets(iid, s_sep.shape[2]), s_sep)
        assert [d["tests"] for d in st] == [d["tests"] for d in s_stats]


def _check_list(lst, dense, label):
    """The compact sepset list holds exactly the dense layout's records at
    levels >= 1, sorted by (i, j)."""
    import heterogpc

    rec = heterogpc.sepset_records(lst)
    keys = list(rec)
    assert keys == sorted(keys), f"{label}: sepset list not sorted"
    want = {}
    p = dense.shape[0]
    for i in range(p):
        for j in range(i + 1, p):
            s = dense[i, j]
            if s[0] >= 0:
                mem = [int(v) for v in s if v >= 0]
                want[(i, j)] = (len(mem), mem)
    assert rec == want, f"{label}: sepset list differs from the dense sepsets"


def _single(hpc, x, alpha, ml):
    p, n = x.shape
    with hpc.DeviceSkeleton(p, n, alpha, max_level=ml) as dev:
        dev.pmcc(x)
        stats = dev.run(ml)
        return dev.adj(), dev.sepsets(), stats


def test_batch_mixed_matches_oracle_and_single(hpc, orc):
    inputs = _mixed_inputs(hpc)
    singles = [_single(hpc, x, a, ml) for _, x, a, ml in inputs]
    with hpc.Batch() as b:
        ids = [b.add(x, a, ml) for _, x, a, ml in inputs]
        steps = b.run()
        print(f"[batch] {len(ids)} instances in {steps} steps", flush=True)
        for iid, (label, x, a, ml), single in zip(ids, inputs, singles):
            _check_instance(b, iid, label, orc, x, a, ml, single)


def test_batch_continuous_admission_and_slot_reuse(hpc, orc):
    inputs = _mixed_inputs(hpc)
    first, later = inputs[:6], inputs[6:]
    with hpc.Batch() as b:
        # launch shapes other than the defaults; results must not change
        b.apply_tuning([{"level": 1, "edges_per_block": 8},
                        {"level": 2, "ranks_per_task": 128, "tasks_per_block": 4},
                        {"level": 3, "ranks_per_task": 2048, "tasks_per_block": 32}])
        with pytest.raises(hpc.HpcError):
            b.apply_tuning([{"level": 17, "ranks_per_task": 128, "tasks_per_block": 4}])
        ids = {}
        for label, x, a, ml in first:
            ids[label] = b.add(x, a, ml)
        b.step()
        b.step()
        # drop one mid-run; its slot is reused by a later instance
        gone = ids.pop("dense")
        b.remove(gone)
        for label, x, a, ml in later:
            ids[label] = b.add(x, a, ml)
            b.step()
        b.run()
        for label, x, a, ml in inputs:
            if label in ids:
                _check_instance(b, ids[label], label, orc, x, a, ml)


def test_batch_many_small_instances(hpc, orc):
    rng = np.random.default_rng(5)
    inputs = []
    for k in range(64):
        p = int(rng.integers(10, 80))
        n = int(rng.integers(100, 1500))
        x = hpc.sem_generate(p, n, float(rng.uniform(1.5, 5)), 0.1, 1.0, seed=50 + k, fp32=True)
        inputs.append((f"tiny{k}", x, float(rng.choice([0.01, 0.05])), -1))
    with hpc.Batch() as b:
        ids = [b.add(x, a, ml) for _, x, a, ml in inputs]
        b.run()
        for iid, (label, x, a, ml) in zip(ids, inputs):
            _check_instance(b, iid, label, orc, x, a, ml)


def _serve(hpc, inputs, budget=0, max_active=0, one_at_a_time=False):
    out = {}
    with hpc.Server(hbm_budget=budget, max_active=max_active) as s:
        reqs = []
        for label, x, a, ml in inputs:
            # both the copying and the in-place submit
            r = s.submit(x, a, ml, copy=len(reqs) % 2 == 1)
            if one_at_a_time:
                assert s.wait(r, 120.0)
            reqs.append((label, r))
        for label, r in reqs:
            assert s.wait(r, 120.0), f"{label} did not finish"
            lst = s.sepset_list(r)
            out[label] = s.result(r, mc=16) + (lst,)
        stats = s.stats()
    return out, stats


def test_server_result_same_alone_and_in_mix(hpc, orc):
    inputs = _mixed_inputs(hpc)
    alone, _ = _serve(hpc, inputs, one_at_a_time=True)
    mixed, st = _serve(hpc, inputs)
    print(f"[server] mixed: {st}", flush=True)
    biggest = max(hpc.footprint(x.shape[0], x.shape[1], ml) for _, x, _, ml in inputs)
    tight, st2 = _serve(hpc, inputs, budget=int(biggest * 1.5), max_active=4)
    print(f"[server] tight budget: {st2}", flush=True)
    assert st2["peak_active"] <= 4 and st2["peak_bytes"] <= st2["budget"]
    for label, x, a, ml in inputs:
        cor, ref, recs = _oracle(orc, x, a, ml)
        mc = ref.mc()
        so = ref.sepsets()
        for mode, res in (("alone", alone), ("mixed", mixed), ("tight", tight)):
            adj, sep, info, lst = res[label]
            assert np.array_equal(adj, ref.adj()), f"{label} ({mode}): adjacency differs"
            assert np.array_equal(sep[:, :, :mc], so), f"{label} ({mode}): sepsets"
            assert (sep[:, :, mc:] == -1).all()
            _check_list(lst, so, f"{label} ({mode})")
            assert info["tests"] == sum(o["tests"] for o in recs), (label, mode)
            assert info["latency_ms"] >= info["queue_ms"] >= 0.0


def test_server_rejects_request_beyond_budget(hpc):
    x, n, alpha, _ = workloads.sem("small")
    with hpc.Server(hbm_budget=1 << 20) as s:
        with pytest.raises(hpc.HpcError):
            s.submit(x, alpha, 4)


def test_server_applied_tuning_keeps_results(hpc, orc):
    """Launch shapes handed to a running server apply from its next step on
    and leave every result unchanged."""
    inputs = _mixed_inputs(hpc)[:6]
    with hpc.Server() as s:
        with pytest.raises(hpc.HpcError):
            s.apply_tuning([{"level": 2, "ranks_per_task": 96, "tasks_per_block": 4}])
        reqs = []
        for k, (label, x, a, ml) in enumerate(inputs):
            if k == 2:
                s.apply_tuning([{"level": 1, "edges_per_block": 64},
                                {"level": 2, "ranks_per_task": 256, "tasks_per_block": 8},
                                {"level": 4, "ranks_per_task": 1024, "tasks_per_block": 2}])
            reqs.append((label, x, a, ml, s.submit(x, a, ml, copy=True)))
        for label, x, a, ml, r in reqs:
            assert s.wait(r, 120.0), f"{label} did not finish"
            adj, sep, info = s.result(r, mc=16)
            cor, ref, recs = _oracle(orc, x, a, ml)
            mc = ref.mc()
            assert np.array_equal(adj, ref.adj()), f"{label}: adjacency differs"
            assert np.array_equal(sep[:, :, :mc], ref.sepsets()), f"{label}: sepsets differ"
            assert (sep[:, :, mc:] == -1).all()
            assert info["tests"] == sum(o["tests"] for o in recs), label


def test_batch_autotune_keeps_results(hpc, orc):
    """The batch tuner times each level's kernel on the batch's own steps and
    leaves every instance at its start: running the same batch afterwards,
    and another batch with the tuned shapes, gives the oracle's results."""
    inputs = _mixed_inputs(hpc)[:8]
    with hpc.Batch() as b:
        ids = [b.add(x, a, ml) for _, x, a, ml in inputs]
        tune = b.autotune()
        print(f"[batch] tuned: {[(t['level'], t['ranks_per_task'], t['tasks_per_block'], t['queue'], t['edges_per_block']) for t in tune]}", flush=True)
        levels = [t["level"] for t in tune]
        assert levels == sorted(set(levels)) and 1 in levels and 2 in levels
        for t in tune:
            if t["level"] >= 2:
                assert t["queue"] in (1, 2, 3) and t["ms_best"] <= t["ms_default"] + 1e-9
        for iid in ids:
            finished, lv, err = b.status(iid)
            assert not finished and lv == 0 and err == 0
        b.run()
        for iid, (label, x, a, ml) in zip(ids, inputs):
            _check_instance(b, iid, f"{label}/tuned-self", orc, x, a, ml)
    with hpc.Batch() as b2:
        b2.apply_tuning(tune)
        ids = [b2.add(x, a, ml) for _, x, a, ml in inputs]
        b2.run()
        for iid, (label, x, a, ml) in zip(ids, inputs):
            _check_instance(b2, iid, f"{label}/tuned", orc, x, a, ml)
