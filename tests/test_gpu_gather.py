"""Rows too long for LDS staging: the gather kernels (search_l1_kernel,
search_ln_kernel) against the oracle.  Level 1 takes the pivot-row kernel
whenever a whole correlation row fits the budget, so budgets below 8 p bytes
drive its staged and gather kernels.

Each level routes row by row: a row whose neighbourhood fits the workgroup's
LDS budget is staged, a longer one is searched by the gather kernel of the
same launch sequence.  The budget is a tuning knob (HPC_TUNE_LDS_LIMIT), so
these tests force the gather kernels on every row (1 byte), on part of the
rows (a few hundred bytes to a few KiB) and, with the default 64 KiB, on the
dense rows of a latent-factor workload whose level-2 rows hold more than a
hundred neighbours.  Bar: the first passing rank of every ordered edge, the
adjacency after every level, the canonical test counts and the sepsets equal
the oracle's, in exact and FP32-MFMA mode, for one state and for a batch.
"""
import numpy as np
import pytest

import workloads
from test_gpu_batch import _check_instance, _single

pytestmark = pytest.mark.gpu

F32 = 1
KIB64 = 64 * 1024


def lds_side(limit, lv):
    """Largest staged block side (row length + 1) within `limit` bytes: the
    binomial rows, the block of doubles and its int ids."""
    d = 0
    while True:
        e = d + 1
        if e * (lv + 1) * 8 + e * e * 8 + ((e * 4 + 7) & ~7) > limit:
            return d
        d = e


def routes(limit, level, max_deg, p):
    """(staged rows possible, gathered rows present) at a level.  Level 1
    stages whole correlation rows (the pivot kernel) when one fits the
    budget, else rows of up to limit / 16 neighbours."""
    if level == 1:
        if p * 8 <= limit:
            return True, False
        cap = min(max_deg, limit // 16)
        return cap >= 2, max_deg > cap
    dmax = min(max_deg + 1, lds_side(limit, level))
    return dmax >= level + 2, max_deg + 1 > dmax


def _run_vs_oracle(hpc, orc, x, n, alpha, ml, limit, mode=0, label=""):
    p = x.shape[0]
    print(f"[gather] {label}: p={p} n={n} alpha={alpha} max_level={ml} lds_limit={limit} "
          f"mode={mode}", flush=True)
    dev = hpc.DeviceSkeleton(p, n, alpha, max_level=ml)
    dev.set_lds_limit(limit)
    dev.pmcc(x, mode)
    cor = orc.corr(x)
    if mode == 0:
        assert np.array_equal(dev.cor(), cor, equal_nan=True)
    ref = orc.Skeleton(cor, n, alpha, max_level=ml)
    gathered = staged = 0
    level = 0
    while ml < 0 or level <= ml:
        d = dev.level(level)
        o = ref.level(level)
        if level >= 1:
            assert np.array_equal(dev.adj_compact(), ref.adj_compact()), (label, level)
            if ml < 0 and o["max_deg"] - 1 < level:
                break
            s, g = routes(limit if limit else KIB64, level, o["max_deg"], p)
            staged += s
            gathered += g
            fp_d, rp_d = dev.first_pass()
            fp_o, rp_o = ref.first_pass()
            assert np.array_equal(rp_d, rp_o)
            assert np.array_equal(fp_d, fp_o), f"{label}: first passing ranks differ at level {level}"
        assert np.array_equal(dev.adj(), ref.adj()), f"{label}: adjacency differs after level {level}"
        for key in ("tests", "tests_max", "removed"):
            assert d[key] == o[key], (label, level, key, d, o)
        level += 1
        if level > 8:
            break
    assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets()), f"{label}: sepsets differ"
    dev.close()
    return staged, gathered


@pytest.mark.parametrize("limit", [1, 300, 600, 1200])
@pytest.mark.parametrize("name", ["small", "dense", "deep"])
def test_forced_gather_exact(hpc, orc, name, limit):
    x, n, alpha, ml = workloads.sem(name)
    staged, gathered = _run_vs_oracle(hpc, orc, x, n, alpha, ml, limit, 0, name)
    assert gathered > 0, "no level took the gather kernel"
    if limit == 1:
        assert staged == 0


@pytest.mark.parametrize("limit", [1, 1200])
def test_forced_gather_fast(hpc, orc, limit):
    x, n, alpha, ml = workloads.sem("small")
    _, gathered = _run_vs_oracle(hpc, orc, x, n, alpha, ml, limit, F32, "small/f32")
    assert gathered > 0


@pytest.mark.parametrize("mode", [0, F32])
def test_dense_rows_take_gather_at_default_limit(hpc, orc, mode):
    """Level-2 rows of more than a hundred neighbours (past the 87 that 64 KiB
    stages) go to the gather kernel while the short rows stay staged."""
    x = workloads.latent()
    staged, gathered = _run_vs_oracle(hpc, orc, x, x.shape[1], 0.05, 2, 0, mode, "latent")
    assert gathered >= 1 and staged >= 1


def test_pinv_fallback_on_gather_path(hpc, orc):
    """Collinear columns send tests to the pseudoinverse form inside the
    gather kernel."""
    x, n, alpha, _ = workloads.sem("small")
    x = x[:60].copy()
    x[7] = x[3]
    x[11] = 2.0 * x[3] - 1.0
    x[20] = x[21] + x[22]
    _run_vs_oracle(hpc, orc, x, n, alpha, -1, 1, 0, "collinear")


@pytest.mark.parametrize("limit", [1, 1200])
def test_batch_forced_gather(hpc, orc, limit):
    """A batch routes every instance's rows the same way: each instance equals
    the oracle and its own single-state run at the default limit."""
    inputs = []
    for name in ("c1", "small", "dense", "deep"):
        x, n, alpha, ml = workloads.sem(name)
        inputs.append((name, x, alpha, ml))
    inputs.append(("latent", workloads.latent(), 0.05, 2))
    singles = [_single(hpc, x, a, ml) for _, x, a, ml in inputs]
    with hpc.Batch() as b:
        b.set_lds_limit(limit)
        # one instance on fast correlations: same results
        ids = [b.add(x, a, ml, corr_mode=F32 if k == 1 else 0)
               for k, (_, x, a, ml) in enumerate(inputs)]
        steps = b.run()
        print(f"[gather] batch lds_limit={limit}: {len(ids)} instances in {steps} steps",
              flush=True)
        for k, (iid, (label, x, a, ml), single) in enumerate(zip(ids, inputs, singles)):
            _check_instance(b, iid, f"{label}/lds{limit}", orc, x, a, ml, single, fast=k == 1)


def test_server_forced_gather(hpc, orc):
    """The server takes the knob for its later steps; results stay the
    oracle's."""
    inputs = [("small", *workloads.sem("small")[0:1], 0.05, 4),
              ("latent", workloads.latent(), 0.05, 2)]
    with hpc.Server() as s:
        s.set_tuning(4, 1)
        reqs = [(label, x, a, ml, s.submit(x, a, ml, copy=True)) for label, x, a, ml in inputs]
        for label, x, a, ml, r in reqs:
            assert s.wait(r, 120.0), f"{label} did not finish"
            adj, sep, info = s.result(r, mc=16)
            cor = orc.corr(x)
            ref = orc.Skeleton(cor, x.shape[1], a, max_level=ml)
            recs = ref.run(ml)
            assert np.array_equal(adj, ref.adj()), f"{label}: adjacency differs"
            assert np.array_equal(sep[:, :, :ref.mc()], ref.sepsets()), f"{label}: sepsets"
            assert info["tests"] == sum(o["tests"] for o in recs), label
