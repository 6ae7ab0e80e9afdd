"""The benchmark's workloads at their full sizes through the C ABI against the
oracle:

- the precision form's p = 4000 SEM (n = 2000, degree 4, weights +-[0.05, 1],
  alpha 0.01) through hpc_run with no level cap, in the exact and the
  FP32-MFMA correlation modes;
- the batched form's 256 instances (p in [20, 400], n in [200, 5000]) in one
  batch;
- the first 48 requests of the served form's trace (p in [50, 2000], n in
  [500, 10000], levels 0..4) served together.

Bar: adjacency, sepsets and per-level test counts identical to the oracle's;
the correlation bit-identical in the exact mode and within the FP32 form's
stated bound in the other.
"""
import numpy as np
import pytest

import workloads
from test_gpu_batch import _check_instance, _check_list, _oracle, _serve
from test_gpu_precision import EPS

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode", [0, 1])  # HPC_CORR_CANON, HPC_CORR_F32
def test_precision_workload_full_size(hpc, orc, mode):
    x, n, alpha, ml = workloads.sem("c3")
    p = x.shape[0]
    print(f"[c3] p={p} n={n} alpha={alpha} mode={mode}", flush=True)
    cor = orc.corr(x)
    ref = orc.Skeleton(cor, n, alpha, max_level=ml)
    recs = ref.run(ml)
    # the oracle's record of the level that finds no testable row, which
    # hpc_run does not write
    assert recs[-1]["tests"] == 0
    recs = recs[:-1]
    with hpc.DeviceSkeleton(p, n, alpha, max_level=ml) as dev:
        dev.pmcc(x, mode)
        got = dev.cor()
        if mode == 0:
            assert np.array_equal(got, cor)
        else:
            off = ~np.eye(p, dtype=bool)
            assert np.max(np.abs(got - cor)[off]) <= EPS
        st = dev.run(ml)
        print(f"[c3] levels {len(st)}, rechecks {sum(s.get('rechecks', 0) for s in st)}", flush=True)
        assert [s["tests"] for s in st] == [r["tests"] for r in recs]
        assert [s["removed"] for s in st] == [r["removed"] for r in recs]
        assert np.array_equal(dev.adj(), ref.adj())
        assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets())


def test_batched_workload_full_size(hpc, orc):
    insts = workloads.batch_mix(256)
    with hpc.Batch() as b:
        ids = [b.add(x, a, ml) for x, a, ml in insts]
        steps = b.run()
        print(f"[batch256] {steps} steps", flush=True)
        for k, (iid, (x, a, ml)) in enumerate(zip(ids, insts)):
            _check_instance(b, iid, f"mix{k}", orc, x, a, ml)


def test_served_workload_trace(hpc, orc):
    trace = workloads.served_trace(48)
    inputs = [(f"req{k}", x, a, ml) for k, (_, x, a, ml) in enumerate(trace)]
    got, stats = _serve(hpc, inputs)
    print(f"[served48] {stats}", flush=True)
    for label, x, a, ml in inputs:
        cor, ref, recs = _oracle(orc, x, a, ml)
        mc = ref.mc()
        so = ref.sepsets()
        adj, sep, info, lst = got[label]
        assert np.array_equal(adj, ref.adj()), f"{label}: adjacency differs"
        assert np.array_equal(sep[:, :, :mc], so), f"{label}: sepsets differ"
        _check_list(lst, so, label)
        assert info["tests"] == sum(o["tests"] for o in recs), label
