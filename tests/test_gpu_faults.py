"""Failure isolation, forced from the host: every kernel runs and exits
normally; only the host side of one instance's level fails.

Bar: the failed instance (or served request) reports its error, and every
other one finishes with the oracle's adjacency, sepsets and test counts.  A
level whose conditioning sets cannot be planned fails with HPC_ELEVEL before
any search runs."""
import numpy as np
import pytest

import workloads
from test_gpu_batch import _check_instance, _oracle

pytestmark = pytest.mark.gpu

HPC_ENOMEM, HPC_ELEVEL = -3, -5


def _inputs():
    out = []
    for name in ("c1", "small", "dense", "deep"):
        x, n, alpha, ml = workloads.sem(name)
        out.append((name, x, alpha, ml))
    return out


def test_batch_isolates_a_failed_instance(hpc, orc):
    inputs = _inputs()
    with hpc.Batch() as b:
        ids = [b.add(x, a, ml) for _, x, a, ml in inputs]
        b.inject_failure(ids[1], 2)
        steps = b.run()
        print(f"[faults] batch: {steps} steps", flush=True)
        finished, levels, err = b.status(ids[1])
        assert finished and err == HPC_ENOMEM and levels == 2, (finished, levels, err)
        for k, (iid, (label, x, a, ml)) in enumerate(zip(ids, inputs)):
            if k != 1:
                _check_instance(b, iid, label, orc, x, a, ml)


def test_server_isolates_a_failed_request(hpc, orc):
    inputs = _inputs()
    with hpc.Server() as s:
        s.inject_failure(3, 1)  # the third request (ids count up from 1)
        reqs = [(label, x, a, ml, s.submit(x, a, ml, copy=True)) for label, x, a, ml in inputs]
        assert reqs[2][4] == 3
        for label, x, a, ml, r in reqs:
            assert s.wait(r, 120.0), f"{label} did not finish"
            if r == 3:
                with pytest.raises(hpc.HpcError, match="injected host failure at level 1"):
                    s.result(r)
                continue
            adj, sep, info = s.result(r, mc=16)
            cor, ref, recs = _oracle(orc, x, a, ml)
            assert np.array_equal(adj, ref.adj()), f"{label}: adjacency differs"
            assert np.array_equal(sep[:, :, :ref.mc()], ref.sepsets()), f"{label}: sepsets"
            assert info["tests"] == sum(o["tests"] for o in recs), label
        # the server keeps serving after the failure
        x, n, a, ml = workloads.sem("c1")
        r = s.submit(x, a, ml, copy=True)
        assert s.wait(r, 120.0)
        adj, _, _ = s.result(r)
        assert np.array_equal(adj, _oracle(orc, x, a, ml)[1].adj())


def test_unplannable_level_fails_before_search(hpc):
    """On the complete graph of p = 200, level 9 holds 199 * ceil(C(198, 9) /
    512) > 2^46 tasks per row and level 16's C(198, 16) saturates 64 bits:
    both fail with HPC_ELEVEL, and the state stays usable."""
    p, n = 200, 1000
    rng = np.random.default_rng(1)
    a = rng.standard_normal((p, 3))
    cor = a @ a.T + np.diag(rng.uniform(1.0, 2.0, p))
    d = np.sqrt(np.diag(cor))
    cor = cor / d[:, None] / d[None, :]
    np.fill_diagonal(cor, 1.0)
    with hpc.DeviceSkeleton(p, n, 0.05, max_level=-1) as dev:
        dev.set_corr(cor)
        assert dev.compact() == p - 1
        for level in (9, 16):
            with pytest.raises(hpc.HpcError, match="rc=-5"):
                dev.search_level(level)
        # a plannable level still runs on the same state
        dev.search_level(1)
        dev.commit_level(1)


def test_wait_limit_knob(hpc, orc):
    x, n, alpha, ml = workloads.sem("small")
    with hpc.DeviceSkeleton(x.shape[0], n, alpha, max_level=4) as dev:
        dev.lib.hpc_set_tuning(dev.h, 5, 60)
        assert dev.lib.hpc_set_tuning(dev.h, 5, -1) == -1
        dev.pmcc(x)
        dev.run(4)
        ref = orc.Skeleton(orc.corr(x), n, alpha, max_level=4)
        ref.run(4)
        assert np.array_equal(dev.adj(), ref.adj())
