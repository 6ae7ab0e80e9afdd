"""The level's task queue: a persistent device kernel that claims blocks of
(edge, rank range) tasks, alone (HPC_TUNE_QUEUE) and shared with host threads
that take blocks from the other end (hpc_search_level_shared, the `-w`
work-stealing mode).

Bar: first passing ranks, adjacency after every level, canonical test counts
and sepsets equal the oracle's (and the device-only run's) whichever side
ran a block."""
import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


def _shared_vs_oracle(hpc, orc, x, n, alpha, top, threads, tuning=None, label=""):
    p = x.shape[0]
    print(f"[queue] {label}: p={p} n={n} alpha={alpha} levels 0..{top} threads={threads}",
          flush=True)
    dev = hpc.DeviceSkeleton(p, n, alpha, max_level=top)
    if tuning:
        dev.apply_tuning(tuning)
        dev.set_ranks_per_task(64)  # level 1's tasks (its entries set only edges per block)
    dev.pmcc(x)
    cor = orc.corr(x)
    ref = orc.Skeleton(cor, n, alpha, max_level=top)
    dev.level(0)
    ref.level(0)
    host_blocks = 0
    for level in range(1, top + 1):
        dev.compact()
        d, h = dev.search_level_shared(level, threads)
        c = dev.commit_level(level)
        o = ref.level(level)
        host_blocks += h["removed"]
        fp_d, rp_d = dev.first_pass()
        fp_o, rp_o = ref.first_pass()
        assert np.array_equal(rp_d, rp_o)
        assert np.array_equal(fp_d, fp_o), f"{label}: first passing ranks differ at level {level}"
        assert np.array_equal(dev.adj(), ref.adj()), f"{label}: adjacency differs at level {level}"
        for key in ("tests", "tests_max", "removed"):
            assert c[key] == o[key], (label, level, key, c, o)
    assert np.array_equal(dev.sepsets(ref.mc()), ref.sepsets()), f"{label}: sepsets differ"
    dev.close()
    return host_blocks


# one task per block of at most 64 sets: host threads finish blocks while the
# device works through the rest
SMALL_BLOCKS = [{"level": lv, "ranks_per_task": 64, "tasks_per_block": 1} for lv in range(2, 7)]


def test_shared_queue_matches_oracle(hpc, orc):
    host = 0
    for name, top in (("small", 5), ("dense", 5), ("deep", 6)):
        x, n, alpha, _ = workloads.sem(name)
        host += _shared_vs_oracle(hpc, orc, x, n, alpha, top, 8, label=name)
        host += _shared_vs_oracle(hpc, orc, x, n, alpha, top, 16, SMALL_BLOCKS, f"{name}/small")
    x = workloads.latent()
    host += _shared_vs_oracle(hpc, orc, x, x.shape[1], 0.05, 2, 16, SMALL_BLOCKS, "latent")
    print(f"[queue] blocks finished on the host: {host}", flush=True)
    assert host > 0, "the host side never finished a block"


def test_shared_queue_with_gather_rows(hpc, orc):
    """Rows too long for LDS go to the gather kernel beside the queue."""
    x, n, alpha, _ = workloads.sem("small")
    p = x.shape[0]
    dev = hpc.DeviceSkeleton(p, n, alpha, max_level=4)
    dev.set_lds_limit(1200)
    dev.pmcc(x)
    ref = orc.Skeleton(orc.corr(x), n, alpha, max_level=4)
    dev.level(0)
    ref.level(0)
    for level in range(1, 5):
        dev.compact()
        dev.search_level_shared(level, 4)
        dev.commit_level(level)
        ref.level(level)
        assert np.array_equal(dev.first_pass()[0], ref.first_pass()[0]), level
        assert np.array_equal(dev.adj(), ref.adj()), level


@pytest.mark.parametrize("name", ["small", "c2k"])
def test_device_queue_on_off_same_results(hpc, name):
    x, n, alpha, ml = workloads.sem(name)
    ml = 4
    out = []
    for on in (True, False):
        with hpc.DeviceSkeleton(x.shape[0], n, alpha, max_level=ml) as dev:
            dev.set_queue(on)
            dev.pmcc(x)
            st = dev.run(ml)
            out.append((dev.adj(), dev.sepsets(), [s["tests"] for s in st]))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert out[0][2] == out[1][2]


def test_queue_back_to_back_levels_and_batch(hpc, orc):
    """The queue head resets itself after every launch: many launches in a
    row (a batch step runs one per level present) keep every result."""
    from test_gpu_batch import _check_instance

    inputs = []
    for name in ("c1", "small", "dense", "deep"):
        x, n, alpha, ml = workloads.sem(name)
        inputs.append((name, x, alpha, ml))
    with hpc.Batch() as b:
        b.lib.hpc_batch_set_tuning(b.h, 6, 1)
        ids = [b.add(x, a, ml) for _, x, a, ml in inputs]
        b.step()
        b.step()
        # instances at different levels from here on
        late = b.add(*workloads.sem("small")[0:1], 0.05, 4)
        b.run()
        for iid, (label, x, a, ml) in zip(ids, inputs):
            _check_instance(b, iid, label, orc, x, a, ml)
        x, _, _, _ = workloads.sem("small")
        _check_instance(b, late, "small/late", orc, x, 0.05, 4)
