"""Benchmark of the PC-stable skeleton hot path on one MI355X.

A step is one full skeleton of the plain p=2000 workload on device-resident
observations: correlation (moments + pairwise sums), then levels 0..4
(compaction, CI tests, removal) -- the reference's default maximum level
(main.cpp:31), at which this workload still holds 30-neighbour rows.  The
level loop of each timed step is a replay of the hipGraph captured from the
first warm-up run (the graph's launch sizes are checked against what the
device computed; a mismatch would re-run the levels eagerly).

value = CI tests/s = sum over levels of the canonical test count T_l divided
by the level wall time (the reference's "Execution duration" scope,
skeleton.cpp:167-180).  Beside it: every kernel of the step against the
roofline that bounds it, the HBM copy bandwidth measured here, a second
plain leg at p=1000 with no level cap, the FP32-MFMA end-to-end run of the
same p=2000 skeleton, the batched and served forms, and the reference's
OpenMP CPU path (median of three runs).  Prints one JSON line.

`python bench.py --copy-probe` runs only device-to-device copies of a known
size (the counter calibration of the committed rocprofv3 summaries).
"""
from __future__ import annotations

import argparse
import json

import numpy as np
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
for _p in (REPO, os.path.join(REPO, "masterthesis-heterogeneous-pc_amd"), os.path.join(REPO, "tests")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "CI tests/sec + skeleton wall time at p=2000; served p50/p99 vs plain path"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def flops_per_test(level: int) -> float:
    """FP64 work of one level >= 2 test in the LDL^T / Schur form: the
    factorisation of the l x l conditioning block, two triangular solves,
    the 2 x 2 Schur block and the statistic, l^3/3 + 2 l^2 + 6 l + 30."""
    return level ** 3 / 3.0 + 2.0 * level ** 2 + 6.0 * level + 30.0


def hbm_copy(nbytes: int = 1 << 30, reps: int = 10, probe_only: bool = False) -> dict:
    """HBM bandwidth of a device-to-device copy (the HIP runtime's copy kernel,
    through the runtime libheterogpc_hip.so links): 2 * nbytes moved per
    copy, timed with HIP events on the stream the copies go to."""
    import ctypes as C

    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    hip = C.CDLL(os.path.join(rocm, "lib", "libamdhip64.so.7"))
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
    hip.hipEventSynchronize.argtypes = [C.c_void_p]
    hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
    hip.hipEventDestroy.argtypes = [C.c_void_p]
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), nbytes) or hip.hipMalloc(C.byref(dst), nbytes):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, nbytes):
            raise RuntimeError("hipMemset failed")
        for _ in range(2):
            hip.hipMemcpyAsync(dst, src, nbytes, 3, None)  # hipMemcpyDeviceToDevice
        e0, e1 = C.c_void_p(), C.c_void_p()
        hip.hipEventCreate(C.byref(e0))
        hip.hipEventCreate(C.byref(e1))
        hip.hipEventRecord(e0, None)
        for _ in range(reps):
            if hip.hipMemcpyAsync(dst, src, nbytes, 3, None):
                raise RuntimeError("hipMemcpyAsync failed")
        hip.hipEventRecord(e1, None)
        hip.hipEventSynchronize(e1)
        ms = C.c_float()
        hip.hipEventElapsedTime(C.byref(ms), e0, e1)
        hip.hipEventDestroy(e0)
        hip.hipEventDestroy(e1)
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    per = ms.value / reps
    return {"bytes_per_copy": 2 * nbytes, "ms_per_copy": per,
            "GBs": 2 * nbytes / (per * 1e-3) / 1e9,
            "note": "device-to-device hipMemcpyAsync of %d MiB, read + write" % (nbytes >> 20)}


def bytes_per_test(level: int) -> int:
    """Algorithmic bytes of one CI test: level 0 reads one correlation (8 B)
    and one adjacency entry (4 B); level l >= 1 reads the (l+2)(l+1)/2
    off-diagonal correlations of its submatrix and l neighbour ids."""
    if level == 0:
        return 12
    return 8 * (level + 2) * (level + 1) // 2 + 4 * level


PROFILE = os.path.join(REPO, "profiles", "skeleton_p2000_kernels.md")


def _profiled_traffic() -> dict:
    """Calibrated HBM bytes per step of the CI-test kernels and the
    correlation, from the committed rocprofv3 summary of this workload."""
    try:
        with open(PROFILE) as f:
            for line in f:
                if line.startswith("machine-readable: "):
                    d = json.loads(line[len("machine-readable: "):])
                    d["source"] = os.path.relpath(PROFILE, REPO)
                    return d
    except OSError:
        pass
    return {}


def run_step(dev):
    t0 = time.perf_counter()
    dev.reset()
    dev.corr()
    ms_mom, ms_cor = dev.corr_ms()
    t1 = time.perf_counter()
    stats = dev.run(-2)
    dev.sync()
    t2 = time.perf_counter()
    return {
        "e2e_ms": (t2 - t0) * 1e3,
        "levels_ms": (t2 - t1) * 1e3,
        "ms_moments": ms_mom,
        "ms_corr": ms_cor,
        "levels": stats,
    }


def host_cpu() -> dict:
    """Model, sockets, cores and threads of the host (lscpu)."""
    import subprocess

    info = {}
    try:
        txt = subprocess.run(["lscpu"], stdout=subprocess.PIPE, text=True, timeout=10).stdout
        for line in txt.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"),
            "threads_per_core": info.get("Thread(s) per core"),
            "cpus_online": info.get("CPU(s)"), "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(x, n, alpha, max_level, gpu_adj, runs=3):
    """The oracle's reference-faithful CPU path (per-test pseudoinverse,
    immediate deletion with early exit, OpenMP `collapse(2)
    schedule(dynamic, 10)` over rows sorted by length, cpuExecutor.cpp:88-109)
    on all host threads of this box, on the same workload: median of `runs`
    runs of the level loop (the correlation once, as the reference's timer
    excludes it).  value = the CPU's own tests (it stops testing an edge once
    either side removed it) / its level time.  The CPU's adjacency is
    compared with the GPU's."""
    host = host_cpu()  # before OpenMP binds this thread to its first place
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    t0 = time.perf_counter()
    cor = oracle.corr(x, threads=threads)
    t_corr = time.perf_counter() - t0
    times, tests, adj = [], [], None
    for _ in range(runs):
        sk = oracle.Skeleton(cor, n, alpha, max_level=max_level, mode=oracle.MODE_PINV,
                             flags=oracle.FLAG_REF_EARLY_EXIT, threads=threads)
        t1 = time.perf_counter()
        lv = sk.run(max_level)
        times.append(time.perf_counter() - t1)
        tests.append(sum(l["tests"] for l in lv))
        adj = sk.adj()
    k = int(np.argsort(times)[len(times) // 2])
    diff = int(np.count_nonzero(np.triu(adj != gpu_adj, 1)))
    return {
        "value": tests[k] / times[k],
        "unit": "CI tests/s",
        "cores": threads,
        "kind": "port",
        "sample": "the whole p=2000 skeleton, levels 0..%d, median of %d runs (%s s); the "
                  "correlation once (%.2f s)" % (max_level, runs,
                                                 ", ".join("%.2f" % t for t in times), t_corr),
        "runs_s": times,
        "tests_per_run": tests,
        "skeleton_ms_levels": times[k] * 1e3,
        "skeleton_ms_e2e": (times[k] + t_corr) * 1e3,
        "adjacency_identical_to_gpu": diff == 0,
        "adjacency_edges_differing": diff,
        "omp": {"OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
                "OMP_PLACES": os.environ.get("OMP_PLACES")},
        "host": host,
    }


def _pct(vals, q):
    import numpy as np

    return float(np.percentile(np.asarray(vals, dtype=np.float64), q)) if vals else 0.0


def _plain_skeleton(heterogpc, x, alpha, max_level):
    """One instance through the plain single-instance path: state, upload,
    correlation, levels, results (adjacency matrix and sepset list) to the
    host."""
    p, n = x.shape
    with heterogpc.DeviceSkeleton(p, n, alpha, max_level=max_level) as dev:
        dev.pmcc(x)
        stats = dev.run(-2)
        return dev.adj(), dev.sepset_list(), stats


def bench_batched(heterogpc, insts):
    """Batched form: the same instances one per launch sequence (plain) and all
    in one batch; skeletons/hour of both and whether every result agrees.
    The batch's launch shapes come from the batch autotuner, run beforehand
    on a separate batch of the same instances (untimed)."""
    for x, a, ml in insts[:4]:
        _plain_skeleton(heterogpc, x, a, ml)
    with heterogpc.Batch() as b:
        for x, a, ml in insts:
            b.add(x, a, ml)
        tuned = b.autotune()
    heterogpc.load().hpc_device_synchronize(0)
    t0 = time.perf_counter()
    plain = [_plain_skeleton(heterogpc, x, a, ml) for x, a, ml in insts]
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    with heterogpc.Batch() as b:
        b.apply_tuning(tuned)
        ids = [b.add(x, a, ml) for x, a, ml in insts]
        steps = b.run()
        got = [(b.adj(i), b.sepset_list(i)) for i in ids]
    t_batch = time.perf_counter() - t0
    same = all(np.array_equal(ga, pa) and np.array_equal(gs, ps)
               for (ga, gs), (pa, ps, _) in zip(got, plain))
    tests = sum(lv["tests"] for _, _, st in plain for lv in st)
    k = len(insts)
    return {
        "workload": "%d SEM instances, p U[20,400], n U[200,5000], degree U[2,6], alpha "
                    "{0.01,0.05}, FP32 data, levels until no edge has enough neighbours" % k,
        "instances": k,
        "plain_s": t_plain,
        "batched_s": t_batch,
        "batched_steps": steps,
        "plain_skeletons_per_hour": k / t_plain * 3600.0,
        "batched_skeletons_per_hour": k / t_batch * 3600.0,
        "speedup": t_plain / t_batch,
        "tests_total": tests,
        "batched_ci_tests_per_s": tests / t_batch,
        "results_identical": bool(same),
        "scope": "host observations -> device, correlation, all levels, adjacency matrix "
                 "and sepset list back on the host",
        "autotune": [{k: t[k] for k in ("level", "ranks_per_task", "tasks_per_block", "queue",
                                        "edges_per_block", "ms_best", "ms_default")}
                     for t in tuned],
    }


def _serve_trace(heterogpc, reqs, arrivals):
    """Replays (x, alpha, max_level) requests arriving at `arrivals` (s)
    through the continuous-batching server.  Returns per-request
    (adj, sepset list, completion s) and the server's statistics."""
    import threading

    n_req = len(reqs)
    sub = [None] * n_req
    out = [None] * n_req
    with heterogpc.Server() as srv:
        t0 = time.perf_counter()

        def feeder():
            for k, ((x, a, ml), ta) in enumerate(zip(reqs, arrivals)):
                d = ta - (time.perf_counter() - t0)
                if d > 0:
                    time.sleep(d)
                sub[k] = srv.submit(x, a, ml)

        th = threading.Thread(target=feeder)
        th.start()
        left = set(range(n_req))
        while left:
            got = 0
            for k in sorted(left):
                if sub[k] is None:
                    break
                if srv.wait(sub[k], 0.0):
                    adj, lst, _ = srv.result(sub[k], dense_sepsets=False)
                    out[k] = (adj, lst, time.perf_counter() - t0)
                    left.discard(k)
                    got += 1
            if not got:
                time.sleep(2e-4)
        th.join()
        stats = srv.stats()
    return out, stats


def _plain_trace(heterogpc, reqs, arrivals):
    """The same trace through the plain path, one request at a time in
    arrival order."""
    t0 = time.perf_counter()
    out = []
    for (x, a, ml), ta in zip(reqs, arrivals):
        d = ta - (time.perf_counter() - t0)
        if d > 0:
            time.sleep(d)
        adj, lst, _ = _plain_skeleton(heterogpc, x, a, ml)
        out.append((adj, lst, time.perf_counter() - t0))
    return out


def bench_served(heterogpc, trace, load, label):
    """Served form.  The plain path's capacity C (requests/s, one at a time)
    is measured on the requests first; Poisson arrivals at `load` x C are
    then replayed in real time through the continuous-batching server and
    through the plain path.  Latency = arrival -> adjacency matrix and sepset
    list in the caller's hands.  Saturation (every request at t = 0) gives
    each path's skeletons/hour."""
    reqs = [(x, a, ml) for _, x, a, ml in trace]
    gaps = np.diff([0.0] + [t for t, _, _, _ in trace])  # unit-rate exponential gaps
    gaps = gaps / gaps.mean()
    n_req = len(reqs)
    _plain_skeleton(heterogpc, *reqs[0])
    sat_plain = _plain_trace(heterogpc, reqs, [0.0] * n_req)
    plain_sat_s = sat_plain[-1][2]
    rate = load * n_req / plain_sat_s
    arrivals = list(np.cumsum(gaps) / rate)
    sat_served, _ = _serve_trace(heterogpc, reqs, [0.0] * n_req)
    served_sat_s = max(v[2] for v in sat_served)
    served, stats = _serve_trace(heterogpc, reqs, arrivals)
    plain = _plain_trace(heterogpc, reqs, arrivals)
    s_lat = [(served[k][2] - arrivals[k]) * 1e3 for k in range(n_req)]
    p_lat = [(plain[k][2] - arrivals[k]) * 1e3 for k in range(n_req)]
    same = all(np.array_equal(sa, pa) and np.array_equal(ss, ps)
               for (sa, ss, _), (pa, ps, _) in zip(served + sat_served, plain + sat_plain))
    return {
        "workload": "%d requests, %s; Poisson arrivals at %.1f/s = %.2f x the plain path's "
                    "capacity" % (n_req, label, rate, load),
        "requests": n_req,
        "arrival_rate_per_s": rate,
        "served_p50_ms": _pct(s_lat, 50),
        "served_p99_ms": _pct(s_lat, 99),
        "plain_p50_ms": _pct(p_lat, 50),
        "plain_p99_ms": _pct(p_lat, 99),
        "served_skeletons_per_hour_at_rate": n_req / max(v[2] for v in served) * 3600.0,
        "plain_skeletons_per_hour_at_rate": n_req / plain[-1][2] * 3600.0,
        "served_skeletons_per_hour_saturated": n_req / served_sat_s * 3600.0,
        "plain_skeletons_per_hour_saturated": n_req / plain_sat_s * 3600.0,
        "saturated_speedup": plain_sat_s / served_sat_s,
        "server": {k: stats[k] for k in ("steps", "peak_active", "peak_bytes", "budget", "busy_ms")},
        "results_identical": bool(same),
    }


def bench_plain(heterogpc, x, n, alpha, max_level, mode=0, steps=5):
    """A plain leg: one state, device-resident observations, `steps` timed
    skeletons (correlation in `mode`, then the levels), after one untimed
    one.  Returns CI tests/s, level and end-to-end wall times (medians) and
    the last run's adjacency and sepset list."""
    p = x.shape[0]
    with heterogpc.DeviceSkeleton(p, n, alpha, max_level=max_level) as dev:
        dev.set_kernel_timing(False)
        dev.set_observations(x)
        dev.corr(mode)
        dev.run(-2)
        lev, e2e, st = [], [], None
        for _ in range(steps):
            dev.reset()
            t0 = time.perf_counter()
            dev.corr(mode)
            t1 = time.perf_counter()
            st = dev.run(-2)
            dev.sync()
            t2 = time.perf_counter()
            lev.append((t2 - t1) * 1e3)
            e2e.append((t2 - t0) * 1e3)
        tests = int(sum(l["tests"] for l in st))
        lv_ms = float(np.median(lev))
        return {"ci_tests_per_s": tests / (lv_ms * 1e-3), "tests": tests,
                "levels_run": len(st), "skeleton_ms_levels": lv_ms,
                "skeleton_ms_e2e": float(np.median(e2e)),
                "rechecks": int(sum(l["rechecks"] for l in st)),
                "edges_left": int(st[-1]["edges_left"]) if st else 0}, dev.adj(), dev.sepset_list()


FP32_MFMA_PEAK_TFLOPS = 157.3


def bench_precision(heterogpc, x, n, alpha, max_level, steps=3):
    """Precision form: the same skeleton from the exact FP64 correlation and
    from the MFMA FP32 one with exact rechecks.  Reports the correlation
    kernels' time and throughput, the max |fast - exact| correlation error,
    the rechecks, and whether the skeletons are identical."""
    p = x.shape[0]
    out = {}
    res = {}
    for mode, name in ((0, "fp64_exact"), (1, "f32_mfma")):
        with heterogpc.DeviceSkeleton(p, n, alpha, max_level=max_level) as dev:
            dev.set_observations(x)
            dev.corr(mode)
            dev.run(-2)  # warm-up
            cors, e2e, lev = [], [], []
            for _ in range(steps):
                dev.reset()
                t0 = time.perf_counter()
                dev.corr(mode)
                ms_m, ms_c = dev.corr_ms()
                t1 = time.perf_counter()
                st = dev.run(-2)
                dev.sync()
                t2 = time.perf_counter()
                cors.append(ms_c)
                e2e.append((t2 - t0) * 1e3)
                lev.append((t2 - t1) * 1e3)
            dev.reset()
            dev.corr(mode)
            cor = dev.cor()
            st = dev.run(-2)
            res[name] = (cor, dev.adj(), dev.sepset_list())
            flops = 2.0 * n * p * (p + 1) / 2.0
            ms_c = float(np.median(cors))
            out[name] = {"corr_ms": ms_c, "corr_tflops": flops / (ms_c * 1e-3) / 1e12,
                         "levels_ms": float(np.median(lev)), "e2e_ms": float(np.median(e2e)),
                         "tests": int(sum(l["tests"] for l in st)),
                         "rechecks": int(sum(l["rechecks"] for l in st))}
    out["f32_mfma"]["corr_frac_of_fp32_mfma_peak"] = \
        out["f32_mfma"]["corr_tflops"] / FP32_MFMA_PEAK_TFLOPS
    out["fp64_exact"]["corr_frac_of_fp64_peak"] = \
        out["fp64_exact"]["corr_tflops"] / FP64_PEAK_TFLOPS
    ce, cf = res["fp64_exact"][0], res["f32_mfma"][0]
    off = ~np.eye(p, dtype=bool)
    d = np.abs(cf - ce)[off]
    big = np.abs(ce)[off] >= 0.1
    out["max_abs_cor_error"] = float(np.nanmax(d))
    out["max_rel_cor_error_abs_r_ge_0.1"] = float(np.nanmax(d[big] / np.abs(ce[off][big])))
    out["cor_error_bound"] = 74.0 * 2.0 ** -24 * 1.01
    out["skeleton_identical"] = bool(np.array_equal(res["fp64_exact"][1], res["f32_mfma"][1]) and
                                     np.array_equal(res["fp64_exact"][2], res["f32_mfma"][2]))
    out["workload"] = ("SEM p=%d, n=%d, expected degree 4, weights +-U[0.05,1], alpha=%g, FP32 "
                       "data, levels 0..%d" % (p, n, alpha, max_level))
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2k")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-count", type=int, default=256,
                    help="instances of the batched form (0 = skip)")
    ap.add_argument("--served-count", type=int, default=200,
                    help="requests of the served form (0 = skip)")
    ap.add_argument("--precision", type=int, default=1,
                    help="run the precision form (p=4000 SEM, exact vs MFMA FP32)")
    ap.add_argument("--served-load", type=float, default=0.7,
                    help="arrival rate of the served form as a fraction of the plain "
                         "path's measured capacity")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only device-to-device copies of a known size (counter calibration)")
    ap.add_argument("--legs", type=int, default=1,
                    help="the p=1000 uncapped and the FP32 p=2000 plain legs")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(every instance runs on a single MI355X).", file=sys.stderr)
        return 2
    if args.copy_probe:
        print(json.dumps(hbm_copy(reps=20)))
        return 0
    # the CPU baseline's OpenMP placement (read when the oracle library loads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")

    import heterogpc
    import workloads

    x, n, alpha, max_level = workloads.sem(args.workload)
    p = x.shape[0]
    dev = heterogpc.DeviceSkeleton(p, n, alpha, max_level=max_level, device=0)
    dev.set_observations(x)  # inputs resident in HBM before timing
    # launch shapes per level from the autotuner, before any timing
    dev.corr()
    tuned = dev.autotune(max_level if max_level >= 0 else 16)

    # The timed steps run the level loop as a caller does, without the
    # per-level kernel-time events (each event marker between two kernels
    # delays the second by a few microseconds).
    dev.set_kernel_timing(False)
    for _ in range(args.warmup):
        run_step(dev)
    heterogpc.load().hpc_device_synchronize(0)
    t0 = time.perf_counter()
    steps = [run_step(dev) for _ in range(args.steps)]
    heterogpc.load().hpc_device_synchronize(0)
    wall = time.perf_counter() - t0
    graph = dev.graph_stats()

    tests_total = sum(lv["tests"] for lv in steps[-1]["levels"])
    levels_ms = sum(s["levels_ms"] for s in steps) / len(steps)
    e2e_ms = wall * 1e3 / len(steps)
    value = tests_total / (levels_ms * 1e-3)

    # Kernel times (roofline, per-level figures): the same number of steps
    # again with the events on, after one untimed step.
    dev.set_kernel_timing(True)
    run_step(dev)
    ksteps = [run_step(dev) for _ in range(args.steps)]
    last = ksteps[-1]["levels"]
    if sum(lv["tests"] for lv in last) != tests_total:
        raise RuntimeError("the kernel-timing pass counted different tests")
    levels_ms_events = sum(s["levels_ms"] for s in ksteps) / len(ksteps)

    # per-kernel device time, averaged over the kernel-timing steps
    def avg(fn):
        return sum(fn(s) for s in ksteps) / len(ksteps)

    k_corr = avg(lambda s: s["ms_corr"])
    k_mom = avg(lambda s: s["ms_moments"])
    k_test = avg(lambda s: sum(lv["ms_test"] for lv in s["levels"]))
    k_compact = avg(lambda s: sum(lv["ms_compact"] for lv in s["levels"]))
    k_commit = avg(lambda s: sum(lv["ms_commit"] for lv in s["levels"]))

    def lvl_avg(key, level):
        return avg(lambda s: sum(lv[key] for lv in s["levels"] if lv["level"] == level))

    test_bytes = sum(lv["tests"] * bytes_per_test(lv["level"]) for lv in last)
    ci_gbs = test_bytes / (k_test * 1e-3) / 1e9 if k_test > 0 else 0.0
    corr_flops = 2.0 * n * p * (p + 1) / 2.0  # distinct pairs x 2n
    corr_tflops = corr_flops / (k_corr * 1e-3) / 1e12 if k_corr > 0 else 0.0
    copy = hbm_copy()
    traffic = _profiled_traffic()

    # Every kernel of the step against the roofline that bounds it.  The
    # level >= 2 searches read their operands from an LDS-staged block, so
    # their algorithmic bytes (B(l) per test) exceed what HBM could deliver
    # in their time: they are bound by FP64 issue, F(l) flops per canonical
    # test; their HBM fraction at algorithmic bytes is given beside it.
    npad = (n + 63) // 64 * 64
    rooflines = []

    def entry(kernel, bound, achieved, peak, unit, ms, work, **kw):
        d = {"kernel": kernel, "bound": bound, "achieved": achieved, "peak": peak, "unit": unit,
             "frac": achieved / peak if peak else None, "ms": ms, "work_per_launch": work}
        d.update(kw)
        rooflines.append(d)

    # NaN-free data: no NaN mask is written (only the masked kernel reads it)
    mom_bytes = 8 * p * n + 8 * p * npad + 16 * p
    entry("stats_kernel + zscore_kernel", "hbm", mom_bytes / (k_mom * 1e-3) / 1e9 if k_mom else 0.0,
          HBM_PEAK_GBS, "GB/s", k_mom, "%d B: observations read once, z-scores, means and standard "
          "deviations written (the two kernels read the observations twice)" % mom_bytes)
    entry("corr_xmfma_kernel", "mfma", corr_tflops, FP64_PEAK_TFLOPS, "TFLOP/s", k_corr,
          "%.3e flops: 2n per pair, p(p+1)/2 pairs" % corr_flops)
    for lv in last:
        l, T = lv["level"], lv["tests"]
        ms = lvl_avg("ms_test", l)
        if ms <= 0:
            continue
        if l == 0:
            entry("level0_kernel + level0_fold_kernel", "hbm", 12 * T / (ms * 1e-3) / 1e9,
                  HBM_PEAK_GBS, "GB/s", ms, "%d pairs x 12 B" % T, level=0)
        elif l == 1:
            entry("search_l1_pivot_kernel", "hbm", 28 * T / (ms * 1e-3) / 1e9, HBM_PEAK_GBS,
                  "GB/s", ms, "%d canonical tests x 28 B (operands from an LDS-staged "
                  "correlation row and L2-resident compacted rows)" % T, level=1)
        else:
            fl = flops_per_test(l) * T
            hb = bytes_per_test(l) * T / (ms * 1e-3) / 1e9
            entry("search_ls_kernel<%d>" % l, "fp64", fl / (ms * 1e-3) / 1e12,
                  FP64_PEAK_TFLOPS, "TFLOP/s", ms,
                  "%d canonical tests x %.1f flops" % (T, flops_per_test(l)), level=l,
                  hbm_algorithmic_GBs=hb, hbm_algorithmic_frac=hb / HBM_PEAK_GBS)
        if l >= 1:
            mc, mm = lvl_avg("ms_compact", l), lvl_avg("ms_commit", l)
            rooflines.append({"kernel": "compaction + removal, level %d" % l, "bound": "latency",
                              "ms_compact": mc, "ms_commit": mm, "edges": lv["edges"],
                              "note": "dependent loads and atomics per row / edge; a few "
                                      "microseconds of publication each"})

    # The headline: the step's dominant kernel (the exact correlation, about
    # 45% of the step) on its own launch time, HIP events on the engine
    # stream.  Its per-launch HBM bytes come from the committed rocprofv3
    # FETCH_SIZE / WRITE_SIZE summary of this command.
    roof = {"kernel": "corr_xmfma_kernel", "bound": "mfma", "achieved": corr_tflops,
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": corr_tflops / FP64_PEAK_TFLOPS,
            "traffic": traffic.get("corr_traffic_bytes"),
            "algorithmic": "%.3e flops per launch (n p (p+1))" % corr_flops,
            "traffic_source": traffic.get("source"),
            "note": "FP64 matrix cores (v_mfma_f64_16x16x4f64: each output is four fused "
                    "multiply-adds in sample order, so the canonical summation order is kept "
                    "bit for bit); 78.6 TFLOP/s is MI355X's dense FP64 MFMA peak"}

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "CI tests/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": e2e_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic linear-Gaussian SEM (seeded), FP32-rounded samples",
        "config": {"workload": "plain PC-stable at p=2000: SEM p=2000, n=5000, expected "
                               "degree 3, alpha=0.05, levels 0..4 (max level 4)"
                               if args.workload == "c2k" else args.workload,
                   "p": p, "n": n, "alpha": alpha, "max_level": max_level},
        "skeleton_ms_levels": levels_ms,
        "skeleton_ms_e2e": e2e_ms,
        "kernel_timing_pass": {
            "note": "value, ms_per_step and skeleton_ms_* come from the timed steps, run "
                    "without per-level HIP events; kernel_ms, levels[].ms_* and roofline from "
                    "a second pass of the same steps with the events on",
            "skeleton_ms_levels_with_events": levels_ms_events},
        "level_loop": {
            "graph": graph,
            "note": "each timed step resets the state and runs correlation then levels 0..4; the "
                    "first warm-up run records the level loop's launch sizes and captures it "
                    "into a hipGraph, later runs replay it (one launch, one host wait) and check "
                    "every level's sizes against the recorded ones (a mismatch re-runs eagerly); "
                    "the kernel-timing pass runs the loop eagerly"},
        "tests_total": tests_total,
        "levels": [{k: lv[k] for k in ("level", "max_deg", "tests", "tests_max", "edges",
                                       "removed", "edges_left", "ms_compact", "ms_test",
                                       "ms_commit", "ms_wall")} for lv in last],
        "kernel_ms": {"moments": k_mom, "corr": k_corr, "ci_tests": k_test,
                      "compact": k_compact, "commit": k_commit},
        "ci_tests_algorithmic": {"bytes_per_step": test_bytes, "GBs": ci_gbs,
                                 "note": "sum of B(l) x T_l over the levels' searches; above "
                                         "HBM's rate because levels >= 2 reuse an LDS-staged "
                                         "block (rooflines has each kernel's own bound)"},
        "level_overhead_ms": levels_ms - k_test,
        "corr_tflops": corr_tflops,
        "hbm_copy": copy,
        "autotune": tuned,
        "roofline": roof,
        "rooflines": rooflines,
    }
    gpu_adj = dev.adj()
    dev.close()
    if args.legs:
        print("[bench] p=1000 uncapped and FP32 p=2000 legs", file=sys.stderr, flush=True)
        x2, n2, a2, ml2 = workloads.sem("c2")
        c2, _, _ = bench_plain(heterogpc, x2, n2, a2, ml2)
        c2["workload"] = ("plain PC-stable at p=1000: SEM p=1000, n=5000, expected degree 3, "
                          "alpha=0.05, levels until no edge has enough neighbours")
        out["c2_uncapped"] = c2
        del x2
        f32, f32_adj, f32_list = bench_plain(heterogpc, x, n, alpha, max_level, mode=1)
        ex, ex_adj, ex_list = bench_plain(heterogpc, x, n, alpha, max_level, mode=0)
        f32["workload"] = ("the headline p=2000 skeleton with the FP32-MFMA correlation and "
                           "exact rechecks (HPC_CORR_F32)")
        f32["exact_same_run"] = {k: ex[k] for k in ("ci_tests_per_s", "skeleton_ms_levels",
                                                    "skeleton_ms_e2e")}
        f32["skeleton_identical_to_exact"] = bool(np.array_equal(f32_adj, ex_adj) and
                                                  np.array_equal(f32_list, ex_list))
        out["c2k_f32_mfma"] = f32
    if args.precision:
        print("[bench] precision form", file=sys.stderr, flush=True)
        xc, nc, ac, _ = workloads.sem("c3")
        out["precision"] = bench_precision(heterogpc, xc, nc, ac, 4)
        del xc
    if args.batch_count > 0:
        print("[bench] batched form", file=sys.stderr, flush=True)
        out["batched"] = bench_batched(heterogpc, workloads.batch_mix(args.batch_count))
    if args.served_count > 0:
        print("[bench] served form", file=sys.stderr, flush=True)
        out["served"] = bench_served(
            heterogpc, workloads.served_trace(args.served_count, 1.0), args.served_load,
            "p log-uniform [50,2000], n U[500,10000], degree U[2,5], alpha {0.01,0.05}, FP32 "
            "data, max level 4")
    if args.batch_count > 0 and args.served_count > 0:
        print("[bench] served form, batched mix", file=sys.stderr, flush=True)
        mix = workloads.batch_mix(args.batch_count)
        rng = np.random.default_rng(4)
        t = np.cumsum(rng.exponential(1.0, len(mix)))
        out["served_c4"] = bench_served(
            heterogpc, [(float(ta), xx, a, ml) for ta, (xx, a, ml) in zip(t, mix)],
            args.served_load,
            "the batched form's %d instances as requests (p U[20,400], n U[200,5000], degree "
            "U[2,6], alpha {0.01,0.05}, no level cap)" % len(mix))
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(x, n, alpha, max_level, gpu_adj)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
