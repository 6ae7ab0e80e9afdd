"""Benchmark of the PC-stable skeleton hot path on one MI355X.

A step is one full skeleton of the plain p=2000 workload on device-resident
observations: correlation (moments + pairwise sums), then levels 0, 1, ...
(compaction, CI tests, removal) until no edge has enough neighbours.

value = CI tests/s = sum over levels of the canonical test count T_l divided
by the level wall time (the reference's "Execution duration" scope,
skeleton.cpp:167-180).  Prints one JSON line.
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


def cpu_baseline(x, n, alpha, max_level, tests_total):
    """The oracle's reference-faithful CPU path (per-test pseudoinverse,
    immediate deletion with early exit, OpenMP over rows sorted by length)
    on all host threads, on the same workload."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()
    t0 = time.perf_counter()
    cor = oracle.corr(x, threads=threads)
    t1 = time.perf_counter()
    sk = oracle.Skeleton(cor, n, alpha, max_level=max_level, mode=oracle.MODE_PINV,
                         flags=oracle.FLAG_REF_EARLY_EXIT, threads=threads)
    lv = sk.run(max_level)
    t2 = time.perf_counter()
    return {
        "value": tests_total / (t2 - t1),
        "unit": "CI tests/s",
        "cores": threads,
        "kind": "port",
        "sample": "the whole p=2000 skeleton once (correlation %.1f s, levels %.2f s, %d levels)"
                  % (t1 - t0, t2 - t1, len(lv)),
        "skeleton_ms_levels": (t2 - t1) * 1e3,
        "skeleton_ms_e2e": (t2 - t0) * 1e3,
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
    in one batch; skeletons/hour of both and whether every result agrees."""
    for x, a, ml in insts[:4]:
        _plain_skeleton(heterogpc, x, a, ml)
    with heterogpc.Batch() as b:
        for x, a, ml in insts[:8]:
            b.add(x, a, ml)
        b.run()
    heterogpc.load().hpc_device_synchronize(0)
    t0 = time.perf_counter()
    plain = [_plain_skeleton(heterogpc, x, a, ml) for x, a, ml in insts]
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    with heterogpc.Batch() as b:
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


def bench_served(heterogpc, trace, load):
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
        "workload": "%d requests, p log-uniform [50,2000], n U[500,10000], degree U[2,5], alpha "
                    "{0.01,0.05}, FP32 data, max level 4; Poisson arrivals at %.1f/s = %.2f x the "
                    "plain path's capacity" % (n_req, rate, load),
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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(every instance runs on a single MI355X).", file=sys.stderr)
        return 2

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
    test_bytes = sum(lv["tests"] * bytes_per_test(lv["level"]) for lv in last)
    ci_gbs = test_bytes / (k_test * 1e-3) / 1e9 if k_test > 0 else 0.0
    corr_flops = 2.0 * n * p * (p + 1) / 2.0  # distinct pairs x 2n
    corr_tflops = corr_flops / (k_corr * 1e-3) / 1e12 if k_corr > 0 else 0.0

    # The metric's timed region is the level loop, so the roofline object is
    # the CI-test kernels' (levels 0..4 together: algorithmic bytes per step
    # over their summed HIP-event time on the engine stream); the correlation
    # has its own entry.  HBM traffic per step comes from the committed
    # rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload.
    traffic = _profiled_traffic()
    roof = {"kernel": "CI-test kernels, levels 0..4 (level0_kernel, search_l1_lds_kernel, "
                      "search_ln_lds_kernel<2..4>)",
            "bound": "hbm", "achieved": ci_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ci_gbs / HBM_PEAK_GBS,
            "traffic": traffic.get("ci_tests_traffic_bytes_per_step"),
            "algorithmic_bytes_per_step": test_bytes,
            "traffic_source": traffic.get("source")}
    if roof["traffic"] and k_test > 0:
        # what the kernels actually move through HBM, at the measured time
        rate = roof["traffic"] / (k_test * 1e-3) / 1e9
        roof["hbm_traffic_GBs"] = rate
        roof["hbm_traffic_frac"] = rate / HBM_PEAK_GBS
        roof["note"] = ("achieved counts algorithmic bytes; the level >= 2 kernels read their "
                        "operands from an LDS-staged row block, so HBM traffic is %.0f%% of them "
                        "and frac can exceed 1; those kernels are bound by FP64 VALU issue"
                        % (100.0 * roof["traffic"] / max(test_bytes, 1)))
    corr_roof = {"kernel": "corr_tile_kernel (fp64, canonical summation order)", "bound": "fp64",
                 "achieved": corr_tflops, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": corr_tflops / FP64_PEAK_TFLOPS,
                 "traffic": traffic.get("corr_traffic_bytes"),
                 "algorithmic": "%.3e flops per launch (n * p * (p+1))" % corr_flops}

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
        "tests_total": tests_total,
        "levels": [{k: lv[k] for k in ("level", "max_deg", "tests", "tests_max", "edges",
                                       "removed", "edges_left", "ms_compact", "ms_test",
                                       "ms_commit", "ms_wall")} for lv in last],
        "kernel_ms": {"moments": k_mom, "corr": k_corr, "ci_tests": k_test,
                      "compact": k_compact, "commit": k_commit},
        "ci_tests_hbm": {"achieved_GBs": ci_gbs, "frac": ci_gbs / HBM_PEAK_GBS,
                         "algorithmic_bytes": test_bytes},
        "corr_tflops": corr_tflops,
        "autotune": tuned,
        "roofline": roof,
        "corr_roofline": corr_roof,
    }
    dev.close()
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
        out["served"] = bench_served(heterogpc, workloads.served_trace(args.served_count, 1.0),
                                     args.served_load)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(x, n, alpha, max_level, tests_total)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
