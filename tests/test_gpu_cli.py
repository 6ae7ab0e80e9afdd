"""The heterogpc command line (C++ host code above the C ABI) on the GPU:
every execution mode gives the oracle's skeleton, printed in the reference's
formats (sepset_util.hpp:9-60, skeleton.cpp:72-102)."""
import os
import re
import subprocess

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(REPO, "masterthesis-heterogeneous-pc_amd", "bin", "heterogpc")
SEP_RE = re.compile(r"^Separation from (\d+) to (\d+)(?: via ((?:\d+ )+))?$")


def run_cli(args, timeout=120):
    print("[cli]", " ".join(args), flush=True)
    proc = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=timeout)
    assert proc.returncode == 0, proc.stdout + proc.stderr
    return proc.stdout


def parse_output(out, p):
    adj = np.ones((p, p), dtype=np.int32) - np.eye(p, dtype=np.int32)
    seps = {}
    total = None
    for line in out.splitlines():
        m = SEP_RE.match(line)
        if m:
            i, j = int(m.group(1)), int(m.group(2))
            adj[i, j] = adj[j, i] = 0
            seps[(i, j)] = [int(v) for v in m.group(3).split()] if m.group(3) else None
        if line.startswith("Total number of edges: "):
            total = int(line.split(": ")[1])
    return adj, seps, total


def oracle_skeleton(orc, x, alpha, levels):
    cor = orc.corr(x)
    sk = orc.Skeleton(cor, x.shape[1], alpha, max_level=max(levels, 1))
    for level in range(levels + 1):
        sk.level(level)
    return cor, sk


def check(orc, out, x, alpha, tested_levels):
    p, n = x.shape
    cor, sk = oracle_skeleton(orc, x, alpha, tested_levels)
    adj, seps, total = parse_output(out, p)
    want = sk.adj()
    # every removed pair is printed with a sepset; the remaining edges match
    assert np.array_equal(adj, want)
    assert total == int(np.triu(want, 1).sum())
    for (i, j), S in seps.items():
        if S is None:
            assert orc.pval(cor[j, i], n) >= alpha
        else:
            assert max(orc.pval_ids(cor, n, i, j, S), orc.pval_ids(cor, n, j, i, S)) >= alpha


@pytest.fixture(scope="module")
def sem_csv(tmp_path_factory):
    x, n, alpha, _ = workloads.sem("small")
    x = x[:120]
    path = tmp_path_factory.mktemp("cli") / "sem.csv"
    np.savetxt(path, x.T, delimiter=",", fmt="%.17g")
    return str(path), np.loadtxt(path, delimiter=",").T.copy()


def test_cooling_house_default_mode(orc):
    path = os.path.join(REPO, "tests", "golden", "cooling_house.csv")
    out = run_cli(["-i", path, "-v", "-p"])
    x = workloads.cooling_house()
    check(orc, out, x, 0.05, 3)  # levels >= 4 are not tested by the row executors


@pytest.mark.parametrize("mode,levels", [(["--gpu-only"], 3), ([], 3), (["--cpu-only"], 3),
                                         (["-w"], 4), (["--all-levels"], 4),
                                         (["--gpu-only", "--all-levels", "-m", "6"], 6)])
def test_sem_modes(orc, sem_csv, mode, levels):
    path, x = sem_csv
    out = run_cli(["-i", path, "-a", "0.05", "-v", "-p", "-t", "4", *mode])
    check(orc, out, x, 0.05, levels)


def test_correlation_matrix_input(orc, sem_csv, tmp_path):
    path, x = sem_csv
    cor = orc.corr(x)
    cpath = tmp_path / "cor.csv"
    np.savetxt(cpath, cor, delimiter=",", fmt="%.17g")
    out = run_cli(["-i", str(cpath), "--corr", "-o", str(x.shape[1]), "-v", "-p"])
    adj, seps, total = parse_output(out, x.shape[0])
    _, sk = oracle_skeleton(orc, x, 0.05, 3)
    assert np.array_equal(adj, sk.adj())


def test_metrics_csv(sem_csv, tmp_path):
    path, x = sem_csv
    mpath = tmp_path / "metrics.csv"
    run_cli(["-i", path, "--csv-export", str(mpath)])
    run_cli(["-i", path, "--csv-export", str(mpath), "--gpu-only"])
    rows = open(mpath).read().strip().splitlines()
    assert len(rows) == 2
    for row in rows:
        cells = row.split(",")
        # numGPUs, threads, edges, 5 levels x (level, balance, cpu, gpu), total
        assert len(cells) == 3 + 5 * 4 + 1
        assert cells[0] == "1"
        assert all(c.strip().lstrip("-").isdigit() for c in cells)
