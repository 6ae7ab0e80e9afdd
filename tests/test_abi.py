"""The C ABI library loads and exports every entry point its header declares;
host-only entry points behave without a GPU.  No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = [os.path.join(REPO, "include", f) for f in os.listdir(os.path.join(REPO, "include"))
           if f.endswith(".h")]


def declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(hpc_\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("hpc_state_create", "hpc_pmcc", "hpc_compact", "hpc_test_level", "hpc_level",
                 "hpc_run", "hpc_get_sepsets", "hpc_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol(hpc):
    lib = hpc.load()
    missing = [n for n in sorted(declared_functions()) if not hasattr(lib, n)]
    assert not missing, missing


def test_bindings_cover_every_declared_symbol():
    from heterogpc import _lib

    assert declared_functions() <= set(_lib.SIGNATURES)


def test_r_crit_host_equals_oracle_boundary(hpc, orc):
    for n, a in [(1000, 0.01), (5000, 0.05), (2000, 0.01), (10000, 0.05), (3, 0.05), (2, 0.05)]:
        assert hpc.r_crit(n, a) == orc.r_crit(n, a)


def test_sem_generator_is_deterministic_and_shaped(hpc):
    x1, dag = hpc.sem_generate(300, 2000, 3.0, seed=5, fp32=True, with_dag=True)
    x2 = hpc.sem_generate(300, 2000, 3.0, seed=5, fp32=True)
    assert np.array_equal(x1, x2)
    assert np.array_equal(x1.astype(np.float32).astype(np.float64), x1)  # FP32-rounded
    assert np.array_equal(dag, dag.T) and not dag.diagonal().any()
    edges = dag.sum() // 2
    assert abs(edges - 300 * 3 / 2) < 4 * np.sqrt(450)  # Binomial(C(300,2), 3/299)
    # a root node (no parents) is pure N(0,1) noise
    roots = [j for j in range(300) if not dag[:j, j].any()]
    z = x1[roots[0]]
    assert abs(z.mean()) < 0.1 and abs(z.std() - 1) < 0.1
    x3 = hpc.sem_generate(300, 2000, 3.0, seed=6, fp32=True)
    assert not np.array_equal(x1, x3)


def test_errors_are_reported_not_fatal(hpc):
    lib = hpc.load()
    h = C.c_void_p()
    rc = lib.hpc_state_create(1, 10, 0.05, 4, 0, C.byref(h))
    assert rc != 0 and b"outside" in lib.hpc_last_error()
    rc = lib.hpc_sem_generate(1, 1, 3.0, 0.1, 1.0, 1, 1, None, None)
    assert rc != 0
    with pytest.raises(hpc.HpcError):
        hpc.sem_generate(1, 5, 2.0)


def test_footprint_is_host_only_and_monotone(hpc):
    b = hpc.footprint(2000, 5000, 4)
    # correlation + level-1 sepset records + observations are a lower bound
    assert b >= 2000 * 2000 * (8 + 8) + 2000 * 5000 * 8
    assert hpc.footprint(2001, 5000, 4) > b
    assert hpc.footprint(2000, 5001, 4) > b
    # sepsets are records of removed edges: no per-level planes
    assert hpc.footprint(2000, 5000, 8) == b
    with pytest.raises(hpc.HpcError):
        hpc.footprint(1, 10, 4)
