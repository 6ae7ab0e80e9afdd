"""Build recipes for libheterogpc_hip.so (gfx950) and the CPU oracle.

Both land in-tree so that a copy of the repository carries them:
  masterthesis-heterogeneous-pc_amd/libheterogpc_hip.so
  oracle/liboracle.so
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG_DIR)
CSRC = os.path.join(PKG_DIR, "csrc")
BUILD = os.path.join(PKG_DIR, "_build")
LIB = os.path.join(PKG_DIR, "libheterogpc_hip.so")
ORACLE_DIR = os.path.join(REPO, "oracle")
HOST_DIR = os.path.join(PKG_DIR, "host")
CLI = os.path.join(PKG_DIR, "bin", "heterogpc")
HOST_SOURCES = ["state.cpp", "io.cpp", "executors.cpp", "balancer.cpp", "skeleton.cpp", "main.cpp"]

SOURCES = ["abi.hip", "engine.hip", "corr.hip", "graph.hip", "search.hip", "hostshare.hip",
           "batch.hip", "recheck.hip", "noalloc.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # every fused multiply-add is an explicit fma: bit-identical to the oracle
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

# per-source flags: the matrix-core kernels keep their accumulators in
# ordinary VGPRs (the default AGPR form made the compiler copy them between
# register files around every step of the exact correlation's loop)
FILE_FLAGS = {"corr.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _run(cmd: list[str], cwd: str | None = None) -> None:
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE,
                          stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))


def build_hip(jobs: int = 4) -> str:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        cmds.append([HIPCC, *HIP_FLAGS, *FILE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src),
                     "-o", obj])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_cli() -> str:
    """The heterogpc command line: C++ host code above the C ABI."""
    os.makedirs(os.path.dirname(CLI), exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    _run([cxx, "-O2", "-std=c++17", "-Wall", "-Wno-unused-variable", "-pthread",
          *[os.path.join(HOST_DIR, f) for f in HOST_SOURCES],
          "-L" + PKG_DIR, "-lheterogpc_hip", "-Wl,-rpath,$ORIGIN/..", "-o", CLI])
    return CLI


def build_oracle() -> str:
    make = shutil.which("make")
    if make:
        _run([make, "-s", "-B", "-C", ORACLE_DIR])
    else:
        _run(["gcc", "-O3", "-std=gnu11", "-fPIC", "-fopenmp", "-ffp-contract=off",
              "-mavx2", "-mfma", "-shared", "-o", "liboracle.so", "oracle.c", "-lm"],
             cwd=ORACLE_DIR)
    return os.path.join(ORACLE_DIR, "liboracle.so")


def build_all() -> None:
    build_oracle()
    build_hip()
    build_cli()


if __name__ == "__main__":
    build_all()
    print(LIB)
