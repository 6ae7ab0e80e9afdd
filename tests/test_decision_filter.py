"""The by-set search decides most tests from s01^2 against q r_crit^2
(removes_sq); this host program checks that every such decision equals the
exact statistic's, near the boundary and far from it (see
tests/decision_filter.cpp)."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "masterthesis-heterogeneous-pc_amd", "csrc")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc is not installed")
def test_squared_filter_equals_exact_decision():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "decision_filter")
        subprocess.run([HIPCC, "-x", "hip", "--offload-arch=gfx950", "-O2", "-std=c++17",
                        "-ffp-contract=off", "-I", CSRC, os.path.join(HERE, "decision_filter.cpp"),
                        "-o", exe], check=True, capture_output=True, timeout=300)
        out = subprocess.run([exe], capture_output=True, text=True, timeout=300)
        print(out.stdout)
        assert out.returncode == 0, out.stdout + out.stderr
