"""hpc_corr_noalloc (the correlation without allocation) against hpc_corr.

Bar: bit-identical correlation matrices (both triangles) on two calls in a
row on one device, the second with a smaller input, plus an input with NaN
cells (the masked kernel) after them; a shape beyond the static scratch is
refused with an error before anything runs.  The test's device buffers come
from the HIP runtime the library itself links (plumbing only: torch ships a
runtime of its own, whose pointers this one cannot use); the entry point
allocates nothing."""
import ctypes as C
import os

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libheterogpc_hip.so
    links (/opt/rocm/lib)."""

    def __init__(self):
        path = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so.7")
        self.lib = C.CDLL(path)
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.lib.hipSetDevice.argtypes = [C.c_int]
        assert self.lib.hipSetDevice(0) == 0

    def alloc(self, nbytes):
        ptr = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(ptr), C.c_size_t(nbytes)) == 0
        return ptr

    def upload(self, a):
        a = np.ascontiguousarray(a)
        ptr = self.alloc(a.nbytes)
        assert self.lib.hipMemcpy(ptr, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return ptr

    def download(self, ptr, shape, dtype=np.float64):
        out = np.empty(shape, dtype=dtype)
        assert self.lib.hipMemcpy(out.ctypes.data, ptr, out.nbytes, 2) == 0  # device to host
        return out

    def free(self, ptr):
        assert self.lib.hipFree(ptr) == 0

    def sync(self):
        assert self.lib.hipDeviceSynchronize() == 0


def _reference(hpc, x):
    p, n = x.shape
    with hpc.DeviceSkeleton(p, n, 0.05, max_level=1) as dev:
        dev.pmcc(x)
        return dev.cor()


def test_noalloc_equals_hpc_corr_twice_then_nan(hpc, orc):
    lib = hpc.load()
    hip = Hip()
    x1, _, _, _ = workloads.sem("c2")                                     # p = 1000, n = 5000
    x2 = np.ascontiguousarray(workloads.sem("small")[0][:257, :1500])    # smaller p and n
    xs, _, _, _ = workloads.sem("small")
    x3 = np.ascontiguousarray(xs[:120].copy())
    rng = np.random.default_rng(3)
    x3[rng.random(x3.shape) < 0.01] = np.nan
    x3[5] = 1.25
    out = hip.alloc(1000 * 1000 * 8)
    try:
        for label, x in (("c2", x1), ("smaller", x2), ("nan", x3)):
            p, n = x.shape
            print(f"[noalloc] {label}: p={p} n={n}", flush=True)
            dx = hip.upload(x)
            rc = lib.hpc_corr_noalloc(0, dx, p, n, out, None)
            assert rc == 0, lib.hpc_last_error()
            hip.sync()
            got = hip.download(out, (p, p))
            hip.free(dx)
            want = _reference(hpc, x)
            assert np.array_equal(got, want, equal_nan=True), (
                f"{label}: max |d| = {np.nanmax(np.abs(got - want))}")
            assert np.array_equal(want, orc.corr(x), equal_nan=True)
    finally:
        hip.free(out)


def test_noalloc_refuses_shapes_beyond_scratch(hpc):
    lib = hpc.load()
    hip = Hip()
    buf = hip.alloc(64)
    try:
        # p = 2000, n = 5000 needs about 91 MB of scratch: refused before any launch
        assert lib.hpc_corr_noalloc(0, buf, 2000, 5000, buf, None) == -1
        assert b"scratch" in lib.hpc_last_error()
        assert lib.hpc_corr_noalloc(0, None, 10, 10, buf, None) == -1
        assert lib.hpc_corr_noalloc(0, buf, 1, 10, buf, None) == -1
        assert lib.hpc_corr_noalloc(99, buf, 10, 10, buf, None) == -1
    finally:
        hip.free(buf)
