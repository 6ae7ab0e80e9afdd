"""The exact correlation on the FP64 matrix cores (corr_xmfma_kernel) against
the oracle, bit for bit, across the shapes its layout depends on.

The kernel reads class-major z-scores whose classes start with
P0 = M8 - M zero steps (M = npad / 64 steps per lane class, M8 = M rounded up
to 8), so sample counts are chosen to give every P0 from 0 to 7 and very
short classes; p is ragged against the 32 x 32 pair blocks and the 8 x 8
super-blocks.  Signs of zero are compared too (an exact zero sum must come
out +0).  Instances with NaN cells read the same layout from the masked
kernel.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# npad / 64 = M for each n: P0 = (8 - M % 8) % 8
NS = [2, 3, 64, 65, 130, 200, 320, 448, 500, 512, 513, 1000, 2049]


def _same_bits(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64))


def _check(hpc, orc, x, label, mode=0):
    p, n = x.shape
    print(f"[corr] {label}: p={p} n={n} mode={mode}", flush=True)
    with hpc.DeviceSkeleton(p, n, 0.05, max_level=0) as dev:
        dev.pmcc(x, mode)
        got = dev.cor()
    want = orc.corr(x)
    if np.isnan(want).any():
        assert np.array_equal(np.isnan(got), np.isnan(want)), label
        m = ~np.isnan(want)
        assert _same_bits(got[m], want[m]), label
    else:
        assert _same_bits(got, want), f"{label}: correlation bits differ"


@pytest.mark.parametrize("n", NS)
def test_xmfma_matches_oracle_bits(hpc, orc, n):
    rng = np.random.default_rng(1000 + n)
    p = 70 if n < 2049 else 300
    x = rng.standard_normal((p, n))
    x[5] = x[3] * 0.5 + 0.01 * x[5]
    _check(hpc, orc, x, f"gauss n={n}")


def test_xmfma_integer_data_exact_zeros(hpc, orc):
    """Small-integer data: many class sums and totals are exactly zero."""
    rng = np.random.default_rng(7)
    for n in (65, 448, 1000):
        x = rng.integers(-1, 2, size=(40, n)).astype(np.float64)
        x[:, 0] = 0.0
        _check(hpc, orc, x, f"ternary n={n}")


def test_xmfma_ragged_p_and_superblocks(hpc, orc):
    rng = np.random.default_rng(3)
    for p in (2, 31, 33, 257, 290):
        x = rng.standard_normal((p, 333))
        _check(hpc, orc, x, f"ragged p={p}")


def test_masked_kernel_reads_class_major_layout(hpc, orc):
    rng = np.random.default_rng(5)
    for n in (200, 513):
        x = rng.standard_normal((50, n))
        x[3, 7] = np.nan
        x[10, ::17] = np.nan
        _check(hpc, orc, x, f"nan n={n}")


def test_fast_mode_then_exact_mode_same_state(hpc, orc):
    """A state that ran the FP32 form and then the exact form holds both
    layouts; the exact correlation is still the oracle's."""
    rng = np.random.default_rng(11)
    p, n = 90, 700
    x = rng.standard_normal((p, n))
    with hpc.DeviceSkeleton(p, n, 0.05, max_level=0) as dev:
        dev.pmcc(x, 1)
        dev.pmcc(x, 0)
        got = dev.cor()
    assert _same_bits(got, orc.corr(x))


@pytest.mark.parametrize("kind", ["f64", "f32", "mixed"])
@pytest.mark.parametrize("p,n", [(20, 1000), (131, 1001), (513, 1025), (97, 40961)])
def test_staged_upload_odd_sizes(hpc, orc, p, n, kind):
    """Observations go through the pinned staging ring in slices split over
    the host threads; sizes that leave a short last slice (and empty thread
    parts in it) still arrive whole.  FP32-exact data ("f32", with -0 and
    FP32 subnormals in it) crosses at half width and is widened on the
    device; "mixed" has one value FP32 cannot hold in its last slice, so the
    slices before it go narrowed and the rest as doubles.  Either way the
    correlation equals the oracle's bit for bit."""
    rng = np.random.default_rng(p * 7 + n)
    x = rng.standard_normal((p, n))
    if kind != "f64":
        x = x.astype(np.float32).astype(np.float64)
        x[0, 1] = -0.0
        x[1, 2] = float(np.float32(1e-40))
        x[2, 3] = float(np.float32(-3e-42))
    if kind == "mixed":
        x[-1, -1] = 1.0 + 2.0 ** -40
    _check(hpc, orc, x, f"staged upload {kind} p={p} n={n}")
