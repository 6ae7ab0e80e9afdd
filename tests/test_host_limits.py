"""Host-side limits of the library, without a GPU: the size check of a
batch or server step (int32 neighbour indexing), the search plan of a row at
the edge of 64-bit binomials, and the deadline of the host wait for device
counts.  Exact integers (Python) pin the library's saturating arithmetic."""
import ctypes as C
import math
import threading
import time

import pytest

HPC_EINVAL, HPC_EHIP, HPC_ENOMEM, HPC_ELEVEL = -1, -2, -3, -5
SAT = 2 ** 64 - 1
MAX_ROW_TASKS = 2 ** 46


def test_pack_check_int32_edges(hpc):
    assert hpc.pack_check([46340]) == 0                      # 46340 * 46339 < 2^31
    assert hpc.pack_check([46341]) == HPC_EINVAL
    assert hpc.pack_check([1]) == HPC_EINVAL
    assert hpc.pack_check([30000, 30000]) == 0               # 1.80e9
    assert hpc.pack_check([30000, 30000, 30000]) == HPC_ENOMEM  # 2.70e9 > 2^31 - 1
    # exactly at the limit: p (p - 1) sums to 2^31 - 1 or one past it
    ps = [46340, 367]                                        # 2147349260 + 134322
    total = sum(p * (p - 1) for p in ps)
    assert total <= 2 ** 31 - 1 and hpc.pack_check(ps) == 0
    ps2 = [46340, 368]
    assert sum(p * (p - 1) for p in ps2) > 2 ** 31 - 1 and hpc.pack_check(ps2) == HPC_ENOMEM
    assert hpc.pack_check([2000] * 500) == 0                 # 2.0e9
    assert hpc.pack_check([2000] * 540) == HPC_ENOMEM        # 2.16e9


def _expected_tasks(deg, level, rpt):
    c = math.comb(deg - 1, level)
    if c == 0:
        return 0
    if c >= SAT:  # the table saturates
        return None
    t = deg * -(-c // rpt)
    return None if t > MAX_ROW_TASKS else t


@pytest.mark.parametrize("rpt", [64, 512, 4096])
def test_row_plan_matches_exact_integers(hpc, rpt):
    degs = list(range(2, 130)) + [150, 199, 200, 300, 500, 1000, 5000, 20000, 46340]
    flips = 0
    for level in range(2, 17):
        prev = True
        for deg in degs:
            want = _expected_tasks(deg, level, rpt)
            if want is None:
                with pytest.raises(hpc.HpcError):
                    hpc.plan_tasks(deg, level, rpt)
            else:
                assert hpc.plan_tasks(deg, level, rpt) == want, (deg, level, rpt)
            ok = want is not None
            flips += prev and not ok
            prev = ok
    assert flips >= 10  # the boundary is crossed at most levels


def test_row_plan_boundary_is_exact(hpc):
    """Around each level's first unplannable row length, both sides."""
    rpt = 512
    for level in range(2, 17):
        deg = level + 1
        while _expected_tasks(deg, level, rpt) is not None and deg < 46340:
            deg += 1
        if deg >= 46340:
            continue
        assert hpc.plan_tasks(deg - 1, level, rpt) == _expected_tasks(deg - 1, level, rpt)
        with pytest.raises(hpc.HpcError):
            hpc.plan_tasks(deg, level, rpt)
    # saturation itself: C(m, 16) first exceeds 2^64 - 1 near m = 270
    m = 16
    while math.comb(m, 16) < SAT:
        m += 1
    with pytest.raises(hpc.HpcError):
        hpc.plan_tasks(m + 1, 16, 64)


def test_wait_word_deadline(hpc):
    lib = hpc.load()
    word = C.c_uint64(0)
    t0 = time.perf_counter()
    rc = lib.hpc_wait_word(C.byref(word), 7, 0.05)
    dt = time.perf_counter() - t0
    assert rc == HPC_EHIP and b"did not arrive within" in lib.hpc_last_error()
    assert 0.04 < dt < 5.0
    # a word that arrives in time ends the wait with success
    def later():
        time.sleep(0.05)
        word.value = 7
    th = threading.Thread(target=later)
    th.start()
    assert lib.hpc_wait_word(C.byref(word), 7, 10.0) == 0
    th.join()
    # already there: no wait; no limit (0) with the word present
    assert lib.hpc_wait_word(C.byref(word), 7, 0.0) == 0
    assert lib.hpc_wait_word(None, 7, 1.0) == HPC_EINVAL
