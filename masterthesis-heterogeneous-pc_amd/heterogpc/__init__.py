"""heterogpc on MI355X: PC-stable skeleton through libheterogpc_hip.so.

Python-side view of the C ABI (include/heterogpc_hip.h) used by the tests and
the benchmark.  The C++ command-line tool `heterogpc` drives the same ABI.
"""
from ._lib import HpcError, LevelStats, load  # noqa: F401
from .batch import Batch, Server, footprint, pack_check, plan_tasks  # noqa: F401
from .skeleton import DeviceSkeleton, r_crit, sem_generate, sepset_records  # noqa: F401
