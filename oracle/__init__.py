"""CPU oracle of heterogpc's PC-stable skeleton path.

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package, and only as the checker (or the timed
CPU baseline).  The product (libheterogpc_hip.so and its bindings) never does.
"""
from .oracle import *  # noqa: F401,F403
