import os
import sys

import pytest

# the oracle's dense algebra works on small matrices in these tests: one BLAS thread is
# faster than several (and keeps parallel test workers from oversubscribing the host)
try:
    from threadpoolctl import threadpool_limits
    threadpool_limits(limits=1, user_api="blas")
except ImportError:  # pragma: no cover
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
for p in (ROOT, os.path.join(ROOT, "optimode.jl_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running CPU test")
