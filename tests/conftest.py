import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def gpu():
    """A GpuUnbinder factory on cuda:0 (fails loudly without the HIP library)."""
    from hbtplus_amd.device import GpuUnbinder
    made = []

    def make(params):
        u = GpuUnbinder(params, device=0)
        made.append(u)
        return u

    yield make
    for u in made:
        u.close()
