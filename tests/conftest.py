import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgdt.so on the GPU)")
    config.addinivalue_line("markers", "slow: long CPU-only oracle checks")


@pytest.fixture(scope="session")
def ref_grid():
    """The reference's grid: 16^3 on [-5,5]^3, R = 16.5 (no node cut), determ.cu:47-53."""
    import oracle as O
    return O.Grid((16, 16, 16), (-5, -5, -5), (5, 5, 5), 16.5)
