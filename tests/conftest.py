import os
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "berkeleygw-kernels-cpp_amd" / "python"))
sys.path.insert(0, str(ROOT / "oracle"))
sys.path.insert(0, str(ROOT / "tests"))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (runs on the MI355X box)")
    config.addinivalue_line("markers", "slow: full benchmark-size parity (GPU box)")


@pytest.fixture(scope="session")
def ctx():
    import bgw

    c = bgw.Context(0)
    yield c
    c.close()


def rel_err(a, b):
    """max_k |a_k - b_k| / |b_k| over complex outputs (|b_k| > 0)."""
    a = np.asarray(a).reshape(-1)
    b = np.asarray(b).reshape(-1)
    return float(np.max(np.abs(a - b) / np.abs(b)))
