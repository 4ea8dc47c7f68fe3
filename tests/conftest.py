import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C ABI)")


def normwise(a, b):
    """(max|a-b|/max|b|, ||a-b||_F/||b||_F) -- the error metrics we report."""
    import numpy as np

    d = np.abs(a - b)
    return float(d.max() / np.abs(b).max()), float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="session")
def ctx():
    """One GPU context for the whole GPU session (never created on CPU runs)."""
    from project1_amd import spexhip

    c = spexhip.Context(0)
    yield c
    c.close()
