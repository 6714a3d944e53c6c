import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")
    config.addinivalue_line("markers", "slow: long-running")


@pytest.fixture(scope="session")
def oracle():
    from cbind import load_oracle
    return load_oracle()


@pytest.fixture(scope="session")
def ref():
    from cbind import load_ref
    r = load_ref()
    if r is None:
        pytest.skip("reference build oracle/_ref not present (built only where the reference checkout exists)")
    return r


@pytest.fixture(scope="session")
def ref_linear():
    from cbind import load_ref_linear
    r = load_ref_linear()
    if r is None:
        pytest.skip("reference build oracle/_ref not present (built only where the reference checkout exists)")
    return r


@pytest.fixture(scope="session")
def golden():
    import numpy as np
    p = Path(__file__).resolve().parent / "golden" / "golden_small.npz"
    return dict(np.load(p, allow_pickle=False))
