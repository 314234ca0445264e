import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long-running")


def load_golden(name):
    """A golden fixture written by tests/golden/make_golden.py (plain arrays only)."""
    return dict(np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False))


def golden_names():
    return sorted(f[:-4] for f in os.listdir(GOLDEN) if f.endswith(".npz"))


def bits_equal(a, b) -> bool:
    a = np.ascontiguousarray(a, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    return a.shape == b.shape and np.array_equal(a.view(np.uint64), b.view(np.uint64))


def momentum_scale(m, vel):
    """Normaliser for momentum comparisons: sum_i m_i |v_i| (total p is ~0 by construction)."""
    return float(np.sum(m * np.linalg.norm(vel, axis=1)))


def angmom_scale(m, moi, pos, vel, w):
    return float(np.sum(m * np.linalg.norm(np.cross(pos, vel), axis=1) + moi * np.linalg.norm(w, axis=1)))


@pytest.fixture(scope="session")
def oracle_mod():
    from oracle import oracle
    if not os.path.exists(oracle.LIB_PATH):
        oracle.build()
    return oracle
