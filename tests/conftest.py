import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
TESTS = os.path.dirname(os.path.abspath(__file__))
if TESTS not in sys.path:
    sys.path.insert(0, TESTS)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def random_strings(rng: np.random.Generator, n: int, L: int, alphabet: str) -> list[str]:
    """n random strings of length L over `alphabet` (uniform)."""
    if L == 0:
        return [""] * n
    idx = rng.integers(0, len(alphabet), size=(n, L))
    table = np.frombuffer(alphabet.encode("latin-1"), dtype=np.uint8)
    chars = table[idx]
    return [bytes(row).decode("latin-1") for row in chars]


@pytest.fixture
def rng():
    return np.random.default_rng(12345)
