"""Test configuration: the `gpu` marker and shared system generators."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU (runs the HIP library)")
    config.addinivalue_line("markers", "slow: CPU test that takes more than a few seconds")


@pytest.fixture(scope="session")
def orc():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def sdt():
    import stokesdt_amd as S
    S.lib()
    return S


def uniform_system(n, box, seed):
    """Uniform positions in [0, box)^3 (overlaps allowed, like the
    reference's tests/data/N1000_Phi0.1.xyz), unit radii."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, box, 3 * n), np.ones(n)


def poly_system(n, box, seed):
    """Polydisperse radii drawn from 15 sizes spanning 21..115 (the spread of
    tests/data/poly.mod), uniform positions."""
    rng = np.random.default_rng(seed)
    sizes = np.geomspace(21.4, 115.2, 15)
    rdi = rng.choice(sizes, size=n)
    return rng.uniform(0.0, box, 3 * n), rdi
