"""Shared pytest configuration.

Markers (mirroring pycc's ``setup.cfg:32-36``): ``gpu`` needs an MI355X and the built HIP library;
``slow`` is opt-in. CPU tests cover the oracle (against the reference's known answers and exact
brute-force results), the host logic, and that the C-ABI library loads and exports its symbols.
"""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running; run explicitly")


def random_amplitudes(no, nv, seed, scale=0.05):
    """(t1, t2) with t2 symmetric under the pair swap (ij)(ab), like CCSD amplitudes."""
    rng = np.random.default_rng(seed)
    t1 = rng.standard_normal((no, nv)) * scale
    t2 = rng.standard_normal((no, no, nv, nv)) * scale
    return t1, t2 + t2.transpose(1, 0, 3, 2)


@pytest.fixture
def amplitudes():
    return random_amplitudes


def reference_constant(system, quantity):
    """(value, tolerance) of a known answer held by the reference's tests (tests/golden)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_constants.json")) as fh:
        d = json.load(fh)[system][quantity]
    return d["value"], d["tolerance"]
