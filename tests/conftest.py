import os
import sys

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
for p in (os.path.join(ROOT, "sponge2.0_amd", "python"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU and the built HIP library")
    config.addinivalue_line("markers", "slow: long-running (full-size) case")


_cache = {}


def _cached(key, fn):
    if key not in _cache:
        _cache[key] = fn()
    return _cache[key]


@pytest.fixture(scope="session")
def gromacs_iter0():
    from sponge_amd import native_io
    return _cached("gmx0", lambda: native_io.system_from_prefix(
        os.path.join(GOLDEN, "gromacs_charmm_tip3p", "iter0"), "sys_flexible"))


@pytest.fixture(scope="session")
def tip3p_1011():
    from sponge_amd import native_io
    return _cached("tip1011", lambda: native_io.system_from_prefix(
        os.path.join(GOLDEN, "tip3p_water_1011"), "tip3p"))


@pytest.fixture(scope="session")
def water_9k():
    from sponge_amd import systems
    return _cached("w9k", lambda: systems.config_cpu_9k())


@pytest.fixture(scope="session")
def small_water():
    from sponge_amd import systems
    return _cached("w648", lambda: systems.tip3p_box(216, 18.7, seed=7, cutoff=8.0))


def gromacs_case(it):
    from sponge_amd import native_io
    return _cached("gmx%d" % it, lambda: native_io.system_from_prefix(
        os.path.join(GOLDEN, "gromacs_charmm_tip3p", "iter%d" % it), "sys_flexible"))


def rel_rms(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))
