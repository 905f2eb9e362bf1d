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


@pytest.fixture(scope="session")
def mixed_small():
    """4,000 waters + 2 chains x 300 beads in a 52 A box (bead exclusions,
    one LJ type per bead): the windowed pair path with a large type table."""
    from sponge_amd import systems
    return _cached("mix13k", lambda: systems.mixed_system(
        n_waters=4000, n_chains=2, beads_per_chain=300, box=52.0, seed=11))


def gromacs_water_atoms(sysd):
    """Atoms of the fixture's TIP3P waters (O at -0.834 e followed by two H at
    +0.417 e, charges in the files' units of 18.2223): their reference forces
    hold only terms this path and the oracle's bond / angle aids compute."""
    q = np.asarray(sysd["charge"], np.float64) / 18.2223
    n = len(q)
    mask = np.zeros(n, bool)
    o = np.nonzero(np.abs(q + 0.834) < 1e-4)[0]
    for i in o:
        if i + 2 < n and abs(q[i + 1] - 0.417) < 1e-4 and abs(q[i + 2] - 0.417) < 1e-4:
            mask[i:i + 3] = True
    return mask


def force_stats(ref, got):
    """benchmarks/comparison/utils.py:122-152: max |dF|, RMS dF, mean cosine."""
    ref = np.asarray(ref, np.float64)
    got = np.asarray(got, np.float64)
    d = got - ref
    cos = np.mean(np.sum(ref * got, 1) / (np.linalg.norm(ref, axis=1) * np.linalg.norm(got, axis=1)))
    return float(np.abs(d).max()), float(np.sqrt(np.mean(d * d))), float(cos)
