"""Host-side logic of the serving layer (no GPU): the seeded request stream
and the segment systems it describes."""
import numpy as np
from scipy import stats

from sponge_amd import serving


def test_request_stream_distributions():
    """Atom counts log-uniform in [1500, 50000] (KS test on the logs), Poisson
    arrivals at 40 per second (mean gap and exponential KS test), steps as
    asked, seeded (the same stream twice)."""
    reqs = serving.make_requests(count=4000, rate=40.0, steps=1000, seed=7)
    n = np.array([r.n_atoms for r in reqs], np.float64)
    assert n.min() >= 1500 and n.max() <= 50000
    u = (np.log(n) - np.log(1500)) / (np.log(50000) - np.log(1500))
    assert stats.kstest(u, "uniform").pvalue > 1e-3
    t = np.array([r.arrival for r in reqs])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert abs(gaps.mean() - 1 / 40.0) < 5 * (1 / 40.0) / np.sqrt(len(gaps))
    assert stats.kstest(gaps * 40.0, "expon").pvalue > 1e-3
    assert all(r.steps == 1000 for r in reqs)
    again = serving.make_requests(count=4000, rate=40.0, steps=1000, seed=7)
    assert [r.n_atoms for r in again] == [r.n_atoms for r in reqs]


def test_segment_system_matches_request():
    """A request's water box holds round(N/3) waters at 29.915 A^3 each
    (1.0 g/cm^3) with the 9 A cutoff of the serving workload."""
    r = serving.Request(0, 4000, 10, 0.0, 3)
    s = r.system()
    nw = int(round(4000 / 3.0))
    assert s["n"] == 3 * nw and s["n_waters"] == nw
    assert abs(float(s["box"][0]) ** 3 - nw * serving.WATER_VOLUME) < 1e-3 * nw * 29.915
    assert s["cutoff"] == 9.0
    assert r.system() is s
    assert serving.estimate_footprint(50000) > serving.estimate_footprint(1500) > 0


def test_grid_key_matches_cell_rule():
    """GRIDS::Initial (neighbor_list.cpp:58-60): floor(L / ((rc + skin) / 2)) cells per axis,
    in float; a 27.35 A box with rc 9, skin 2 has 4 cells (27.35 / 5.5 = 4.97)."""
    from sponge_amd import autotune

    class Ctx:
        cutoff = 9.0
        box = np.array([27.35, 33.0, 55.0], np.float32)
        n = 2000

    assert autotune.grid_key(Ctx()) == (4, 6, 10)
    assert autotune.grid_key(Ctx(), occupancy=True) == (4, 6, 10, 8)
