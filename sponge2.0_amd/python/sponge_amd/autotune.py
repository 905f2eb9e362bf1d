"""Tile-shape tuner for the windowed pair path.

The pair kernel's workgroups own blocks of a x b x s neighbour-list cells and
keep the force accumulators of the block's window in LDS.  Which shape is
fastest depends on the box (cell counts), the density and the device: large
blocks cut the window/stage overhead, small ones give more workgroups to fill
the chip.  The library picks a shape from a fixed preference list; this tuner
times every shape the library accepts for a context on that context's own
workload (a full neighbour-list period of force steps at fixed coordinates)
and keeps the fastest.  For each windowed shape it also times both ways the
blocks hand their window forces to the accumulators on force-only steps
(a staged per-atom gather, or atomic adds per block: the first is cheaper
where the stage stays in the last-level cache and the gather's launch is
hidden, the second removes a launch from small systems' dependent chain).
Every choice gives the same forces, energies and virials (integer
fixed-point sums), so it affects time only.
"""
import numpy as np

# (a, b, s) cells per block along x, y, z; (0, 0, 0) is the plain path
SHAPES = [(0, 0, 0), (1, 1, 1), (1, 1, 2), (1, 1, 4), (1, 2, 2), (1, 2, 4), (2, 2, 2),
          (2, 2, 3), (2, 2, 4), (2, 2, 6), (2, 2, 8), (2, 3, 4), (3, 3, 3), (3, 3, 4),
          (2, 4, 4), (4, 4, 4)]

_CACHE = {}


def grid_key(nb, skin=2.0, occupancy=False):
    """Cell counts of a context's neighbour-list grid (GRIDS::Initial,
    neighbor_list.cpp:58-60, in float), with the mean atoms per cell when
    systems of different densities share a cache."""
    f = np.float32
    gl = (f(nb.cutoff) + f(skin)) * f(0.5)
    nx, ny, nz = (int(np.floor(f(x) / gl)) for x in nb.box)
    if occupancy:
        return nx, ny, nz, int(round(nb.n / float(nx * ny * nz)))
    return nx, ny, nz


def time_period(nb, interval, reps=2):
    """Mean ms per force step over `reps` neighbour-list periods (one rebuild
    each), replayed eagerly on the context's stream at fixed coordinates."""
    torch = nb.torch
    with torch.cuda.stream(nb.stream):
        for k in range(interval):                 # warm: build for this shape
            nb.step(step=k, forced_nl=(k == 0))
        ev0 = torch.cuda.Event(enable_timing=True)
        ev1 = torch.cuda.Event(enable_timing=True)
        ev0.record(nb.stream)
        for r in range(reps):
            for k in range(interval):
                nb.step(step=k)
        ev1.record(nb.stream)
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / (reps * interval)


def tune_window(nb, interval=10, reps=2, shapes=SHAPES, modes=(0, 1), use_cache=True):
    """Times each accepted shape (and, for windowed shapes, each hand-off mode
    of snb_lj_set_direct_j) and sets the fastest; returns (shape, {(shape,
    mode): ms per step}, mode).  Results are cached per grid key."""
    key = grid_key(nb)
    if use_cache and key in _CACHE:
        best, _, mode = _CACHE[key]
        if nb.set_window_shape(*best):
            nb.set_direct_j(mode)
            return _CACHE[key]
    times = {}
    for shp in shapes:
        if not nb.set_window_shape(*shp):
            continue
        for mode in (modes if any(shp) else (-1,)):
            nb.set_direct_j(mode)
            times[(shp, mode)] = time_period(nb, interval, reps)
    best, mode = min(times, key=times.get)
    nb.set_window_shape(*best)
    nb.set_direct_j(mode)
    nb.zero_outputs()
    nb.clear_fixed_forces()
    _CACHE[key] = (best, times, mode)
    return _CACHE[key]
