"""GPU tests of the serving layer: every window shape of the pair path gives
the same forces, energies and virials (so the tuner's choice changes time
only), and a segment's result is the same alone and in a mix of concurrent
segments."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_window_shapes_give_identical_results(water_9k):
    """Every shape the library accepts, and both hand-off modes of the window
    forces on force-only steps, give the same results bit for bit."""
    from sponge_amd import autotune
    from sponge_amd.lib import Nonbonded
    nb = Nonbonded(water_9k)
    ref = None
    tried = []
    for shp in autotune.SHAPES:
        if not nb.set_window_shape(*shp):
            continue
        tried.append(shp)
        nb.zero_outputs()
        nb.step(step=0, need_energy=True, need_virial=True, forced_nl=True)
        nb.sync()
        got = (nb.frc.cpu().numpy().copy(), nb.atom_energy.cpu().numpy().copy(),
               nb.atom_virial.cpu().numpy().copy(), nb.energies())
        # force-only steps in both hand-off modes of the window forces
        for mode in (0, 1):
            nb.set_direct_j(mode)
            nb.zero_outputs()
            nb.step(step=1)
            nb.sync()
            got += (nb.frc.cpu().numpy().copy(),)
        nb.set_direct_j(-1)
        print("shape", shp, nb.window_info(), flush=True)
        np.testing.assert_array_equal(got[4], got[5])
        if ref is None:
            ref = got
            continue
        np.testing.assert_array_equal(got[0], ref[0])
        np.testing.assert_array_equal(got[1], ref[1])
        np.testing.assert_array_equal(got[2], ref[2])
        assert got[3] == ref[3]
        np.testing.assert_array_equal(got[4], ref[4])
    assert (0, 0, 0) in tried and len(tried) >= 5
    assert nb.status() == 0
    nb.close()


@pytest.mark.parametrize("launch", ["streams", "batched"])
def test_segment_result_same_alone_and_in_mix(launch):
    """Three requests of 1,500 / 2,400 / 3,600 atoms, 30 steps each: served
    one at a time on the per-segment path and all three at once (per-segment
    streams, or one step batch), the final coordinates and velocities agree
    bit for bit."""
    from sponge_amd import serving

    def requests():
        return [serving.Request(k, n, 30, 0.0, 900 + k) for k, n in enumerate((1500, 2400, 3600))]

    alone = requests()
    out1 = serving.Server(slots=1, launch="streams").run(alone, realtime=False)
    mixed = requests()
    out3 = serving.Server(slots=3, launch=launch).run(mixed, realtime=False)
    print(out1, out3, flush=True)
    for a, b in zip(alone, mixed):
        assert a.result["status"] == 0 and b.result["status"] == 0
        np.testing.assert_array_equal(a.result["crd"], b.result["crd"])
        np.testing.assert_array_equal(a.result["vel"], b.result["vel"])
        assert np.isfinite(a.result["crd"]).all()
    assert out3["requests"] == 3 and out3["p99_latency_s"] >= out3["p50_latency_s"] > 0


def test_batched_server_staggered_admissions():
    """Requests of different lengths (20-60 steps) arriving while others run:
    the step batch is re-planned as segments join and leave, and each result
    equals the segment run alone."""
    from sponge_amd import serving

    def requests():
        return [serving.Request(k, n, st, 0.0, 700 + k)
                for k, (n, st) in enumerate(((1600, 60), (5200, 20), (2600, 40), (7000, 30),
                                             (3300, 50)))]

    alone = requests()
    serving.Server(slots=1, launch="streams").run(alone, realtime=False)
    mixed = requests()
    out = serving.Server(slots=3, launch="batched").run(mixed, realtime=False)
    print(out, flush=True)
    assert out["plans"] >= 3
    for a, b in zip(alone, mixed):
        print("case request", b.rid, b.n_atoms, b.steps, flush=True)
        assert b.result["status"] == 0
        np.testing.assert_array_equal(a.result["crd"], b.result["crd"])
        np.testing.assert_array_equal(a.result["vel"], b.result["vel"])
