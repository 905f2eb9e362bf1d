"""GPU tests of step batches (snb_batch_*): the launches of several
contexts' steps recorded, merged into one dispatch per kernel and issued on
one stream (eagerly and from a captured HIP graph) give every context the
same coordinates and velocities, bit for bit, as its own md_step calls."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CHUNK = 10


def _member(n, seed, overlap=False):
    import torch
    from sponge_amd import autotune, serving
    from sponge_amd.lib import Nonbonded
    s = serving.Request(0, n, 4 * CHUNK, 0.0, seed).system()
    st = torch.cuda.Stream()
    nb = Nonbonded(s, cutoff=s["cutoff"], refresh_interval=CHUNK, stream=st,
                   max_atom_in_grid_numbers=serving.MAX_ATOMS_IN_CELL)
    nb.set_pme_overlap(overlap)
    nb.set_export(False)
    autotune.tune_window(nb, interval=CHUNK)
    with torch.cuda.stream(st):
        nb.md_initial(s, temperature=300.0, tau=0.1, seed=seed, dt=0.002)
        for k in range(CHUNK):
            nb.md_step(step=k, forced_nl=(k == 0))
    st.synchronize()
    return nb, st


# plain pair path (1,600 atoms), windowed paths of several shapes; the largest
# grid takes the unfused inverse transform
SIZES = (1600, 5200, 12000, 30000)


def test_step_batch_matches_each_system_alone():
    import torch
    from sponge_amd.lib import StepBatch
    ref = []
    for i, n in enumerate(SIZES):
        nb, st = _member(n, 300 + i)
        with torch.cuda.stream(st):
            for k in range(CHUNK, 4 * CHUNK):
                nb.md_step(step=k)
        st.synchronize()
        ref.append((nb.crd.cpu().numpy(), nb.vel.cpu().numpy()))
        print("alone", n, nb.window_info(), flush=True)
        nb.close()
    mem = [_member(n, 300 + i) for i, n in enumerate(SIZES)]
    b = StepBatch()
    for nb, st in mem:
        b.record(nb, CHUNK, CHUNK)
    launches, batched = b.plan()
    info = b.info()
    print("plan", info, flush=True)
    assert info["members"] == len(SIZES) and batched <= launches
    assert info["merged"] > launches      # launches of several members share dispatches
    bs = torch.cuda.Stream()
    b.launch(bs)                          # period 2, eager
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(bs):
        g.capture_begin()
        b.launch(bs)
        g.capture_end()
    for _ in range(2):                    # periods 3 and 4, graph replays
        with torch.cuda.stream(bs):
            g.replay()
        bs.synchronize()                  # eager host work between the replays
    for (nb, st), (c, v) in zip(mem, ref):
        print("case", nb.n, flush=True)
        assert nb.status() == 0
        np.testing.assert_array_equal(nb.crd.cpu().numpy(), c)
        np.testing.assert_array_equal(nb.vel.cpu().numpy(), v)
    g.reset()
    b.close()
    for nb, st in mem:
        nb.close()


def test_step_batch_rejects_second_stream_and_double_record():
    from sponge_amd.lib import SnbError, StepBatch
    nb, st = _member(1600, 11, overlap=True)
    b = StepBatch()
    with pytest.raises(SnbError) as e:
        b.record(nb, CHUNK, CHUNK)
    assert e.value.code == 4   # SNB_ERR_UNSUPPORTED: the reciprocal stream is on
    nb.set_pme_overlap(False)
    b.record(nb, CHUNK, CHUNK)
    with pytest.raises(SnbError):
        b.record(nb, CHUNK, CHUNK)   # a context is a member once
    b.plan()
    with pytest.raises(SnbError):
        b.plan()                     # planned once
    b.close()
    nb.close()


def test_row_kernel_lds_accumulators_same_forces():
    """The row kernel's LDS-accumulator form (small systems, force-only
    steps) gives the same forces bit for bit as its global-atomic form."""
    from sponge_amd import serving
    from sponge_amd.lib import Nonbonded
    s = serving.Request(0, 1600, 10, 0.0, 77).system()
    nb = Nonbonded(s, cutoff=s["cutoff"], max_atom_in_grid_numbers=serving.MAX_ATOMS_IN_CELL)
    assert nb.window_info()["windowed"] == 0      # a box of 4 cells per axis: the row kernel
    got = []
    for mode in (0, 1):
        nb.set_row_lds(mode)
        nb.zero_outputs()
        nb.step(step=0, forced_nl=True)
        nb.sync()
        got.append(nb.frc.cpu().numpy().copy())
    np.testing.assert_array_equal(got[0], got[1])
    assert np.abs(got[0]).max() > 0 and nb.status() == 0
    nb.close()
