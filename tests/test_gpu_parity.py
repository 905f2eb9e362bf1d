"""GPU parity of the HIP path against the CPU oracle.

Bit-exact: neighbor-list sets (vs the brute-force restatement of the
reference's criterion) and every int64 fixed-point accumulator (vs the FP32
twin given the same inputs).  Tolerance: forces 5e-5 relative RMS and
energies/virials 1e-6 relative against the FP64 oracle; the reference's own
GROMACS-comparison tolerances on the committed fixture.
"""
import json
import os

import numpy as np
import pytest

from conftest import GOLDEN, gromacs_case, rel_rms

pytestmark = pytest.mark.gpu

FORCE_TOL = 5e-5        # per-atom force, relative RMS vs FP64
ENERGY_TOL = 1e-6       # energy terms, relative vs FP64


def _nb(sysd, **kw):
    from sponge_amd.lib import Nonbonded
    return Nonbonded(sysd, **kw)


def _excl(sysd):
    return (sysd["excl_start"], sysd["excl_numbers"], sysd["excl_list"])


def _sorted_rows(counts, rows):
    return [np.sort(rows[i, :counts[i]]) for i in range(len(counts))]


SYSTEMS = ["tip3p_1011", "gromacs_iter0", "water_9k", "small_water"]


def _get(request, name):
    return request.getfixturevalue(name)


@pytest.mark.parametrize("name", SYSTEMS)
def test_neighbor_list_sets_equal_oracle(request, name):
    import oracle
    s = _get(request, name)
    nb = _nb(s)
    print("case", name, "n", s["n"], flush=True)
    nb.nl_update(forced=True)
    gc, gr = nb.nl_rows()
    assert nb.status() == 0
    oc, orows, ov = oracle.nl_build(s["crd"], s["cell"], s["rcell"], nb.cutoff, 2.0, _excl(s),
                                    max_nb=nb.max_nb)
    assert not ov
    np.testing.assert_array_equal(gc, oc)
    for i, (a, b) in enumerate(zip(_sorted_rows(gc, gr), _sorted_rows(oc, orows))):
        assert np.array_equal(a, b), "row %d differs" % i
    nb.close()


@pytest.mark.parametrize("name", SYSTEMS)
@pytest.mark.parametrize("energy,virial,owner,stream", [(False, False, False, 0),
                                                        (False, False, False, 1),
                                                        (True, True, False, 0),
                                                        (True, True, True, 0)])
def test_pair_kernel_bit_exact_with_twin(request, name, energy, virial, owner, stream):
    """Default: per-atom terms on the lower id, as the reference's rows hold
    them (twin over the public list).  owner=True: per-atom terms on the row
    owner of the internal list (twin over the internal rows).  stream: the
    force-only loop pipelined within rows (0) or across row ends (1)."""
    import oracle
    s = _get(request, name)
    nb = _nb(s)
    print("case", name, energy, virial, owner, stream, flush=True)
    nb.set_owner_attribution(owner)
    nb.set_row_stream(stream)
    nb.nl_update(forced=True)
    counts, rows = nb.internal_rows() if owner else nb.nl_rows()
    nb.lj_direct(need_energy=energy, need_virial=virial, frc=False)
    g_f = nb.fixed_forces()
    t_f, t_elj, t_ec, t_v = oracle.pair_twin(s["crd"], s["charge"], s["lj_type"], s["lj_a"],
                                             s["lj_b"], s["cell"], s["rcell"], nb.cutoff,
                                             nb.beta, counts, rows, virial=True)
    np.testing.assert_array_equal(g_f, t_f)
    # every pair adds +f to i and -f to j: the fixed-point total is exactly 0
    assert np.all(g_f.sum(axis=1) == 0)
    if energy:
        np.testing.assert_array_equal(nb.fixed_atom_terms(0), t_elj)
        np.testing.assert_array_equal(nb.fixed_atom_terms(1), t_ec)
    if virial:
        np.testing.assert_array_equal(nb.fixed_atom_terms(3), t_v)
    nb.close()


@pytest.mark.parametrize("name", ["gromacs_iter0", "water_9k"])
def test_exclusion_correction_bit_exact_with_twin(request, name):
    import oracle
    s = _get(request, name)
    nb = _nb(s)
    nb.excluded(need_energy=True, need_virial=True, frc=False)
    g_f = nb.fixed_forces()
    t_f, t_e, t_v = oracle.excl_twin(s["crd"], s["charge"], s["cell"], s["rcell"], nb.beta,
                                     _excl(s))
    np.testing.assert_array_equal(g_f, t_f)
    np.testing.assert_array_equal(nb.fixed_atom_terms(2), t_e)
    np.testing.assert_array_equal(nb.fixed_atom_terms(3), t_v)
    nb.close()


@pytest.mark.parametrize("name", SYSTEMS)
def test_pme_spread_and_gather_bit_exact_with_twin(request, name):
    """Charge grid and gathered forces bit-exact with the twin; the second of
    two calls (the first set the brick order the prep kernel visits)."""
    import oracle
    s = _get(request, name)
    nb = _nb(s)
    for rep in range(2):
        nb.clear_fixed_forces()
        nb.reciprocal(step=0, frc=False)
    print("case", name, flush=True)
    q_gpu = nb.pme_grid(0)
    q_twin = oracle.pme_spread_twin(s["crd"], s["charge"], s["rcell"], nb.grid)
    q_twin_f = (q_twin.astype(np.float64) * 2.0 ** -32).astype(np.float32)
    np.testing.assert_array_equal(q_gpu.view(np.int32), q_twin_f.view(np.int32))
    phi = nb.pme_grid(1)
    g_f = nb.fixed_forces()
    t_f = oracle.pme_gather_twin(s["crd"], s["charge"], s["rcell"], nb.grid, phi)
    np.testing.assert_array_equal(g_f, t_f)
    nb.close()


@pytest.mark.parametrize("overlap", [True, False])
def test_pme_update_interval_two(small_water, overlap):
    """PME_update_interval = 2 (PM_force.cpp:1079-1163): on step 0 the
    reciprocal force is twice the interval-1 gather (bit-exact with the twin
    at update_interval 2) and energy and virial match the FP64 oracle; on
    step 1, with the atoms moved, no force and no virial are added and the
    reported reciprocal energy is still that of step 0's grids; step 2 updates
    again at the moved coordinates, equal to a fresh interval-1 context there."""
    import torch
    import oracle
    s = small_water
    nb = _nb(s, update_interval=2)
    nb.set_pme_overlap(overlap)
    nb.zero_outputs()
    nb.reciprocal(step=0, need_energy=True, need_virial=True, frc=False)
    phi = nb.pme_grid(1)
    g0 = nb.fixed_forces()
    np.testing.assert_array_equal(
        g0, oracle.pme_gather_twin(s["crd"], s["charge"], s["rcell"], nb.grid, phi, 2.0))
    ref = oracle.nonbonded_terms_f64(s, cutoff=nb.cutoff, grid=nb.grid)
    e0 = nb.energies()
    assert abs(e0["PM_reciprocal"] - ref["PM_reciprocal"]) <= ENERGY_TOL * abs(ref["PM"])
    assert abs(e0["PM_self"] - ref["PM_self"]) <= ENERGY_TOL * abs(ref["PM"])
    v0 = nb.reciprocal_virial()
    assert np.max(np.abs(v0 - ref["virial_recip"])) <= 1e-5 * np.max(np.abs(ref["virial_recip"]))
    pot0 = float(nb.potential.cpu()[0])
    assert np.any(nb.virial.cpu().numpy() != 0)
    # step 1: atoms moved, no update
    rng = np.random.default_rng(5)
    moved = (s["crd"] + rng.uniform(-0.05, 0.05, s["crd"].shape)).astype(np.float32)
    nb.crd.copy_(torch.from_numpy(moved).to(nb.crd.device))
    nb.clear_fixed_forces()
    nb.zero_outputs()
    nb.reciprocal(step=1, need_energy=True, need_virial=True, frc=False)
    assert not np.any(nb.fixed_forces())
    assert not np.any(nb.virial.cpu().numpy())
    e1 = nb.energies()
    assert e1["PM_reciprocal"] == e0["PM_reciprocal"]
    assert e1["PM_self"] == e0["PM_self"]
    assert float(nb.potential.cpu()[0]) == pot0
    # step 2: update at the moved coordinates
    nb.zero_outputs()
    nb.reciprocal(step=2, need_energy=True, need_virial=True, frc=False)
    g2 = nb.fixed_forces()
    phi2 = nb.pme_grid(1)
    np.testing.assert_array_equal(
        g2, oracle.pme_gather_twin(moved, s["charge"], s["rcell"], nb.grid, phi2, 2.0))
    s2 = dict(s)
    s2["crd"] = moved
    nb1 = _nb(s2)
    nb1.set_pme_overlap(overlap)
    nb1.reciprocal(step=0, need_energy=True, need_virial=True, frc=False)
    np.testing.assert_array_equal(phi2, nb1.pme_grid(1))
    e2, e21 = nb.energies(), nb1.energies()
    assert e2["PM_reciprocal"] == e21["PM_reciprocal"]
    assert list(nb.reciprocal_virial()) == list(nb1.reciprocal_virial())
    nb.close()
    nb1.close()


@pytest.mark.parametrize("grid", [(48, 48, 48), (24, 24, 20), (28, 40, 36), (72, 72, 72),
                                  (96, 96, 96), (30, 42, 56)])
def test_fft_matches_float64_dft_and_rocfft(water_9k, grid):
    """The library's radix-4/2/3/5/7 R2C/C2R against numpy's float64 FFT and,
    as the independent cross-check, torch.fft (rocFFT)."""
    import ctypes

    import torch
    from sponge_amd import lib
    nb = _nb(water_9k, grid=grid)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(grid).astype(np.float32)
    xd = torch.as_tensor(x).cuda()
    fc = grid[2] // 2 + 1
    spec = torch.empty(grid[0] * grid[1] * fc * 2, dtype=torch.float32, device="cuda")
    lib.check(nb.L.snb_fft_r2c(nb.ctx, ctypes.c_void_p(xd.data_ptr()),
                               ctypes.c_void_p(spec.data_ptr())), "snb_fft_r2c")
    nb.sync()
    got = spec.view(grid[0], grid[1], fc, 2).cpu().numpy()
    got = got[..., 0] + 1j * got[..., 1]
    ref = np.fft.rfftn(x.astype(np.float64))
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < 2e-6, err
    roc = torch.fft.rfftn(xd).cpu().numpy()
    assert np.max(np.abs(got - roc)) / np.max(np.abs(ref)) < 4e-6
    back = torch.empty(grid, dtype=torch.float32, device="cuda")
    lib.check(nb.L.snb_fft_c2r(nb.ctx, ctypes.c_void_p(spec.data_ptr()),
                               ctypes.c_void_p(back.data_ptr())), "snb_fft_c2r")
    nb.sync()
    nall = grid[0] * grid[1] * grid[2]
    rt = back.cpu().numpy() / nall
    assert np.max(np.abs(rt - x)) < 2e-5
    nb.close()


def _full_terms(nb, step=0):
    nb.zero_outputs()
    nb.step(step=step, need_energy=True, need_virial=True, forced_nl=True, fused=True)
    nb.sync()
    e = nb.energies()
    return e, nb.frc.cpu().numpy().astype(np.float64)


@pytest.mark.parametrize("name", ["tip3p_1011", "gromacs_iter0", "water_9k"])
def test_forces_and_energies_vs_fp64_oracle(request, name):
    import oracle
    s = _get(request, name)
    nb = _nb(s)
    e, f = _full_terms(nb)
    ref = oracle.nonbonded_terms_f64(s, cutoff=nb.cutoff, grid=nb.grid)
    err = rel_rms(f, ref["frc"])
    print("case", name, "force rel RMS", err, flush=True)
    assert err < FORCE_TOL
    assert abs(e["LJ_short"] - ref["LJ_short"]) <= ENERGY_TOL * abs(ref["LJ_short"])
    # the PM components cancel heavily (self ~ -correction); each is held to
    # 1e-6 of the printed PM total, and PM itself to 1e-6 relative
    for k in ("PM_direct", "PM_correction", "PM_self", "PM_reciprocal"):
        print(k, e[k], ref[k], (e[k] - ref[k]) / abs(ref[k]), flush=True)
        assert abs(e[k] - ref[k]) <= ENERGY_TOL * abs(ref["PM"]), (k, e[k], ref[k])
    assert abs(e["PM"] - ref["PM"]) <= ENERGY_TOL * abs(ref["PM"])
    assert abs(e["LJ_long"] - ref["LJ_long"]) <= 1e-6 * abs(ref["LJ_long"]) + 1e-9
    # reciprocal virial (reference PME_Sum_Virial convention)
    v = nb.reciprocal_virial()
    scale = np.max(np.abs(ref["virial_recip"]))
    assert np.max(np.abs(v - ref["virial_recip"])) <= 1e-5 * scale
    # per-atom virial of the direct + exclusion terms (fixed point, flushed to float)
    av = nb.atom_virial.cpu().numpy().astype(np.float64)
    assert rel_rms(av, ref["virial_atoms"]) < FORCE_TOL
    nb.close()


@pytest.mark.parametrize("it", [0, 1, 2])
def test_gromacs_fixture_within_reference_tolerances(it):
    """benchmarks/comparison/tests/gromacs/tests/comp_alanine_dipeptide.py:106-115:
    |LJ_short - lj_sr| <= 25 and |PM - (coulomb_sr + coulomb_recip)| <= 2.0."""
    s = gromacs_case(it)
    ref = json.load(open(os.path.join(GOLDEN, "gromacs_charmm_tip3p", "reference_terms.json")))
    entry = [x for x in ref["entries"] if x["iteration"] == it][0]
    nb = _nb(s, cutoff=12.0)
    e, _ = _full_terms(nb)
    print("iter", it, "LJ_short", e["LJ_short"], "PM", e["PM"], flush=True)
    assert abs(e["LJ_short"] - entry["lj_sr"]) <= 25.0
    assert abs(e["PM"] - (entry["coulomb_sr"] + entry["coulomb_recip"])) <= 2.0
    nb.close()


def test_reference_cpu_values_on_fixture(gromacs_iter0, tip3p_1011):
    """Values the reference's own CPU build printed for the committed inputs
    (recorded when the reference was probed): GROMACS iter0 LJ_short 1237.962,
    PM -7437.531 (rc 12); TIP3P 1011 LJ_short 508.1525, PM -3712.9453 (rc 8).
    Agreement to the reference's float precision (~1e-6 relative)."""
    for s, rc, lj, pm in ((gromacs_iter0, 12.0, 1237.962, -7437.531),
                          (tip3p_1011, 8.0, 508.1525, -3712.9453)):
        nb = _nb(s, cutoff=rc)
        e, _ = _full_terms(nb)
        assert abs(e["LJ_short"] - lj) < 2e-6 * abs(lj) + 1e-3
        assert abs(e["PM"] - pm) < 2e-6 * abs(pm) + 1e-3
        nb.close()


def test_determinism_two_runs_identical(water_9k):
    """Forces, energies and (with sorted rows) the public list are the same
    bits on a second run."""
    nb = _nb(water_9k)
    nb.set_export(2)
    outs = []
    for _ in range(2):
        nb.clear_fixed_forces()
        nb.nl_update(forced=True)
        nb.excluded(True, True, frc=False)
        nb.lj_direct(True, True, frc=False)
        nb.reciprocal(0, True, True, frc=False)
        outs.append((nb.fixed_forces(), nb.energies(), nb.nl_rows()))
    np.testing.assert_array_equal(outs[0][0], outs[1][0])
    assert outs[0][1] == outs[1][1]
    np.testing.assert_array_equal(outs[0][2][0], outs[1][2][0])
    np.testing.assert_array_equal(outs[0][2][1], outs[1][2][1])
    nb.close()


def test_translation_invariance_pme(water_9k):
    """Shifting every atom by a lattice-commensurate vector leaves the grid
    relation unchanged: PME energy is invariant to a shift by whole grid cells."""
    s = dict(water_9k)
    nb = _nb(s)
    e0, f0 = _full_terms(nb)
    nb.close()
    h = np.float32(s["box"][0]) / np.float32(48.0) * np.float32(3.0)
    s2 = dict(s)
    s2["crd"] = (s["crd"] + h).astype(np.float32)
    nb = _nb(s2)
    e1, f1 = _full_terms(nb)
    nb.close()
    assert abs(e1["PM_reciprocal"] - e0["PM_reciprocal"]) < 1e-4 * abs(e0["PM"])
    assert rel_rms(f1, f0) < 1e-4


def test_nonfinite_coordinate_raises_flag(small_water):
    s = dict(small_water)
    crd = s["crd"].copy()
    crd[5, 1] = np.nan
    s["crd"] = crd
    nb = _nb(s)
    nb.reciprocal(0, frc=False)   # PME treats the atom as frac 0 (PM_force.cpp:710-714)
    nb.lj_direct(frc=False)        # packing raises the flag
    from sponge_amd.lib import FLAG_NONFINITE_CRD
    assert nb.status(clear=True) & FLAG_NONFINITE_CRD
    assert nb.status() == 0
    nb.close()


def test_neighbor_overflow_flag_and_truncation(small_water):
    from sponge_amd.lib import FLAG_NL_OVERFLOW
    nb = _nb(small_water, max_neighbor_numbers=16)
    nb.nl_update(forced=True)
    counts, rows = nb.nl_rows()
    assert counts.max() == 16
    assert nb.status() & FLAG_NL_OVERFLOW
    nb.close()


def test_box_too_small_is_an_error(small_water):
    from sponge_amd.lib import SnbError
    with pytest.raises(SnbError) as ei:
        _nb(small_water, cutoff=12.0)   # (12+2)/2 = 7 A cells, 18.7 A box -> 2 cells
    assert ei.value.code == 3


def test_conditional_rebuild_policy(small_water):
    """refresh_interval 0: rebuild only when an atom moved more than
    skin * skin_permit (neighbor_list.cpp:166-180); the stale list is kept
    otherwise."""
    import torch
    nb = _nb(small_water, refresh_interval=0)
    nb.nl_update(forced=True)
    c0, r0 = nb.nl_rows()
    # move one atom by 0.9 A (< 1 A): no rebuild, rows unchanged
    nb.crd[0, 0] += 0.9
    nb.nl_update(step=1, forced=False)
    c1, r1 = nb.nl_rows()
    np.testing.assert_array_equal(c0, c1)
    # move it by 1.1 A in total from the rebuild point: rebuild
    nb.crd[0, 0] += 0.2
    nb.nl_update(step=2, forced=False)
    c2, _ = nb.nl_rows()
    import oracle
    s = small_water
    oc, _, _ = oracle.nl_build(nb.crd.cpu().numpy(), s["cell"], s["rcell"], nb.cutoff, 2.0,
                               _excl(s), max_nb=nb.max_nb)
    np.testing.assert_array_equal(c2, oc)
    nb.close()


# --------------------------------------------------------------------------
# windowed pair path (cell-ordered internal list, LDS force accumulators)
# --------------------------------------------------------------------------
def _pairs(counts, rows):
    out = set()
    for i in range(len(counts)):
        for j in rows[i, :counts[i]]:
            out.add((min(i, int(j)), max(i, int(j))))
    return out


@pytest.mark.parametrize("name,windowed", [("water_9k", 1), ("gromacs_iter0", 1),
                                           ("mixed_small", 1), ("small_water", 0),
                                           ("tip3p_1011", 0)])
def test_window_path_selection_and_pair_set(request, name, windowed):
    """Every axis >= 5 cells selects the windowed path; its internal rows hold
    exactly the public half list's pairs, each once."""
    s = _get(request, name)
    nb = _nb(s)
    info = nb.window_info()
    print("case", name, info, flush=True)
    assert info["windowed"] == windowed
    nb.nl_update(forced=True)
    ic, ir = nb.internal_rows()
    pc, pr = nb.nl_rows()
    assert nb.status() == 0
    assert ic.sum() == pc.sum()
    assert _pairs(ic, ir) == _pairs(pc, pr)
    nb.close()


def test_window_pair_bit_exact_mixed(mixed_small):
    """Bead chains (exclusions, 602 LJ types) through the windowed kernel."""
    import oracle
    s = mixed_small
    nb = _nb(s)
    assert nb.window_info()["windowed"] == 1
    nb.nl_update(forced=True)
    counts, rows = nb.nl_rows()
    nb.lj_direct(need_energy=True, need_virial=True, frc=False)
    g_f = nb.fixed_forces()
    t_f, t_elj, t_ec, t_v = oracle.pair_twin(s["crd"], s["charge"], s["lj_type"], s["lj_a"],
                                             s["lj_b"], s["cell"], s["rcell"], nb.cutoff,
                                             nb.beta, counts, rows, virial=True)
    np.testing.assert_array_equal(g_f, t_f)
    np.testing.assert_array_equal(nb.fixed_atom_terms(0), t_elj)
    np.testing.assert_array_equal(nb.fixed_atom_terms(1), t_ec)
    np.testing.assert_array_equal(nb.fixed_atom_terms(3), t_v)
    oc, orows, _ = oracle.nl_build(s["crd"], s["cell"], s["rcell"], nb.cutoff, 2.0, _excl(s),
                                   max_nb=nb.max_nb)
    assert _pairs(counts, rows) == _pairs(oc, orows)
    ic, ir = nb.internal_rows()
    assert _pairs(ic, ir) == _pairs(oc, orows)
    nb.close()


@pytest.mark.parametrize("name", ["water_9k", "mixed_small"])
def test_export_modes_same_rows(request, name):
    """The append-order export (default) and the sorted export hold the same
    set in every row; sorted rows ascend; both equal the oracle's list."""
    import oracle
    s = _get(request, name)
    nb = _nb(s)
    nb.set_export(1)
    nb.nl_update(forced=True)
    c1, r1 = nb.nl_rows()
    nb.set_export(2)
    nb.nl_update(forced=True)
    c2, r2 = nb.nl_rows()
    assert nb.status() == 0
    np.testing.assert_array_equal(c1, c2)
    for i in range(len(c1)):
        row2 = r2[i, :c2[i]]
        assert np.all(np.diff(row2) > 0)
        np.testing.assert_array_equal(np.sort(r1[i, :c1[i]]), row2)
    oc, orows, _ = oracle.nl_build(s["crd"], s["cell"], s["rcell"], nb.cutoff, 2.0, _excl(s),
                                   max_nb=nb.max_nb)
    np.testing.assert_array_equal(c2, oc)
    assert _pairs(c2, r2) == _pairs(oc, orows)
    nb.close()


def test_window_export_off_same_forces(water_9k):
    """With the public list export switched off the pair forces are unchanged."""
    outs = []
    for export in (True, False):
        nb = _nb(water_9k)
        nb.set_export(export)
        nb.nl_update(forced=True)
        nb.lj_direct(True, True, frc=False)
        outs.append((nb.fixed_forces(), nb.fixed_atom_terms(0), nb.fixed_atom_terms(3)))
        nb.close()
    for a, b in zip(*outs):
        np.testing.assert_array_equal(a, b)


def test_window_neighbor_overflow_flag(water_9k):
    from sponge_amd.lib import FLAG_NL_OVERFLOW
    nb = _nb(water_9k, max_neighbor_numbers=16)
    assert nb.window_info()["windowed"] == 1
    nb.nl_update(forced=True)
    counts, _ = nb.internal_rows()
    assert counts.max() == 16
    assert nb.status() & FLAG_NL_OVERFLOW
    nb.close()


def test_window_rows_follow_moving_atoms(water_9k):
    """Between rebuilds atoms keep their cell slots: after a displacement the
    forces still equal the twin on the (stale) internal rows."""
    import oracle
    s = water_9k
    nb = _nb(s, refresh_interval=10)
    nb.nl_update(step=0, forced=True)
    rng = np.random.default_rng(3)
    moved = (s["crd"] + rng.uniform(-0.3, 0.3, s["crd"].shape)).astype(np.float32)
    import torch
    nb.crd.copy_(torch.from_numpy(moved))
    nb.nl_update(step=1, forced=False)   # (1+1) % 10 != 0: no rebuild
    counts, rows = nb.internal_rows()
    nb.clear_fixed_forces()
    nb.lj_direct(frc=False)
    t_f, _, _, _ = oracle.pair_twin(moved, s["charge"], s["lj_type"], s["lj_a"], s["lj_b"],
                                    s["cell"], s["rcell"], nb.cutoff, nb.beta, counts, rows,
                                    virial=False)
    np.testing.assert_array_equal(nb.fixed_forces(), t_f)
    nb.close()


def test_window_93k_workload_pair_exact(request):
    """The 93k-atom workload (30,000 waters + 3,000 beads, 96.5 A box): pair
    forces bit-exact with the twin, fixed-point total exactly zero, on an
    energy step and on force-only steps in the default hand-off of the window
    forces (staged gather at this size) and with per-block atomic adds."""
    import oracle
    from sponge_amd import systems
    s = systems.config_gpu_93k()
    nb = _nb(s, grid=s["pme_grid"])
    print("case 93k", nb.window_info(), flush=True)
    nb.nl_update(forced=True)
    counts, rows = nb.nl_rows()
    nb.lj_direct(True, False, frc=False)
    g_f = nb.fixed_forces()
    assert nb.status() == 0
    assert np.all(g_f.sum(axis=1) == 0)
    t_f, t_elj, t_ec, _ = oracle.pair_twin(s["crd"], s["charge"], s["lj_type"], s["lj_a"],
                                           s["lj_b"], s["cell"], s["rcell"], nb.cutoff,
                                           nb.beta, counts, rows, virial=False)
    np.testing.assert_array_equal(g_f, t_f)
    np.testing.assert_array_equal(nb.fixed_atom_terms(0), t_elj)
    for mode in (-1, 1):
        for stream in (0, 1):
            print("case 93k force-only, hand-off mode", mode, "stream", stream, flush=True)
            nb.set_direct_j(mode)
            nb.set_row_stream(stream)
            nb.clear_fixed_forces()
            nb.lj_direct(False, False, frc=False)
            np.testing.assert_array_equal(nb.fixed_forces(), t_f)
    nb.close()


# --------------------------------------------------------------------------
# HIP graph replay of a whole step
# --------------------------------------------------------------------------
def test_graph_replay_matches_eager_every_replay(water_9k):
    """A captured force step replays to the same float forces as the eager
    step, on every replay, with eager work and synchronisations in between."""
    import torch
    nb = _nb(water_9k, refresh_interval=10)
    nb.set_export(False)
    nb.step(step=0, forced_nl=True)          # builds the list (step 0 rebuilds)
    torch.cuda.synchronize()
    nb.zero_outputs()
    nb.step(step=1)                          # eager reference: no rebuild, no energy
    torch.cuda.synchronize()
    ref = nb.frc.cpu().numpy().copy()
    graph = nb.capture_step(step=1)
    for rep in range(4):
        nb.zero_outputs()
        graph.replay()
        torch.cuda.synchronize()
        got = nb.frc.cpu().numpy()
        print("replay", rep, flush=True)
        np.testing.assert_array_equal(got, ref)
        # eager work between replays
        nb.zero_outputs()
        nb.step(step=1 + rep)
        torch.cuda.synchronize()
        np.testing.assert_array_equal(nb.frc.cpu().numpy(), ref)
    assert nb.status() == 0
    nb.close()


def test_graph_replay_of_rebuild_and_energy_steps(water_9k):
    """Rebuild and energy steps captured as graphs give the eager results."""
    import torch
    nb = _nb(water_9k, refresh_interval=10)
    nb.set_export(False)
    nb.step(step=0, forced_nl=True)
    torch.cuda.synchronize()
    nb.zero_outputs()
    nb.step(step=9, need_energy=True, need_virial=True)   # (9+1) % 10 == 0: rebuild
    torch.cuda.synchronize()
    ref_f = nb.frc.cpu().numpy().copy()
    ref_e = nb.energies()
    ref_v = nb.atom_virial.cpu().numpy().copy()
    graph = nb.capture_step(step=9, need_energy=True, need_virial=True)
    for rep in range(3):
        nb.zero_outputs()
        graph.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(nb.frc.cpu().numpy(), ref_f)
        np.testing.assert_array_equal(nb.atom_virial.cpu().numpy(), ref_v)
        assert nb.energies() == ref_e
    nb.close()


@pytest.mark.parametrize("slots", [64, 1500])
def test_window_overflow_path_bit_exact(water_9k, slots):
    """Blocks whose window exceeds the LDS reservation use global integer
    atomics: the forces and per-atom terms are unchanged."""
    import oracle
    s = water_9k
    nb = _nb(s)
    nb.set_window_capacity(slots)
    print("case slots", slots, nb.window_info(), flush=True)
    nb.nl_update(forced=True)
    counts, rows = nb.nl_rows()
    nb.lj_direct(need_energy=True, need_virial=True, frc=False)
    t_f, t_elj, t_ec, t_v = oracle.pair_twin(s["crd"], s["charge"], s["lj_type"], s["lj_a"],
                                             s["lj_b"], s["cell"], s["rcell"], nb.cutoff,
                                             nb.beta, counts, rows, virial=True)
    np.testing.assert_array_equal(nb.fixed_forces(), t_f)
    np.testing.assert_array_equal(nb.fixed_atom_terms(0), t_elj)
    np.testing.assert_array_equal(nb.fixed_atom_terms(1), t_ec)
    np.testing.assert_array_equal(nb.fixed_atom_terms(3), t_v)
    nb.close()


@pytest.mark.parametrize("slots", [None, 64])
@pytest.mark.parametrize("stream", [0, 1])
def test_window_direct_j_adds_bit_exact(water_9k, slots, stream):
    """Force-only steps with the window forces added atomically per block
    (mode 1) match the twin bit for bit, as the staged gather (mode 0) does,
    with either pair loop; 64 slots puts most blocks on the overflow path."""
    import oracle
    s = water_9k
    t_f = None
    for mode in (0, 1):
        nb = _nb(s)
        if slots:
            nb.set_window_capacity(slots)
        nb.set_direct_j(mode)
        nb.set_row_stream(stream)
        print("case mode", mode, "slots", slots, "stream", stream, nb.window_info(), flush=True)
        nb.nl_update(forced=True)
        if t_f is None:
            counts, rows = nb.nl_rows()
            t_f = oracle.pair_twin(s["crd"], s["charge"], s["lj_type"], s["lj_a"], s["lj_b"],
                                   s["cell"], s["rcell"], nb.cutoff, nb.beta, counts, rows)[0]
        for rep in range(2):                   # the second call adds on top of the first
            nb.lj_direct(need_energy=False, need_virial=False, frc=False)
        np.testing.assert_array_equal(nb.fixed_forces(), 2 * t_f)
        nb.close()
    nb = _nb(s)
    with pytest.raises(RuntimeError):
        nb.set_direct_j(2)
    nb.close()


def test_period_graph_matches_eager_steps(water_9k):
    """Ten whole force steps (a neighbour-list period, the rebuild in its last
    step, energies and virials in its first) captured in one HIP graph give
    the eager steps' accumulated float forces, energies and virials on every
    replay."""
    import torch
    nb = _nb(water_9k, refresh_interval=10)
    nb.set_export(False)
    nb.step(step=0, forced_nl=True)
    torch.cuda.synchronize()

    def period(k0):
        for k in range(k0, k0 + 10):
            e = k % 100 == 0
            nb.step(step=k, need_energy=e, need_virial=e)

    nb.zero_outputs()
    period(100)
    torch.cuda.synchronize()
    ref = (nb.frc.cpu().numpy().copy(), nb.atom_virial.cpu().numpy().copy(), nb.energies())
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream(device=nb.dev)
    home = nb.stream
    nb.sync()
    torch.cuda.synchronize()
    nb.set_stream(side)
    try:
        with torch.cuda.graph(g, stream=side):
            period(200)
    finally:
        nb.set_stream(home)
    torch.cuda.synchronize()
    for rep in range(3):
        nb.zero_outputs()
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(nb.frc.cpu().numpy(), ref[0])
        np.testing.assert_array_equal(nb.atom_virial.cpu().numpy(), ref[1])
        assert nb.energies() == ref[2]
        nb.step(step=5)                       # eager work between the replays
        torch.cuda.synchronize()
    nb.close()


@pytest.mark.parametrize("overlap", [True, False])
def test_pme_stream_overlap_same_results(water_9k, overlap):
    """The reciprocal pipeline on its own stream (and, on force-only fused
    steps, the exclusion correction, 1-4 pairs and bonds with it) gives the
    same fused forces, energies and virials as everything on the context
    stream."""
    import torch
    res = []
    for ov in (False, overlap):
        nb = _nb(water_9k, refresh_interval=10)
        nb.set_pme_overlap(ov)
        nb.step(step=0, forced_nl=True, need_energy=True, need_virial=True)
        nb.zero_outputs()
        nb.step(step=1)                        # force-only: exclusion + reciprocal on the PME stream
        torch.cuda.synchronize()
        f_plain = nb.frc.cpu().numpy().copy()
        nb.zero_outputs()
        nb.step(step=2, need_energy=True, need_virial=True)
        torch.cuda.synchronize()
        res.append((nb.frc.cpu().numpy().copy(), nb.atom_virial.cpu().numpy().copy(),
                    nb.energies(), nb.reciprocal_virial(), f_plain))
        nb.close()
    np.testing.assert_array_equal(res[0][0], res[1][0])
    np.testing.assert_array_equal(res[0][1], res[1][1])
    assert res[0][2] == res[1][2]
    assert list(res[0][3]) == list(res[1][3])
    np.testing.assert_array_equal(res[0][4], res[1][4])
    np.testing.assert_array_equal(res[0][4], res[0][0])   # fixed coordinates: same forces


def _with_zero_charges(sysd):
    """Copy of a system with the charges of the first few 1-4 atoms set to 0
    (the reference raises |q| < 1e-10 to 1e-10 in the 1-4 kernel)."""
    s = dict(sysd)
    q = np.array(s["charge"], np.float32).copy()
    q[np.asarray(s["nb14"][0][:3])] = 0.0
    s["charge"] = q
    return s


@pytest.mark.parametrize("name", ["gromacs_iter0", "mixed_small", "gromacs_q0"])
@pytest.mark.parametrize("energy,virial", [(False, False), (True, True)])
def test_nb14_bit_exact_with_twin(request, name, energy, virial):
    import oracle
    s = _with_zero_charges(request.getfixturevalue("gromacs_iter0")) if name == "gromacs_q0" \
        else _get(request, name)
    nb = _nb(s)
    print("case", name, "pairs", len(s["nb14"][0]), flush=True)
    nb.nb14_force(need_energy=energy, need_virial=virial, frc=False)
    g_f = nb.fixed_forces()
    t_f, t_ep, t_ea, t_v = oracle.nb14_twin(s["crd"], s["charge"], s["cell"], s["rcell"],
                                            s["nb14"], virial=virial)
    np.testing.assert_array_equal(g_f, t_f)
    if energy:
        np.testing.assert_array_equal(nb.fixed_atom_terms(4), t_ea)
        lj, ee = nb.nb14_energy()
        assert lj == float(t_ep[0].sum()) * 2.0 ** -32
        assert ee == float(t_ep[1].sum()) * 2.0 ** -32
        f64, ep64, _, _ = oracle.nb14_f64(s["crd"], s["charge"], s["cell"], s["rcell"],
                                          s["nb14"], virial=False)
        tot = np.abs(ep64).sum()
        assert abs(lj - ep64[0].sum()) <= ENERGY_TOL * tot
        assert abs(ee - ep64[1].sum()) <= ENERGY_TOL * tot
        assert rel_rms(g_f.T * 2.0 ** -32, f64) < FORCE_TOL
    if virial:
        np.testing.assert_array_equal(nb.fixed_atom_terms(3), t_v)
    nb.close()


def _stretched(sysd, seed=12, sd=0.03):
    """Copy of a system with every coordinate moved by N(0, sd) A, so that the
    bonds sit off their rest lengths."""
    s = dict(sysd)
    rng = np.random.default_rng(seed)
    s["crd"] = (np.asarray(sysd["crd"], np.float32) +
                rng.normal(0, sd, np.shape(sysd["crd"]))).astype(np.float32)
    return s


@pytest.mark.parametrize("name", ["gromacs_iter0", "mixed_small"])
@pytest.mark.parametrize("energy,virial", [(False, False), (True, True)])
def test_bond_bit_exact_with_twin(request, name, energy, virial):
    import oracle
    s = _get(request, name)
    if name == "mixed_small":
        s = _stretched(s)
    nb = _nb(s)
    print("case", name, "bonds", len(s["bonds"][0]), flush=True)
    nb.bond_force(need_energy=energy, need_virial=virial, frc=False)
    g_f = nb.fixed_forces()
    t_f, t_ep, t_ea, t_v = oracle.bond_twin(s["crd"], s["cell"], s["rcell"], s["bonds"],
                                            virial=virial)
    np.testing.assert_array_equal(g_f, t_f)
    if energy:
        np.testing.assert_array_equal(nb.fixed_atom_terms(5), t_ea)
        e = nb.bond_energy()
        assert e == float(t_ep.sum()) * 2.0 ** -32
        f64, ep64, _, _ = oracle.bond_f64(s["crd"], s["cell"], s["rcell"], s["bonds"],
                                          virial=False)
        assert abs(e - ep64.sum()) <= ENERGY_TOL * ep64.sum()
        assert rel_rms(g_f.T * 2.0 ** -32, f64) < FORCE_TOL
    if virial:
        np.testing.assert_array_equal(nb.fixed_atom_terms(3), t_v)
    nb.close()


def test_bond_outputs_flush_and_energy_restart(gromacs_iter0):
    """Direct outputs: d_frc += bond forces, d_atom_energy += per-atom bond
    energy, d_atom_virial += f (x) dr; the per-bond energies restart at each
    energy call, so two calls give twice the force and the same energy."""
    import oracle
    s = gromacs_iter0
    nb = _nb(s)
    nb.zero_outputs()
    for _ in range(2):
        nb.bond_force(need_energy=True, need_virial=True, frc=True)
    nb.sync()
    t_f, t_ep, t_ea, t_v = oracle.bond_twin(s["crd"], s["cell"], s["rcell"], s["bonds"])
    one = (t_f.astype(np.float64) * 2.0 ** -32).astype(np.float32).T
    np.testing.assert_array_equal(nb.frc.cpu().numpy(), one + one)
    ea = (t_ea.astype(np.float64) * 2.0 ** -32).astype(np.float32)
    np.testing.assert_array_equal(nb.atom_energy.cpu().numpy(), ea + ea)
    v1 = (t_v.astype(np.float64) * 2.0 ** -32).astype(np.float32).T
    np.testing.assert_array_equal(nb.atom_virial.cpu().numpy(), v1 + v1)
    assert nb.bond_energy() == float(t_ep.sum()) * 2.0 ** -32
    nb.close()


def test_nb14_outputs_flush_and_energy_restart(gromacs_iter0):
    """Direct outputs: d_frc += forces, d_atom_energy += per-atom 1-4 energy
    (restarted at each energy call, so two calls give twice the force and
    the same energy sums)."""
    import oracle
    s = gromacs_iter0
    nb = _nb(s)
    nb.zero_outputs()
    nb.nb14_force(need_energy=True, need_virial=False, frc=True)
    nb.nb14_force(need_energy=True, need_virial=False, frc=True)
    nb.sync()
    t_f, t_ep, t_ea, _ = oracle.nb14_twin(s["crd"], s["charge"], s["cell"], s["rcell"],
                                          s["nb14"], virial=False)
    one = (t_f.astype(np.float64) * 2.0 ** -32).astype(np.float32).T
    f = nb.frc.cpu().numpy()
    np.testing.assert_array_equal(f, one + one)
    e = nb.atom_energy.cpu().numpy()
    ea = (t_ea.astype(np.float64) * 2.0 ** -32).astype(np.float32)
    np.testing.assert_array_equal(e, ea + ea)
    lj, ee = nb.nb14_energy()
    assert lj == float(t_ep[0].sum()) * 2.0 ** -32
    nb.close()


# --------------------------------------------------------------------------
# box change (barostat step)
# --------------------------------------------------------------------------
def test_box_update_matches_fresh_context(water_9k):
    """Particle_Mesh::Update_Box + a neighbour-list update on the new cell:
    after the box and the coordinates are scaled by 1.004, the step equals,
    bit for bit, the step of a context built on the scaled system (same grid),
    and stays within the force/energy bounds of the FP64 oracle; a reciprocal
    call with a stale rcell and a box below the initial cells' width are
    refused before anything is launched."""
    import oracle
    from sponge_amd import native_io
    from sponge_amd.lib import SnbError
    s = water_9k
    nb = _nb(s)
    e0, f0 = _full_terms(nb)
    lam = 1.004
    s2 = dict(s)
    s2["box"] = (np.asarray(s["box"], np.float64) * lam).astype(np.float32)
    s2["crd"] = (np.asarray(s["crd"], np.float64) * lam).astype(np.float32)
    s2["cell"], s2["rcell"] = native_io.box_to_cell(s2["box"], s.get("angle", (90, 90, 90)))
    old_rcell = nb.rcell
    nb.update_box(s2["box"])
    nb.crd.copy_(nb.torch.as_tensor(s2["crd"]))
    e1, f1 = _full_terms(nb)
    nb2 = _nb(s2, grid=nb.grid)
    e2, f2 = _full_terms(nb2)
    print("PM", e0["PM"], e1["PM"], e2["PM"], flush=True)
    np.testing.assert_array_equal(f1, f2)
    assert e1 == e2
    assert list(nb.reciprocal_virial()) == list(nb2.reciprocal_virial())
    ref = oracle.nonbonded_terms_f64(s2, cutoff=nb.cutoff, grid=nb.grid)
    assert rel_rms(f1, ref["frc"]) < FORCE_TOL
    assert abs(e1["PM"] - ref["PM"]) <= ENERGY_TOL * abs(ref["PM"])
    assert abs(e1["LJ_long"] - ref["LJ_long"]) <= 1e-6 * abs(ref["LJ_long"])
    # stale rcell: refused
    with pytest.raises(SnbError):
        from sponge_amd.lib import check, _ptr
        check(nb.L.snb_pme_reciprocal_force(nb.ctx, _ptr(nb.crd), nb.cell, old_rcell,
                                            _ptr(nb.charge), None, 0, 0, None, None, 0),
              "snb_pme_reciprocal_force")
    # a box thinner than the initial cell count allows: refused
    small = np.asarray(s["box"], np.float32) * np.float32(0.9)
    cell, rcell = native_io.box_to_cell(small)
    from sponge_amd.lib import LTMatrix3
    rc = nb.L.snb_nl_update(nb.ctx, None, nb.n, 0, _ptr(nb.crd), LTMatrix3.of(cell),
                            LTMatrix3.of(rcell), 0, 1, None, None, None)
    assert rc == 3                                  # SNB_ERR_BOX_TOO_SMALL
    nb.sync()
    assert nb.status() == 0
    nb.close()
    nb2.close()


@pytest.mark.parametrize("it", [0, 1, 2])
def test_gpu_forces_on_waters_vs_gromacs(it):
    """The library's step forces (path + bonds + 1-4 pairs) plus the oracle's
    angle / Urey-Bradley aid against GROMACS 2025.4's per-atom forces on the
    fixture's water atoms, within the reference's tolerances (max |dF| <= 0.07,
    RMS <= 0.01, cosine >= 0.999; comp_alanine_dipeptide.py:113-115)."""
    import oracle
    from conftest import force_stats, gromacs_water_atoms
    s = gromacs_case(it)
    nb = _nb(s)
    nb.zero_outputs()
    nb.step(forced_nl=True)
    nb.sync()
    f = nb.frc.cpu().numpy().astype(np.float64)
    nb.close()
    fa, _ = oracle.angle_f64(s["crd"], s["cell"], s["rcell"], s["urey_bradley"][0])
    fu, _, _, _ = oracle.bond_f64(s["crd"], s["cell"], s["rcell"], s["urey_bradley"][1])
    gmx = np.load(os.path.join(GOLDEN, "gromacs_charmm_tip3p", "iter%d" % it, "gromacs_forces.npy"),
                  allow_pickle=False)
    w = gromacs_water_atoms(s)
    mx, rms, cos = force_stats(gmx[w], (f + fa + fu)[w])
    print("iteration", it, "water atoms: max", mx, "rms", rms, "cos", cos, flush=True)
    assert mx <= 0.07 and rms <= 0.01 and cos >= 0.999


def _triclinic(s, angle=(85.0, 87.0, 84.0)):
    """The system's box lengths with angles: a lower-triangular cell with
    positive tilts (native_io.box_to_cell, MD_core/pbc.hpp:66-101)."""
    from sponge_amd import native_io
    t = dict(s)
    t["cell"], t["rcell"] = native_io.box_to_cell(s["box"], angle)
    t["angle"] = angle
    return t


@pytest.mark.parametrize("name", ["small_water", "water_9k"])
def test_triclinic_neighbor_list_and_pair_kernel(request, name):
    """A triclinic cell (neighbor_list.cpp:5-48, :206-231): the list holds the
    pairs within rc + skin (the oracle's brute-force set under the triclinic
    minimum image) whose grids the reference's neighbour-grid table pairs up
    (oracle.nl_cell_rule), every row of it within the brute-force row; with
    three grids per axis every grid is a neighbour and the two sets are equal.
    The row kernel's forces over the list equal the float twin's bit for bit."""
    import oracle
    s = _triclinic(_get(request, name))
    nb = _nb(s)
    print("case", name, "cell", s["cell"], flush=True)
    assert nb.window_info()["windowed"] == 0     # triclinic cells take the row kernel
    nb.nl_update(forced=True)
    gc, gr = nb.nl_rows()
    assert nb.status() == 0
    oc, orows, ov = oracle.nl_build(s["crd"], s["cell"], s["rcell"], nb.cutoff, 2.0, _excl(s),
                                    max_nb=nb.max_nb)
    assert not ov
    rule = oracle.nl_cell_rule(s["crd"], s["cell"], nb.cutoff, 2.0)
    dropped = 0
    for i, (a, b) in enumerate(zip(_sorted_rows(gc, gr), _sorted_rows(oc, orows))):
        want = b[rule(np.full(len(b), i), b)]
        dropped += len(b) - len(want)
        assert np.array_equal(a, want), "row %d differs" % i
    print("pairs within rc + skin outside the grid table:", dropped, flush=True)
    if name == "small_water":
        assert dropped == 0
    nb.lj_direct(frc=False)
    g_f = nb.fixed_forces()
    t_f, _, _, _ = oracle.pair_twin(s["crd"], s["charge"], s["lj_type"], s["lj_a"], s["lj_b"],
                                    s["cell"], s["rcell"], nb.cutoff, nb.beta, gc, gr,
                                    virial=True)
    np.testing.assert_array_equal(g_f, t_f)
    nb.close()
