"""GPU edge cases against the oracle: a system without charges (PME and
exclusion terms vanish), a dilute system whose neighbour rows are empty, and
a module with no members (no bonds, no 1-4 pairs, no triangles) being a
no-op."""
import numpy as np
import pytest

from conftest import rel_rms

pytestmark = pytest.mark.gpu


def _nb(sysd, **kw):
    from sponge_amd.lib import Nonbonded
    return Nonbonded(sysd, **kw)


def test_uncharged_system_matches_oracle(water_9k):
    import oracle
    s = dict(water_9k)
    s["charge"] = np.zeros_like(np.asarray(s["charge"], np.float32))
    nb = _nb(s)
    nb.zero_outputs()
    nb.step(need_energy=True, need_virial=True, forced_nl=True)
    nb.sync()
    e = nb.energies()
    ref = oracle.nonbonded_terms_f64(s, cutoff=nb.cutoff, grid=nb.grid)
    f = nb.frc.cpu().numpy().astype(np.float64)
    print("LJ", e["LJ_short"], ref["LJ_short"], "PM", e["PM"], flush=True)
    assert rel_rms(f, ref["frc"]) < 5e-5
    assert e["PM"] == 0.0 and ref["PM"] == 0.0
    assert abs(e["LJ_short"] - ref["LJ_short"]) <= 1e-6 * abs(ref["LJ_short"])
    assert np.array_equal(nb.reciprocal_virial(), np.zeros(6))
    nb.close()


def test_dilute_system_empty_rows():
    """200 waters in a 70 A box (every row of the list empty or nearly so):
    forces and energies still match the oracle, the public rows equal it."""
    import oracle
    from sponge_amd import systems
    s = systems.tip3p_box(200, 70.0, seed=3, jitter=0.2, cutoff=8.0)
    nb = _nb(s)
    nb.zero_outputs()
    nb.step(need_energy=True, need_virial=True, forced_nl=True)
    nb.sync()
    counts, _ = nb.nl_rows()
    ref = oracle.nonbonded_terms_f64(s, cutoff=nb.cutoff, grid=nb.grid)
    assert np.array_equal(counts, ref["nl_counts"])
    e = nb.energies()
    f = nb.frc.cpu().numpy().astype(np.float64)
    print("rows with entries", int((counts > 0).sum()), "of", s["n"], flush=True)
    assert rel_rms(f, ref["frc"]) < 5e-5
    # the PM components nearly cancel here (total ~1 kcal/mol against a self
    # term of ~1e4): the bound is 1e-6 of their magnitudes
    scale = sum(abs(ref[k]) for k in ("PM_direct", "PM_reciprocal", "PM_self", "PM_correction"))
    print("PM", e["PM"], ref["PM"], "scale", scale, flush=True)
    assert abs(e["PM"] - ref["PM"]) <= 1e-6 * scale
    assert nb.status() == 0
    nb.close()


def test_modules_without_members_are_noops(water_9k):
    """bond_numbers = 0, nb14_numbers = 0 and triangle_numbers = 0 leave the
    modules uninitialised, as the reference does; their calls return 0 and
    change nothing."""
    import ctypes
    s = water_9k
    nb = _nb(s)
    L = nb.L
    z = np.zeros(1, np.int32)
    zf = np.zeros(1, np.float32)
    assert L.snb_bond_initial(nb.ctx, 0, z.ctypes.data, z.ctypes.data, zf.ctypes.data,
                              zf.ctypes.data) == 0
    assert L.snb_nb14_initial(nb.ctx, 0, z.ctypes.data, z.ctypes.data, zf.ctypes.data,
                              zf.ctypes.data, zf.ctypes.data) == 0
    nb.zero_outputs()
    nb.clear_fixed_forces()
    P = ctypes.c_void_p
    assert L.snb_bond_force(nb.ctx, P(nb.crd.data_ptr()), nb.cell, nb.rcell,
                            P(nb.frc.data_ptr()), 1, P(nb.atom_energy.data_ptr()), 1,
                            P(nb.atom_virial.data_ptr())) == 0
    assert L.snb_nb14_force(nb.ctx, P(nb.crd.data_ptr()), P(nb.charge.data_ptr()), nb.cell,
                            nb.rcell, P(nb.frc.data_ptr()), 1, P(nb.atom_energy.data_ptr()), 1,
                            P(nb.atom_virial.data_ptr())) == 0
    e = ctypes.c_double(1.0)
    assert L.snb_bond_energy(nb.ctx, ctypes.byref(e)) == 0 and e.value == 0.0
    nb.sync()
    assert not nb.frc.cpu().numpy().any()
    assert not nb.atom_energy.cpu().numpy().any()
    nb.md_initial(s, settle=False)
    assert L.snb_settle_initial(nb.ctx, 0, z.ctypes.data, z.ctypes.data, z.ctypes.data,
                                zf.ctypes.data, zf.ctypes.data, zf.ctypes.data,
                                zf.ctypes.data) == 0
    crd0 = nb.crd.cpu().numpy().copy()
    assert L.snb_settle_do(nb.ctx, P(nb.mass.data_ptr()), P(nb.crd.data_ptr()), nb.cell,
                           nb.rcell, P(nb.vel.data_ptr()), 0, P(nb.stress.data_ptr()),
                           ctypes.c_float(0.04), ctypes.c_float(1.0), ctypes.c_float(1.0)) == 0
    nb.sync()
    assert np.array_equal(nb.crd.cpu().numpy(), crd0)
    nb.close()
