"""The precision and serving workloads on the GPU against the FP64 oracle.

- The 93,000-atom mixed system (30,000 rigid TIP3P waters + 3 x 1,000-bead
  chains, 96.5 A box, PME 96^3): Langevin equilibration, then NVE leap-frog
  steps at 1 fs with SETTLE; the total energy is tracked, and the evolved
  frame's forces, energies and virial are checked against the FP64 oracle.
- Served water-box segments (cutoff 9 A, PME spacing ~1 A) that span the
  plain and the windowed pair paths: each final frame against the FP64
  oracle, and each result the same alone and in the mix.
- The 93k system's neighbour list against the oracle's build, row by row.

Tolerances: per-atom force 5e-5 relative RMS; every energy term and the
total virial tensor 1e-6 relative.
"""
import numpy as np
import pytest

from conftest import rel_rms

pytestmark = pytest.mark.gpu

FORCE_TOL = 5e-5
ENERGY_TOL = 1e-6
VIRIAL_TOL = 1e-6


def _frame_vs_fp64(nb, s):
    """One energy/virial step of the library on s["crd"] against the FP64
    oracle on the same coordinates; returns the errors (and asserts them)."""
    import oracle
    nb.zero_outputs()
    nb.step(step=0, need_energy=True, need_virial=True, forced_nl=True, fused=True)
    nb.sync()
    e = nb.energies()
    if nb.nb14 is not None:
        e["nb14_LJ"], e["nb14_EE"] = nb.nb14_energy()
    if nb.bonds is not None:
        e["bond"] = nb.bond_energy()
    ref = oracle.nonbonded_terms_f64(s, cutoff=nb.cutoff, grid=nb.grid)
    err = {"force": rel_rms(nb.frc.cpu().numpy().astype(np.float64), ref["frc"])}
    for k in ("LJ_short", "LJ_long", "PM", "nb14_LJ", "nb14_EE", "bond"):
        if k in e and ref[k] != 0.0:
            err[k] = abs(e[k] - ref[k]) / abs(ref[k])
    # the PM components cancel heavily: each is held to 1e-6 of the PM total
    for k in ("PM_direct", "PM_reciprocal", "PM_self", "PM_correction"):
        err[k] = abs(e[k] - ref[k]) / abs(ref["PM"])
    vgot = (nb.atom_virial.cpu().numpy().astype(np.float64).sum(axis=0)
            + nb.reciprocal_virial())
    vref = ref["virial_atoms"].sum(axis=0) + ref["virial_recip"]
    err["virial"] = float(np.max(np.abs(vgot - vref)) / np.max(np.abs(vref)))
    print("errors vs FP64:", {k: "%.2e" % v for k, v in err.items()}, flush=True)
    assert err["force"] < FORCE_TOL
    for k, v in err.items():
        if k not in ("force", "virial"):
            assert v <= ENERGY_TOL, (k, v)
    assert err["virial"] <= VIRIAL_TOL
    return err


def test_mixed_93k_nve_evolved_frame_vs_fp64():
    """93k atoms: 200 Langevin steps at 1 fs (tau 0.1 ps), then 1,000 NVE
    leap-frog steps at 1 fs with SETTLE.  The total energy (path potential +
    kinetic, sampled every 100 steps) stays within 0.3% of the kinetic
    energy; the final, dynamically evolved frame meets the force, energy and
    virial tolerances against the FP64 oracle."""
    from sponge_amd import systems
    from sponge_amd.lib import Nonbonded
    s = systems.config_gpu_93k()
    nb = Nonbonded(s, grid=s["pme_grid"], refresh_interval=10)
    nb.set_export(False)
    print("case 93k nve", nb.window_info(), flush=True)
    nb.md_initial(s, temperature=300.0, tau=0.1, seed=20260217, dt=0.001)
    for k in range(200):
        nb.md_step(step=k, forced_nl=(k == 0))
    E, ke_start = [], None
    for k in range(1000):
        e = k % 100 == 0
        if e:
            ke0 = nb.kinetic_energy()
        nb.nve_step(step=k, need_energy=e, forced_nl=(k == 0))
        if e:
            ke = 0.5 * (ke0 + nb.kinetic_energy())
            ke_start = ke if ke_start is None else ke_start
            E.append(nb.potential_energy() + ke)
    E = np.asarray(E)
    print("E_total", E[0], E[-1], "max dev", np.abs(E - E[0]).max(), "KE", ke_start, flush=True)
    assert np.abs(E - E[0]).max() < 3e-3 * ke_start
    assert nb.status() == 0
    s2 = dict(s)
    s2["crd"] = nb.crd.cpu().numpy().copy()
    _frame_vs_fp64(nb, s2)
    nb.close()


def test_served_segments_vs_fp64_and_alone():
    """Eight requests of 1,500-12,000 atoms (cutoff 9 A; the small boxes run
    the plain pair path, the larger ones the windowed path), 100 steps each:
    served together as one step batch and one at a time on the per-segment
    path, every final frame is the same bit for bit, and each one meets the
    tolerances against the FP64 oracle."""
    from sponge_amd import serving
    from sponge_amd.lib import Nonbonded

    def requests():
        return serving.make_requests(count=8, n_min=1500, n_max=12000, rate=40.0, steps=100,
                                     seed=4242)

    serving.pretune(requests())
    mixed = requests()
    out = serving.Server(slots=8, launch="batched").run(mixed, realtime=False)
    alone = requests()
    serving.Server(slots=1, launch="streams").run(alone, realtime=False)
    print(out, flush=True)
    paths = set()
    for a, b in zip(alone, mixed):
        assert a.result["status"] == 0 and b.result["status"] == 0
        np.testing.assert_array_equal(a.result["crd"], b.result["crd"])
        np.testing.assert_array_equal(a.result["vel"], b.result["vel"])
        paths.add(any(b.result["shape"]))
        s = dict(b.system())
        s["crd"] = b.result["crd"]
        nb = Nonbonded(s, cutoff=s["cutoff"], max_atom_in_grid_numbers=serving.MAX_ATOMS_IN_CELL)
        print("case request", b.rid, b.n_atoms, "shape", b.result["shape"], flush=True)
        _frame_vs_fp64(nb, s)
        nb.close()
    assert paths == {False, True}


def test_mixed_93k_neighbor_list_equals_oracle():
    """The 93k system's public half list (row = lower id) equals the oracle's
    cell-list build with the reference's float distance test and exclusion
    rule: the same count in every row and the same entries (order within a
    row aside)."""
    import oracle
    from sponge_amd import systems
    from sponge_amd.lib import Nonbonded
    s = systems.config_gpu_93k()
    nb = Nonbonded(s, grid=(96, 96, 96))
    assert nb.window_info()["windowed"] == 1
    nb.nl_update(forced=True)
    gc, gr = nb.nl_rows()
    assert nb.status() == 0
    oc, orows, _ = oracle.nl_build(s["crd"], s["cell"], s["rcell"], nb.cutoff, 2.0,
                                   (s["excl_start"], s["excl_numbers"], s["excl_list"]),
                                   max_nb=nb.max_nb)
    print("entries", int(gc.sum()), int(oc.sum()), flush=True)
    np.testing.assert_array_equal(gc, oc)
    width = int(gc.max())
    big = np.iinfo(np.int32).max
    cols = np.arange(width)[None, :]
    a = np.where(cols < gc[:, None], gr[:, :width], big)
    b = np.where(cols < oc[:, None], orows[:, :width], big)
    a.sort(axis=1)
    b.sort(axis=1)
    np.testing.assert_array_equal(a, b)
    nb.close()
