"""The path at the size of the one-million-atom water configuration, checked
through properties that hold at any size (within the rounding of the
replicas' larger coordinates).  A 12,346-water box (37,038 atoms,
71.75 A, PME 72^3) is replicated 3 x 3 x 3 into a 1,000,026-atom box of
215.24 A with PME 216^3 (the same grid spacing, so the supercell's charge
grid is the small box's grid tiled): energies are 27x the small box's, each
replica's forces are the small box's forces, every pair and exclusion term
cancels exactly in the int64 accumulators, the list has the mean length of
a homogeneous liquid, and nothing overflows."""
import numpy as np
import pytest

from conftest import rel_rms

pytestmark = pytest.mark.gpu


def _replicate(sysd, m):
    """m x m x m copies of a cubic periodic system (atoms ordered replica by
    replica; exclusion lists shifted with them)."""
    from sponge_amd.native_io import box_to_cell
    L = float(sysd["box"][0])
    n = sysd["n"]
    shifts = np.array([(i, j, k) for i in range(m) for j in range(m) for k in range(m)],
                      np.float64) * L
    crd = (np.asarray(sysd["crd"], np.float64)[None, :, :] + shifts[:, None, :]).reshape(-1, 3)
    r = len(shifts)
    es = np.asarray(sysd["excl_start"], np.int64)
    en = np.asarray(sysd["excl_numbers"], np.int64)
    el = np.asarray(sysd["excl_list"], np.int64)
    tot = int(en.sum())
    out = dict(sysd)
    out.update(
        n=n * r, crd=crd.astype(np.float32),
        charge=np.tile(sysd["charge"], r), lj_type=np.tile(sysd["lj_type"], r),
        mass=np.tile(sysd["mass"], r), n_waters=sysd["n_waters"] * r,
        excl_start=np.concatenate([es + q * tot for q in range(r)]).astype(np.int32),
        excl_numbers=np.tile(en, r).astype(np.int32),
        excl_list=np.concatenate([el + q * n for q in range(r)]).astype(np.int32))
    box = np.array([m * L] * 3, np.float32)
    out["box"] = box
    out["cell"], out["rcell"] = box_to_cell(box)
    return out


def _terms(nb):
    nb.zero_outputs()
    nb.step(need_energy=True, forced_nl=True)
    nb.sync()
    return nb.energies(), nb.frc.cpu().numpy().astype(np.float64)


def test_water_1m_replicated_box_properties():
    from sponge_amd import systems
    from sponge_amd.lib import Nonbonded
    base = systems.tip3p_box(12346, 215.24 / 3.0, seed=9)
    sup = _replicate(base, 3)
    n = sup["n"]
    assert n == 1000026
    # the last cell of each axis takes the remainder of box / cell width, so
    # a corner cell can hold several times the mean ~22 atoms: the per-cell
    # capacity key is raised from its default of 150, as the reference's
    # Check_Overflow would raise it after flagging (to 256, the most the
    # windowed list path takes)
    nbs = Nonbonded(base, grid=(72, 72, 72), max_atom_in_grid_numbers=256)
    e1, f1 = _terms(nbs)
    nbs.close()
    nb = Nonbonded(sup, grid=(216, 216, 216), refresh_interval=20, max_atom_in_grid_numbers=256)
    nb.set_export(False)
    print("window", nb.window_info(), flush=True)
    assert nb.window_info()["windowed"] == 1
    nb.nl_update(forced=True)
    counts = nb.internal_counts()
    rho = n / float(sup["box"][0]) ** 3
    ideal = rho * (2.0 / 3.0) * np.pi * (nb.cutoff + 2.0) ** 3
    print("mean list length", counts.mean(), "ideal", ideal, flush=True)
    assert abs(counts.mean() - ideal) < 0.03 * ideal
    # every pair / exclusion term enters +f and -f: the int64 sums cancel exactly
    nb.excluded(need_energy=False, need_virial=False, frc=False)
    nb.lj_direct(need_energy=False, need_virial=False, frc=False)
    fx = nb.fixed_forces()
    assert np.all(fx.sum(axis=1) == 0), fx.sum(axis=1)
    nb.clear_fixed_forces()
    e27, f27 = _terms(nb)
    assert np.isfinite(f27).all() and nb.status() == 0
    nb.close()
    # a replica's float coordinates are the small box's plus k L, rounded at
    # up to 215 A (half an ulp is 7.6e-6 A): on this unrelaxed lattice the
    # steep LJ contacts turn that into ~1e-6 of the total, hence 1e-5
    print("LJ_short", e27["LJ_short"], 27 * e1["LJ_short"], flush=True)
    assert abs(e27["LJ_short"] - 27 * e1["LJ_short"]) <= 1e-5 * abs(27 * e1["LJ_short"])
    scale = 27 * sum(abs(e1[k]) for k in ("PM_direct", "PM_reciprocal", "PM_self", "PM_correction"))
    for k in ("PM_direct", "PM_reciprocal", "PM_self", "PM_correction", "PM"):
        print(k, e27[k], 27 * e1[k], flush=True)
        assert abs(e27[k] - 27 * e1[k]) <= 1e-5 * scale, k
    err = rel_rms(f27, np.tile(f1, (27, 1)))
    print("replica forces rel RMS", err, flush=True)
    assert err < 5e-5
