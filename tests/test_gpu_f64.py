"""The FP64 evaluation of the path (snb_f64_force, snb_f64_nve_iterate), the
reference the FP32 / int64 path's precision is measured against.

- On the same float coordinates it matches the CPU FP64 oracle with every
  step of the PME taken in double (grid points, spline fractions and the
  influence function's wave vectors; the oracle's default follows the
  kernels' float steps there) and the long-range LJ factor in double: every
  force, energy term (relative to the term, or to 10 kcal/mol when smaller)
  and the total virial within 1e-9 relative (the residue is the 2^-40 /
  2^-32 fixed-point rounding and summation order).
- Its NVE step with SETTLE keeps the rigid waters' geometry to 1e-9 A and
  conserves the total energy.
"""
import numpy as np
import pytest

from conftest import rel_rms

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _f64_frame(nb, crd32):
    import torch
    crd64 = torch.as_tensor(np.asarray(crd32, np.float64)).to(nb.dev).contiguous()
    frc64 = torch.zeros_like(crd64)
    terms, vir = nb.f64_force(crd64, frc64, step=0, forced_nl=True, need_energy=True,
                              need_virial=True)
    if nb.nb14 is None:
        terms.pop("nb14_LJ")
        terms.pop("nb14_EE")
    if nb.bonds is None:
        terms.pop("bond")
    return terms, vir, frc64.cpu().numpy()


@pytest.mark.parametrize("name", ["water_9k", "mixed_small", "gromacs_iter0"])
def test_f64_force_matches_fp64_oracle(request, name):
    import oracle
    from sponge_amd.lib import Nonbonded
    s = request.getfixturevalue(name)
    nb = Nonbonded(s)
    terms, vir, frc = _f64_frame(nb, s["crd"])
    ref = oracle.nonbonded_terms_f64(s, cutoff=nb.cutoff, grid=nb.grid, pme_frac_double=True)
    # the long-range LJ term with the total C6 and its factor in double
    rc = float(np.float32(nb.cutoff))
    ref["LJ_long"] = (oracle.lj_c6_total(s["lj_type"], s["lj_ntypes"], s["lj_b"])
                      * (-2.0 / 3.0 * np.pi / rc ** 3 / 6.0) / ref["volume"])
    ref["LJ"] = ref["LJ_short"] + ref["LJ_long"]
    err = rel_rms(frc, ref["frc"])
    print("case", name, "force rel RMS", err, flush=True)
    for part in ("frc_pair", "frc_excl", "frc_recip", "frc_nb14", "frc_bond"):
        print(part, "rel to total", rel_rms(ref[part], ref["frc"]), flush=True)
    errs = {}
    for k, v in terms.items():
        scale = abs(ref["PM"]) if k.startswith("PM_") else abs(ref[k])
        # (an absolute floor for terms that are nearly zero, such as bonds at r0)
        errs[k] = abs(v - ref[k]) / max(scale, 10.0)
        print(k, v, ref[k], errs[k], flush=True)
    vref = ref["virial_atoms"].sum(axis=0) + ref["virial_recip"]
    verr = np.max(np.abs(vir - vref)) / np.max(np.abs(vref))
    print("virial rel err", verr, flush=True)
    assert err < TOL
    for k, e in errs.items():
        assert e <= TOL, k
    assert verr < TOL
    nb.close()


def test_f64_nve_keeps_constraints_and_energy(water_9k):
    """9,000-atom water box: 200 FP32 Langevin steps at 1 fs, then 500 FP64
    NVE steps at 1 fs from that state; O-H and H-H distances stay at the
    SETTLE lengths to 1e-9 A, the total energy (FP64 potential + kinetic,
    every 50 steps) within 1e-3 of the kinetic energy."""
    import torch
    from sponge_amd import systems
    from sponge_amd.lib import Nonbonded
    s = water_9k
    nb = Nonbonded(s, refresh_interval=10)
    nb.md_initial(s, temperature=300.0, tau=0.1, seed=31, dt=0.001)
    for k in range(200):
        nb.md_step(step=k, forced_nl=(k == 0))
    nb.sync()
    crd = nb.crd.double().contiguous()
    vel = nb.vel.double().contiguous()
    mass = nb.mass.double()
    inv = torch.where(mass > 0, 1.0 / torch.where(mass > 0, mass, torch.ones_like(mass)),
                      torch.zeros_like(mass))
    nb64 = Nonbonded(s, refresh_interval=10)
    nb64.md_initial(s, temperature=300.0, tau=0.1, seed=31, dt=0.001)   # SETTLE triangles
    frc = torch.zeros_like(crd)
    E, ke_start = [], None
    dt = nb.dt_internal
    for k in range(500):
        e = k % 50 == 0
        if e:
            ke0 = 0.5 * float((mass[:, None] * vel * vel).sum())
        frc.zero_()
        terms, _ = nb64.f64_force(crd, frc, step=k, forced_nl=(k == 0), need_energy=e)
        nb64.f64_nve_iterate(mass, inv, frc, vel, crd, dt)
        if e:
            ke1 = 0.5 * float((mass[:, None] * vel * vel).sum())
            ke = 0.5 * (ke0 + ke1)
            ke_start = ke if ke_start is None else ke_start
            E.append(terms["LJ"] + terms["PM"] + ke)
    assert nb64.status() == 0
    E = np.asarray(E)
    print("FP64 E_total", E[0], E[-1], "max dev", np.abs(E - E[0]).max(), "KE", ke_start,
          flush=True)
    assert np.abs(E - E[0]).max() < 1e-3 * ke_start
    x = crd.cpu().numpy()
    a, b, c, rab, rac, rbc = systems.water_triangles(s)
    L = float(s["box"][0])

    def dist(i, j):
        d = x[j] - x[i]
        d -= L * np.round(d / L)
        return np.sqrt((d * d).sum(axis=1))
    assert np.abs(dist(a, b) - rab.astype(np.float64)).max() < 1e-9
    assert np.abs(dist(a, c) - rac.astype(np.float64)).max() < 1e-9
    assert np.abs(dist(b, c) - rbc.astype(np.float64)).max() < 1e-9
    nb.close()
    nb64.close()
