"""GPU parity of the device integrator: leap-frog, middle-scheme Langevin and
SETTLE bit-exact against the oracle's float twins, and a short MD run of the
9,000-atom water box that keeps the waters rigid and the temperature at the
thermostat's."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _nb(sysd, **kw):
    from sponge_amd.lib import Nonbonded
    return Nonbonded(sysd, **kw)


def _host_thermostat_arrays(mass, exp_gamma, temperature):
    """The library's 1/m and noise amplitudes, by the same float operations
    (Middle_Langevin_MD.cpp:226-245)."""
    kb = np.float32(0.00198716)
    eg2 = np.float32(exp_gamma) * np.float32(exp_gamma)
    sg = np.sqrt(np.float32((1.0 - float(eg2)) * float(np.float32(temperature)) * float(kb)))
    inv = np.zeros_like(mass)
    sm = np.zeros_like(mass)
    nz = mass != 0
    inv[nz] = np.float32(1.0) / mass[nz]
    sm[nz] = sg * np.sqrt(np.array([np.float32(1.0 / float(m)) for m in mass[nz]], np.float32))
    return inv, sm


def test_langevin_bit_exact_with_twin(water_9k):
    import oracle
    s = water_9k
    nb = _nb(s)
    nb.md_initial(s, temperature=300.0, tau=1.0, seed=4242, dt=0.002, settle=False)
    rng = np.random.default_rng(5)
    frc = rng.normal(0, 20.0, (s["n"], 3)).astype(np.float32)
    nb.frc.copy_(nb.torch.as_tensor(frc))
    vel = nb.vel.cpu().numpy().copy()
    crd = nb.crd.cpu().numpy().copy()
    inv, sm = _host_thermostat_arrays(np.asarray(s["mass"], np.float32), nb.exp_gamma, 300.0)
    ctr = np.zeros(s["n"], np.uint32)
    for _ in range(3):
        assert nb.L.snb_langevin_iterate(nb.ctx, ctypes.c_void_p(nb.frc.data_ptr()),
                                         ctypes.c_void_p(nb.vel.data_ptr()), None,
                                         ctypes.c_void_p(nb.crd.data_ptr())) == 0
        oracle.langevin_twin(inv, sm, ctr, 4242, frc, vel, crd, np.float32(nb.dt_internal),
                             np.float32(0.5 * nb.dt_internal), np.float32(nb.exp_gamma))
    nb.sync()
    np.testing.assert_array_equal(nb.vel.cpu().numpy(), vel)
    np.testing.assert_array_equal(nb.crd.cpu().numpy(), crd)
    nb.close()


def test_leap_frog_bit_exact_with_twin(water_9k):
    import oracle
    s = water_9k
    nb = _nb(s)
    nb.md_initial(s, settle=False)
    rng = np.random.default_rng(6)
    frc = rng.normal(0, 50.0, (s["n"], 3)).astype(np.float32)
    nb.frc.copy_(nb.torch.as_tensor(frc))
    vel = nb.vel.cpu().numpy().copy()
    crd = nb.crd.cpu().numpy().copy()
    inv = nb.inv_mass.cpu().numpy()
    for vmax in (0.0, 0.5):
        assert nb.L.snb_leap_frog(nb.ctx, ctypes.c_void_p(nb.inv_mass.data_ptr()),
                                  ctypes.c_void_p(nb.frc.data_ptr()),
                                  ctypes.c_void_p(nb.vel.data_ptr()),
                                  ctypes.c_void_p(nb.crd.data_ptr()),
                                  ctypes.c_float(0.04091), ctypes.c_float(vmax)) == 0
        oracle.leap_frog_twin(inv, frc, vel, crd, np.float32(0.04091), np.float32(vmax))
    nb.sync()
    np.testing.assert_array_equal(nb.vel.cpu().numpy(), vel)
    np.testing.assert_array_equal(nb.crd.cpu().numpy(), crd)
    nb.close()


def test_settle_bit_exact_with_twin(water_9k):
    import oracle
    from sponge_amd import systems
    s = water_9k
    nb = _nb(s)
    nb.md_initial(s, settle=True)
    a, b, c, rab, rac, rbc = systems.water_triangles(s)
    tri = np.stack([a, b, c], 1)
    mass = np.asarray(s["mass"], np.float32)
    geo = np.stack([oracle.settle_frame(rab[k], rac[k], rbc[k], mass[a[k]], mass[b[k]],
                                        mass[c[k]]) for k in range(len(a))])
    crd0 = nb.crd.cpu().numpy().copy()
    ba, ca = oracle.settle_remember(tri, crd0, s["cell"], s["rcell"])
    assert nb.L.snb_settle_remember(nb.ctx, ctypes.c_void_p(nb.crd.data_ptr()), nb.cell,
                                    nb.rcell) == 0
    rng = np.random.default_rng(8)
    moved = (crd0 + rng.normal(0, 0.02, crd0.shape)).astype(np.float32)
    vel = nb.vel.cpu().numpy().copy()
    nb.crd.copy_(nb.torch.as_tensor(moved))
    nb.stress.zero_()
    assert nb.L.snb_settle_do(nb.ctx, ctypes.c_void_p(nb.mass.data_ptr()),
                              ctypes.c_void_p(nb.crd.data_ptr()), nb.cell, nb.rcell,
                              ctypes.c_void_p(nb.vel.data_ptr()), 1,
                              ctypes.c_void_p(nb.stress.data_ptr()),
                              ctypes.c_float(nb.dt_internal), ctypes.c_float(nb.v_factor),
                              ctypes.c_float(nb.x_factor)) == 0
    vir = oracle.settle_twin(tri, geo, mass, moved, s["cell"], s["rcell"], ba, ca,
                             np.float32(nb.dt_internal), np.float32(nb.v_factor),
                             np.float32(nb.x_factor), vel)
    nb.sync()
    np.testing.assert_array_equal(nb.crd.cpu().numpy(), moved)
    np.testing.assert_array_equal(nb.vel.cpu().numpy(), vel)
    inv_v = np.float32(np.float32(s["rcell"][0]) * np.float32(s["rcell"][2])) * \
        np.float32(s["rcell"][5])
    want = np.array([inv_v * np.float32(float(v) * 2.0 ** -32) for v in vir], np.float32)
    np.testing.assert_array_equal(nb.stress.cpu().numpy(), want)
    nb.close()


def test_settle_pair_bit_exact_with_twin(water_9k):
    """Rigid two-atom pairs (settle_pair, settle.cpp:258-307): one O-H pair per
    water of the 9k box (no triangles), remembered, moved at random, then
    constrained on the device and by the float twin: coordinates, velocities
    and the stress from the pair virial agree bit for bit; the fused
    iteration refuses a context with pairs."""
    import oracle
    from sponge_amd import systems
    from sponge_amd.lib import SnbError, check
    s = water_9k
    nb = _nb(s)
    nb.md_initial(s, settle=False)
    a, b, c, rab, rac, rbc = systems.water_triangles(s)
    pairs = np.stack([a, b], 1).astype(np.int32)
    r0 = np.asarray(rab, np.float32)
    mass = np.asarray(s["mass"], np.float32)
    pi = np.ascontiguousarray(pairs[:, 0])
    pj = np.ascontiguousarray(pairs[:, 1])
    check(nb.L.snb_settle_pair_initial(nb.ctx, len(pairs), pi.ctypes.data, pj.ctypes.data,
                                       r0.ctypes.data, mass.ctypes.data),
          "snb_settle_pair_initial")
    ck = (np.float32(1.0) / (mass[pi] + mass[pj])).astype(np.float32)
    crd0 = nb.crd.cpu().numpy().copy()
    ab = oracle.settle_pair_remember(pairs, crd0, s["cell"], s["rcell"])
    assert nb.L.snb_settle_remember(nb.ctx, ctypes.c_void_p(nb.crd.data_ptr()), nb.cell,
                                    nb.rcell) == 0
    rng = np.random.default_rng(9)
    moved = (crd0 + rng.normal(0, 0.02, crd0.shape)).astype(np.float32)
    vel = nb.vel.cpu().numpy().copy()
    nb.crd.copy_(nb.torch.as_tensor(moved))
    nb.stress.zero_()
    assert nb.L.snb_settle_do(nb.ctx, ctypes.c_void_p(nb.mass.data_ptr()),
                              ctypes.c_void_p(nb.crd.data_ptr()), nb.cell, nb.rcell,
                              ctypes.c_void_p(nb.vel.data_ptr()), 1,
                              ctypes.c_void_p(nb.stress.data_ptr()),
                              ctypes.c_float(nb.dt_internal), ctypes.c_float(nb.v_factor),
                              ctypes.c_float(nb.x_factor)) == 0
    vir = oracle.settle_pair_twin(pairs, r0, ck, mass, moved, s["cell"], s["rcell"], ab,
                                  np.float32(nb.dt_internal), np.float32(nb.v_factor),
                                  np.float32(nb.x_factor), vel)
    nb.sync()
    np.testing.assert_array_equal(nb.crd.cpu().numpy(), moved)
    np.testing.assert_array_equal(nb.vel.cpu().numpy(), vel)
    inv_v = np.float32(np.float32(s["rcell"][0]) * np.float32(s["rcell"][2])) * \
        np.float32(s["rcell"][5])
    want = np.array([inv_v * np.float32(float(v) * 2.0 ** -32) for v in vir], np.float32)
    np.testing.assert_array_equal(nb.stress.cpu().numpy(), want)
    with pytest.raises(SnbError) as e:
        check(nb.L.snb_md_iterate_fused(nb.ctx, ctypes.c_void_p(nb.mass.data_ptr()),
                                        ctypes.c_void_p(nb.crd.data_ptr()), nb.cell, nb.rcell,
                                        ctypes.c_void_p(nb.vel.data_ptr()), 0, None),
              "snb_md_iterate_fused")
    assert e.value.code == 4
    nb.close()


def _t_kin(nb, s):
    v = nb.vel.cpu().numpy().astype(np.float64)
    m = np.asarray(s["mass"], np.float64)
    ek = 0.5 * (m[:, None] * v * v).sum()
    dof = 6 * (s["n"] // 3) - 3
    return 2 * ek / (dof * 0.00198716)


def test_md_run_keeps_waters_rigid_and_temperature(water_9k):
    """1,500 steps of 2 fs (forces, Langevin at 300 K with tau 0.1 ps, SETTLE;
    the list rebuilt every 10 steps) from Maxwell velocities at 300 K on the
    unrelaxed lattice: every water keeps its geometry to 1e-4 A, and once the
    relaxation heat is gone (last 1 ps, sampled every 20 steps) the kinetic
    temperature of the rigid waters (6 degrees of freedom each) averages
    300 K within 3%; nothing overflows."""
    from sponge_amd import systems
    s = water_9k
    nb = _nb(s, refresh_interval=10)
    nb.md_initial(s, temperature=300.0, tau=0.1, seed=77, dt=0.002)
    temps = []
    for k in range(1500):
        nb.md_step(step=k, forced_nl=(k == 0))
        if k >= 1000 and k % 20 == 0:
            temps.append(_t_kin(nb, s))
        if k % 500 == 0:
            print("step", k, "T_kin", _t_kin(nb, s), flush=True)
    nb.sync()
    print("status", nb.status(), flush=True)
    assert nb.status() == 0
    c = nb.crd.cpu().numpy().astype(np.float64).reshape(-1, 3, 3)
    L = np.asarray(s["box"], np.float64)

    def dist(u, v):
        d = u - v
        d -= np.rint(d / L) * L
        return np.linalg.norm(d, axis=1)

    a, b, cc, rab, rac, rbc = systems.water_triangles(s)
    assert np.abs(dist(c[:, 1], c[:, 0]) - rab[0]).max() < 1e-4
    assert np.abs(dist(c[:, 2], c[:, 0]) - rac[0]).max() < 1e-4
    assert np.abs(dist(c[:, 2], c[:, 1]) - rbc[0]).max() < 1e-4
    t_mean = float(np.mean(temps))
    print("T_kin mean over the last 1 ps", t_mean, flush=True)
    assert abs(t_mean - 300.0) < 9.0
    nb.close()


def test_pressure_assembly_bit_exact_with_twin(water_9k):
    """snb_get_pressure / snb_get_potential_to_stress / snb_get_kinetic_to_stress
    against the oracle's twin: random atom virials (one system-sized entry on
    atom 0), velocities and masses; the stress and virial tensor start from
    non-zero values, as after SETTLE; two calls in a row (the sums restart)."""
    import oracle
    from sponge_amd.lib import LTMatrix3, check
    s = water_9k
    nb = _nb(s)
    torch = nb.torch
    n = s["n"]
    rng = np.random.default_rng(31)
    av = rng.normal(0, 30.0, (n, 6)).astype(np.float32)
    av[0] = [3.0e5, -1.0e4, 2.5e5, 7.0e3, -2.0e3, 2.8e5]
    mass = np.asarray(s["mass"], np.float32)
    vel = (rng.normal(0, 1, (n, 3)) / np.sqrt(mass)[:, None]).astype(np.float32)
    d_av = torch.as_tensor(av).to(nb.dev)
    d_vel = torch.as_tensor(vel).to(nb.dev)
    d_mass = torch.as_tensor(mass).to(nb.dev)
    st0 = np.array([0.5, 0.1, 0.5, -0.2, 0.05, 0.5], np.float32)
    vt0 = np.array([1.0, 0.0, 2.0, 0.0, 0.0, 3.0], np.float32)
    d_st = torch.as_tensor(st0).to(nb.dev)
    d_vt = torch.as_tensor(vt0).to(nb.dev)
    cell = s["cell"]
    inv_v = np.float32(1.0) / ((np.float32(cell[0]) * np.float32(cell[2])) * np.float32(cell[5]))
    st, vt = st0.copy(), vt0.copy()
    P = ctypes.c_void_p
    for rep in range(2):
        hs, hp = LTMatrix3(), ctypes.c_float()
        check(nb.L.snb_get_pressure(nb.ctx, n, P(d_vel.data_ptr()), P(d_mass.data_ptr()),
                                    P(d_av.data_ptr()), nb.cell, P(d_vt.data_ptr()),
                                    P(d_st.data_ptr()), ctypes.byref(hs), ctypes.byref(hp)),
              "snb_get_pressure")
        oracle.stress_twin(inv_v, vt, st, atom_virial=av, vel=vel, mass=mass)
        got = np.array([hs.a11, hs.a21, hs.a22, hs.a31, hs.a32, hs.a33], np.float32)
        print("rep", rep, "stress", got, "pressure", hp.value, flush=True)
        np.testing.assert_array_equal(got, st)
        np.testing.assert_array_equal(d_st.cpu().numpy(), st)
        np.testing.assert_array_equal(d_vt.cpu().numpy(), vt)
        assert hp.value == np.float32((st[0] + st[2]) + st[5]) / np.float32(3.0)
    # the two halves on their own
    check(nb.L.snb_get_potential_to_stress(nb.ctx, n, P(d_av.data_ptr()), nb.cell,
                                           P(d_vt.data_ptr()), P(d_st.data_ptr())),
          "snb_get_potential_to_stress")
    oracle.stress_twin(inv_v, vt, st, atom_virial=av)
    check(nb.L.snb_get_kinetic_to_stress(nb.ctx, n, P(d_vel.data_ptr()), P(d_mass.data_ptr()),
                                         nb.cell, P(d_st.data_ptr())),
          "snb_get_kinetic_to_stress")
    oracle.stress_twin(inv_v, vt, st, vel=vel, mass=mass)
    nb.sync()
    np.testing.assert_array_equal(d_st.cpu().numpy(), st)
    np.testing.assert_array_equal(d_vt.cpu().numpy(), vt)
    nb.close()


def test_md_pressure_steps_match_eager_assembly(water_9k):
    """A pressure step of md_step (virials zeroed, force path with virials,
    Langevin + SETTLE with the constraint virial, Get_pressure on the device)
    gives the stress that the twin assembles from the step's own atom
    virials, reciprocal virial, velocities and SETTLE part; the pressure is
    finite."""
    import oracle
    s = water_9k
    nb = _nb(s, refresh_interval=10)
    nb.md_initial(s, temperature=300.0, tau=0.1, seed=5, dt=0.002)
    for k in range(30):
        nb.md_step(step=k, forced_nl=(k == 0))
    # the same pressure step, taken apart: forces + virials, iteration, then assembly
    nb.frc.zero_()
    nb.atom_virial.zero_()
    nb.virial.zero_()
    nb.stress.zero_()
    nb.virial_tensor.zero_()
    nb.step(step=30, need_virial=True, fused=True)
    nb.md_iterate(need_pressure=True)
    nb.sync()
    av = nb.atom_virial.cpu().numpy().copy()
    rec = nb.virial.cpu().numpy().copy()
    st = nb.stress.cpu().numpy().copy()          # SETTLE's part
    vel = nb.vel.cpu().numpy().copy()
    hs, p = nb.get_pressure(copy=True)
    av[0] += rec
    cell = s["cell"]
    inv_v = np.float32(1.0) / ((np.float32(cell[0]) * np.float32(cell[2])) * np.float32(cell[5]))
    vt = np.zeros(6, np.float32)
    oracle.stress_twin(inv_v, vt, st, atom_virial=av, vel=vel,
                       mass=np.asarray(s["mass"], np.float32))
    np.testing.assert_array_equal(hs, st)
    print("pressure (bar)", p * 6.946827162543585e4, flush=True)
    assert np.isfinite(p)
    nb.close()


def test_nve_conserves_energy_and_final_frame_matches_fp64(water_9k):
    """Precision: after 300 Langevin steps at 1 fs, 1,000 NVE leap-frog steps
    with SETTLE keep the total energy (path potential + kinetic, sampled
    every 50 steps) within 0.3% of the kinetic energy of its start, and the
    fitted change over the 1 ps below 2e-3 kT per degree of freedom (the
    long-run drift per ns is the bench's 20 ps measurement); on the final
    (dynamically evolved) frame the forces are within 5e-5 relative RMS and
    the LJ / PM energies within 1e-6 of the FP64 oracle."""
    import oracle
    from conftest import rel_rms
    s = water_9k
    nb = _nb(s, refresh_interval=10)
    nb.md_initial(s, temperature=300.0, tau=0.1, seed=11, dt=0.001)
    for k in range(300):
        nb.md_step(step=k, forced_nl=(k == 0))
    t, E = [], []
    ke_start = None
    for k in range(1000):
        e = k % 50 == 0
        if e:
            ke0 = nb.kinetic_energy()
        nb.nve_step(step=k, need_energy=e, forced_nl=(k == 0))
        if e:
            ke1 = nb.kinetic_energy()
            ke = 0.5 * (ke0 + ke1)
            ke_start = ke if ke_start is None else ke_start
            t.append(k * 0.001)
            E.append(nb.potential_energy() + ke)
    E = np.asarray(E)
    dof = 6 * (s["n"] // 3) - 3
    slope = np.polyfit(np.asarray(t), E, 1)[0]
    change_kt = slope * 1.0 / dof / (0.00198716 * 300.0)     # over the 1 ps
    print("E_total", E[0], E[-1], "std", E.std(), "change kT/dof over 1 ps", change_kt,
          flush=True)
    assert np.abs(E - E[0]).max() < 3e-3 * ke_start
    assert abs(change_kt) < 2e-3
    assert nb.status() == 0
    # the final frame against the FP64 oracle
    s2 = dict(s)
    s2["crd"] = nb.crd.cpu().numpy().copy()
    nb.zero_outputs()
    nb.step(step=0, need_energy=True, need_virial=True, forced_nl=True, fused=True)
    nb.sync()
    e = nb.energies()
    ref = oracle.nonbonded_terms_f64(s2, cutoff=nb.cutoff, grid=nb.grid)
    err = rel_rms(nb.frc.cpu().numpy().astype(np.float64), ref["frc"])
    print("final frame force rel RMS", err, flush=True)
    assert err < 5e-5
    assert abs(e["LJ_short"] - ref["LJ_short"]) <= 1e-6 * abs(ref["LJ_short"])
    assert abs(e["PM"] - ref["PM"]) <= 1e-6 * abs(ref["PM"])
    nb.close()


def test_fused_iteration_bit_identical_to_separate_calls(mixed_small):
    """snb_md_iterate_fused (forces straight from the fixed-point accumulators,
    one remember + Langevin + SETTLE pass over triangles and free atoms)
    against the flush + snb_settle_remember + snb_langevin_iterate +
    snb_settle_do sequence: 25 steps of the mixed system (waters and bead
    chains) with a pressure step, same seed; coordinates, velocities and the
    stress agree bit for bit."""
    s = mixed_small
    res = []
    for fused in (False, True):
        nb = _nb(s, refresh_interval=10)
        nb.md_initial(s, temperature=300.0, tau=0.1, seed=21, dt=0.002)
        for k in range(25):
            nb.md_step(step=k, forced_nl=(k == 0), need_virial=(k == 12), fused_iterate=fused)
        nb.sync()
        res.append((nb.crd.cpu().numpy().copy(), nb.vel.cpu().numpy().copy(),
                    nb.stress.cpu().numpy().copy(), nb.status()))
        nb.close()
    print("stress", res[0][2], res[1][2], flush=True)
    assert res[0][3] == 0 and res[1][3] == 0
    np.testing.assert_array_equal(res[0][0], res[1][0])
    np.testing.assert_array_equal(res[0][1], res[1][1])
    np.testing.assert_array_equal(res[0][2], res[1][2])
