"""Pins of the oracle's integration twins (CPU only): the random-number
generator against published known answers and distribution tests, SETTLE
against the closed-form triangle frame and the constraint geometry, the
Langevin step against the temperature it must produce."""
import numpy as np
import pytest
from scipy import stats

import oracle


def test_philox_random123_known_answers():
    """Known-answer vectors of Philox4x32-10 from the Random123 distribution
    (Salmon et al., SC 2011): counter, key -> output."""
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF),
         (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in cases:
        assert tuple(int(x) for x in oracle.philox4x32_10(ctr, key)) == want


def test_langevin_normals_are_standard_normal():
    """4M draws (1M atoms x 4): mean, variance, fourth moment and a
    Kolmogorov-Smirnov test against N(0, 1); successive stream positions of
    one atom are uncorrelated."""
    n = 1 << 20
    z = oracle.langevin_normals(12345, np.zeros(n, np.uint32)).astype(np.float64)
    x = z.ravel()
    N = x.size
    assert abs(x.mean()) < 5.0 / np.sqrt(N)
    assert abs(x.var() - 1.0) < 5.0 * np.sqrt(2.0 / N)
    assert abs((x ** 4).mean() - 3.0) < 5.0 * np.sqrt(96.0 / N)
    assert stats.kstest(x[:200000], "norm").pvalue > 1e-3
    # tails: P(|z| > 3) = 0.0027
    assert abs((np.abs(x) > 3).mean() - 0.0026998) < 5 * np.sqrt(0.0027 / N)
    z1 = oracle.langevin_normals(12345, np.ones(n, np.uint32))
    r = np.corrcoef(z[:, 0], z1[:, 0])[0, 1]
    assert abs(r) < 5.0 / np.sqrt(n)
    # different seeds, different streams
    z2 = oracle.langevin_normals(54321, np.zeros(n, np.uint32))
    assert abs(np.corrcoef(z[:, 0], z2[:, 0])[0, 1]) < 5.0 / np.sqrt(n)


def test_settle_frame_closed_form_tip3p():
    """Get_Rabcde_From_SSS (settle.cpp:50-98) for TIP3P (O-H 0.9572 A, HOH
    104.52 deg, masses 15.999 / 1.008): with the centre of mass at the origin
    the oxygen sits on the y axis at 2 mH h / M from it (h = r_OH cos(theta/2)),
    the hydrogens at x = -+r_HH/2, y = -(h - ra)."""
    roh, th = 0.9572, np.deg2rad(104.52)
    rhh = 2 * roh * np.sin(th / 2)
    mo, mh = 15.999, 1.008
    g = oracle.settle_frame(roh, roh, rhh, mo, mh, mh)
    h = roh * np.cos(th / 2)
    ra = 2 * mh * h / (mo + 2 * mh)
    np.testing.assert_allclose(g, [ra, -(h - ra), -rhh / 2, rhh / 2, -(h - ra)], atol=2e-6)


def _waters(n_w, seed=3, L=30.0):
    """n_w rigid TIP3P waters at random positions and orientations."""
    from sponge_amd import systems
    rng = np.random.default_rng(seed)
    o = rng.uniform(0, L, (n_w, 3))
    R = systems._random_rotations(rng, n_w)
    geo = systems._water_geometry()
    crd = (o[:, None, :] + np.einsum("nij,kj->nki", R, geo)).reshape(-1, 3).astype(np.float32)
    mass = np.tile([15.999, 1.008, 1.008], n_w).astype(np.float32)
    tri = np.stack([3 * np.arange(n_w), 3 * np.arange(n_w) + 1, 3 * np.arange(n_w) + 2],
                   1).astype(np.int32)
    cell = np.array([L, 0, L, 0, 0, L], np.float32)
    rcell = np.array([1 / L, 0, 1 / L, 0, 0, 1 / L], np.float32)
    return crd, mass, tri, cell, rcell


def test_settle_restores_geometry_and_conserves_momentum():
    """After an unconstrained move, SETTLE puts every triangle back on its
    O-H / O-H / H-H lengths, leaves each water's centre of mass where the
    move put it, and changes no water's momentum."""
    roh, th = 0.9572, np.deg2rad(104.52)
    rhh = 2 * roh * np.sin(th / 2)
    crd, mass, tri, cell, rcell = _waters(500)
    geo = np.tile(oracle.settle_frame(roh, roh, rhh, 15.999, 1.008, 1.008), (500, 1))
    ba, ca = oracle.settle_remember(tri, crd, cell, rcell)
    rng = np.random.default_rng(7)
    moved = (crd + rng.normal(0, 0.02, crd.shape)).astype(np.float32)
    vel = rng.normal(0, 0.1, crd.shape).astype(np.float32)
    m3 = mass.reshape(-1, 3)[:, :, None].astype(np.float64)
    com0 = (m3 * moved.reshape(-1, 3, 3)).sum(1) / m3.sum(1)
    p0 = (m3 * vel.reshape(-1, 3, 3)).sum(1)
    c = moved.copy()
    v = vel.copy()
    dt = np.float32(2.0 * 0.02045)
    oracle.settle_twin(tri, geo, mass, c, cell, rcell, ba, ca, dt, 1.0, 1.0, v)
    w = c.reshape(-1, 3, 3).astype(np.float64)
    d_oh1 = np.linalg.norm(w[:, 1] - w[:, 0], axis=1)
    d_oh2 = np.linalg.norm(w[:, 2] - w[:, 0], axis=1)
    d_hh = np.linalg.norm(w[:, 2] - w[:, 1], axis=1)
    assert np.abs(d_oh1 - roh).max() < 2e-5 and np.abs(d_oh2 - roh).max() < 2e-5
    assert np.abs(d_hh - rhh).max() < 3e-5
    com1 = (m3 * w).sum(1) / m3.sum(1)
    assert np.abs(com1 - com0).max() < 2e-5
    p1 = (m3 * v.reshape(-1, 3, 3)).sum(1)
    assert np.abs(p1 - p0).max() < 1e-4 * np.abs(p0).max()
    # the velocity change is the position change over dt (v_factor = x_factor = 1)
    dv = (v - vel).astype(np.float64)
    dx = (c - moved).astype(np.float64)
    np.testing.assert_allclose(dv, dx / float(dt), atol=2e-3 * np.abs(dv).max())


def test_langevin_thermalises_free_particles():
    """Free particles (no force) under the middle-scheme thermostat reach
    <m v^2> = kB T per component: start at rest, tau 0.1 ps, 600 steps of
    2 fs, then time-average over 200 more."""
    n = 20000
    T, tau, dt_ps = 300.0, 0.1, 0.002
    kb, tu = 0.00198716, 20.455
    mass = np.where(np.arange(n) % 3 == 0, 15.999, 1.008).astype(np.float32)
    dt = np.float32(dt_ps * tu)
    gamma = (1.0 / tau) / tu
    eg = np.float32(np.exp(-gamma * dt))
    sg = np.float32(np.sqrt((1.0 - float(eg) * float(eg)) * T * kb))
    sm = (sg * np.sqrt(1.0 / mass)).astype(np.float32)
    im = (1.0 / mass).astype(np.float32)
    ctr = np.zeros(n, np.uint32)
    vel = np.zeros((n, 3), np.float32)
    crd = np.zeros((n, 3), np.float32)
    frc = np.zeros((n, 3), np.float32)
    acc = []
    for k in range(800):
        oracle.langevin_twin(im, sm, ctr, 99, frc, vel, crd, dt, np.float32(0.5 * dt), eg)
        if k >= 600:
            acc.append((mass[:, None] * vel.astype(np.float64) ** 2).mean() / kb)
    t_kin = float(np.mean(acc))
    assert abs(t_kin - T) < 0.01 * T, t_kin
    assert np.all(ctr == 800)


# ------------------------------------------------------------ pressure assembly
def test_stress_twin_kinetic_part_is_twice_kinetic_energy():
    """Get_Kinetic_to_stress: V * trace(stress) = sum m |v|^2 = 2 E_kin, with
    E_kin summed independently in FP64; the off-diagonal layout of
    Get_Virial_From_Force_Dis(v, v) (a21 = 2 m vx vy ...) on one atom."""
    rng = np.random.default_rng(21)
    n = 50000
    mass = np.where(np.arange(n) % 3 == 0, 15.999, 1.008).astype(np.float32)
    vel = rng.normal(0, 1, (n, 3)).astype(np.float32) / np.sqrt(mass)[:, None]
    inv_v = np.float32(1.0 / 27000.0)
    vt = np.zeros(6, np.float32)
    st = np.zeros(6, np.float32)
    oracle.stress_twin(inv_v, vt, st, vel=vel, mass=mass)
    ek2 = float((mass.astype(np.float64)[:, None] * vel.astype(np.float64) ** 2).sum())
    tr = (float(st[0]) + float(st[2]) + float(st[5])) / float(inv_v)
    assert abs(tr - ek2) < 2e-6 * ek2
    assert np.all(vt == 0)
    v1 = np.array([[0.3, -0.2, 0.7]], np.float32)
    m1 = np.array([2.0], np.float32)
    st = np.zeros(6, np.float32)
    oracle.stress_twin(np.float32(1.0), vt, st, vel=v1, mass=m1)
    x, y, z = (float(c) for c in v1[0])
    want = 2.0 * np.array([x * x, 2 * x * y, y * y, 2 * x * z, 2 * y * z, z * z])
    np.testing.assert_allclose(st, want, rtol=1e-6)


def test_stress_twin_potential_part_is_the_atom_sum():
    """Get_Potential_to_stress: virial_tensor += sum of the atom virials
    (numpy FP64 sum of the same floats), stress += virial_tensor / V on top of
    what the stress held (a SETTLE contribution, say); one large entry (a
    system virial folded into atom 0) is summed exactly."""
    rng = np.random.default_rng(22)
    av = rng.normal(0, 30.0, (20000, 6)).astype(np.float32)
    av[0] = [3.0e5, -1.0e4, 2.5e5, 7.0e3, -2.0e3, 2.8e5]
    vt = np.array([1.0, 0, 2.0, 0, 0, 3.0], np.float32)
    st = np.array([0.5, 0.1, 0.5, 0, 0, 0.5], np.float32)
    vt0, st0 = vt.copy(), st.copy()
    inv_v = np.float32(1.0 / 8.0e5)
    sums = oracle.stress_twin(inv_v, vt, st, atom_virial=av)
    ref = av.astype(np.float64).sum(0)
    np.testing.assert_allclose(sums[:6] * 2.0 ** -32, ref, rtol=0, atol=1e-6 * np.abs(av).max())
    assert np.all(sums[6:] == 0)
    np.testing.assert_allclose(vt, vt0 + ref, rtol=2e-7, atol=1e-3)
    np.testing.assert_allclose(st, st0 + float(inv_v) * (vt0 + ref), rtol=1e-6, atol=1e-9)
