"""Pins of the CPU oracle (oracle/snb_oracle.c), CPU only.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a sign, an index, a transposed operand) fails it: known answers printed
by the reference for its own fixtures, textbook constants, closed forms and
brute force, and the float twin against the FP64 layer.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy.special import erfc as sp_erfc

import oracle
from conftest import GOLDEN, gromacs_case, rel_rms


# --------------------------------------------------------------------- beta
@pytest.mark.parametrize("rc,beta", [(8.0, 0.348639), (12.0, 0.226635), (10.0, 0.275106),
                                     (9.0, 0.307676)])
def test_beta_known_values(rc, beta):
    """PM_force.cpp:244-272: reference values (rc 8 and 12 printed by the
    reference's CPU build, rc 9 and 10 from a float emulation of it)."""
    assert abs(oracle.get_beta(rc, 1e-5) - beta) < 2e-6


@pytest.mark.parametrize("rc", [8.0, 9.0, 10.0, 12.0])
def test_beta_solves_tolerance_equation(rc):
    b = oracle.get_beta(rc, 1e-5)
    assert abs(sp_erfc(b * rc) / rc - 1e-5) < 1e-9


# ----------------------------------------------------------------- grid size
@pytest.mark.parametrize("length,size", [(44.8, 48), (96.5, 100), (215.24, 216), (24.64, 28),
                                         (79.3, 80), (70.0, 72), (21.8359051, 24),
                                         (23.3691196, 24), (19.9797707, 20), (62.0, 64),
                                         (125.0, 128)])
def test_fft_size_reference_rule(length, size):
    """common.cpp:639-659 (values the reference printed: 72^3 for the 70 A
    GROMACS box, 24x24x20 for the TIP3P box)."""
    n = oracle.fft_size(length)
    assert n == size
    assert oracle.check_2357(n) and n % 4 == 0


def test_check_2357():
    assert [oracle.check_2357(x) for x in (1, 2, 49, 210, 11, 22, 26)] == \
        [True, True, True, True, False, False, False]


# ------------------------------------------------------- B-spline machinery
def test_bspline_moduli_closed_form():
    """order 4: b(0) = 1 (the weights sum to 1), b(K/2) = 9 (|1/6 - 2/3 + 1/6| = 1/3)."""
    b = oracle.bspline_moduli(48)
    assert abs(b[0] - 1.0) < 1e-14
    assert abs(b[24] - 9.0) < 1e-12
    m = np.arange(48)
    w = np.exp(2j * np.pi * np.outer(m, np.arange(3)) / 48) @ np.array([1 / 6, 2 / 3, 1 / 6])
    np.testing.assert_allclose(b, 1.0 / np.abs(w) ** 2, rtol=1e-13)


def test_bc_zero_mode_and_symmetry():
    rcell = np.array([1 / 30.0, 0, 1 / 30.0, 0, 0, 1 / 30.0], np.float32)
    bc = oracle.bc_table(24, 24, 24, rcell, 0.3, 27000.0)
    assert bc[0, 0, 0] == 0.0
    assert np.all(bc.ravel()[1:] > 0)
    np.testing.assert_allclose(bc[1, 2, 3], bc[23, 22, 3], rtol=1e-13)  # k and -k (kz kept)


# ------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(4, 6, 5), (8, 3, 7), (6, 10, 9)])
def test_fft_matches_naive_dft(shape):
    rng = np.random.default_rng(3)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    np.testing.assert_allclose(oracle.fft3d(x), oracle.dft3d_naive(x), atol=1e-10)
    np.testing.assert_allclose(oracle.fft3d(x, +1), oracle.dft3d_naive(x, +1), atol=1e-10)


# ----------------------------------------------------------------- Ewald
def _rock_salt(cells, a=1.0):
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    na = base
    cl = base + np.array([0.5, 0, 0])
    pos, q = [], []
    for i in range(cells):
        for j in range(cells):
            for k in range(cells):
                off = np.array([i, j, k], float)
                pos += list((na + off) * a)
                q += [1.0] * 4
                pos += list((cl + off) * a)
                q += [-1.0] * 4
    return np.array(pos), np.array(q), cells * a


def test_ewald_madelung_constant_rock_salt():
    """Textbook: E per ion pair = -M q^2 / r0 with M = 1.747565 (r0 = a/2)."""
    pos, q, L = _rock_salt(3, a=2.0)
    _, ed, er, es = oracle.ewald_f64(pos, q, np.array([L, L, L]), beta=1.5, rcut=2.99, kmax=14)
    e = ed + er + es
    m = -e / (len(q) / 2) * 1.0   # r0 = 1
    assert abs(m - 1.747565) < 1e-5


def test_pme_f64_matches_exact_ewald():
    """The FP64 PME reciprocal + direct + self reproduces the exact Ewald sum of
    a random neutral box (no exclusions), energies and forces."""
    rng = np.random.default_rng(5)
    n, L = 60, 20.0
    pos = rng.uniform(0, L, (n, 3))
    q = rng.normal(0, 1, n)
    q -= q.mean()
    beta, rc = 0.35, 9.5
    f_ew, ed, er, es = oracle.ewald_f64(pos, q, np.array([L, L, L]), beta, rc, kmax=14)
    rcell = np.array([1 / L, 0, 1 / L, 0, 0, 1 / L], np.float32)
    crd = pos.astype(np.float32)
    qf = q.astype(np.float32)
    f_rec, e_rec, _, _, _ = oracle.pme_f64(crd, qf, rcell, (48, 48, 48), beta, L ** 3)
    # 4th-order B-splines on a 0.42 A grid: ~1e-4 relative reciprocal error
    assert abs(e_rec - er) < 2e-3 * abs(er) + 1e-3
    # reciprocal forces: Ewald total minus its direct part
    f_dir = np.zeros((n, 3))
    for i in range(n):
        for j in range(n):
            if i == j:
                continue
            for a in (-1, 0, 1):
                for b in (-1, 0, 1):
                    for c in (-1, 0, 1):
                        d = pos[j] - pos[i] + np.array([a, b, c]) * L
                        r = np.linalg.norm(d)
                        if r >= rc:
                            continue
                        x = beta * r
                        fa = -q[i] * q[j] * (2 / math.sqrt(math.pi) * x * math.exp(-x * x) +
                                             math.erfc(x)) / r ** 3
                        f_dir[i] += fa * d
    assert rel_rms(f_rec, f_ew - f_dir) < 2e-3


def test_pme_translation_invariance_and_zero_net_force():
    rng = np.random.default_rng(9)
    n, L = 80, 24.0
    pos = rng.uniform(0, L, (n, 3)).astype(np.float32)
    q = rng.normal(0, 1, n).astype(np.float32)
    q -= q.mean()
    rcell = np.array([1 / L, 0, 1 / L, 0, 0, 1 / L], np.float32)
    f0, e0, _, _, _ = oracle.pme_f64(pos, q, rcell, (24, 24, 24), 0.3, L ** 3)
    f1, e1, _, _, _ = oracle.pme_f64((pos + np.float32(L / 24 * 5)).astype(np.float32), q,
                                     rcell, (24, 24, 24), 0.3, L ** 3)
    assert abs(e1 - e0) < 1e-6 * abs(e0)
    assert rel_rms(f1, f0) < 1e-4
    # smooth PME does not conserve momentum exactly (the B-spline gather is not
    # pairwise antisymmetric); the net force stays a small fraction
    assert np.abs(f0.sum(0)).max() < 1e-2 * np.abs(f0).max()


def test_bspline_partition_of_unity():
    """sum_k w_k(t) = 1 and sum_k w_k'(t) = 0 -> a uniform charge sheet gives a
    spread grid summing to the total charge."""
    L = 16.0
    rcell = np.array([1 / L, 0, 1 / L, 0, 0, 1 / L], np.float32)
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, L, (30, 3)).astype(np.float32)
    q = rng.normal(0, 1, 30).astype(np.float32)
    _, _, _, Q, _ = oracle.pme_f64(pos, q, rcell, (16, 16, 16), 0.3, L ** 3, want_grids=True)
    assert abs(Q.sum() - q.astype(np.float64).sum()) < 1e-5


# ------------------------------------------------------------ pair math
def _two_atoms(r, qa=1.3, qb=-0.7, sig=3.2, eps=0.2, L=40.0):
    crd = np.array([[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]], np.float32)
    a = np.array([4 * eps * sig ** 12 * 12], np.float32)
    b = np.array([4 * eps * sig ** 6 * 6], np.float32)
    cell = np.array([L, 0, L, 0, 0, L], np.float32)
    rcell = np.array([1 / L, 0, 1 / L, 0, 0, 1 / L], np.float32)
    return crd, np.array([qa, qb], np.float32), np.zeros(2, np.int32), a, b, cell, rcell


@pytest.mark.parametrize("r", [2.5, 3.5, 5.0, 9.0])
def test_pair_two_atom_closed_form(r):
    """Lennard_Jones_force.h:43-73: E = A/12 r^-12 - B/6 r^-6 + qq erfc(br)/r and
    F_i = -dE/dr_i, checked against a central difference of the closed form."""
    crd, q, t, A, B, cell, rcell = _two_atoms(r)
    beta = 0.3
    counts = np.array([1, 0], np.int32)
    rows = np.array([[1], [0]], np.int32)
    f, elj, ec, vir = oracle.pair_f64(crd, q, t, A, B, cell, rcell, 10.0, beta, counts, rows)
    rr = float(np.float32(5.0 + r) - np.float32(5.0))
    qq = float(q[0]) * float(q[1])

    def energy(x):
        return float(A[0]) / 12 * x ** -12 - float(B[0]) / 6 * x ** -6 + qq * math.erfc(beta * x) / x
    de = (energy(rr + 1e-6) - energy(rr - 1e-6)) / 2e-6
    assert abs(elj[0] + ec[0] - energy(rr)) < 1e-8 * max(1, abs(energy(rr)))
    # force on atom 0 along +x is +dE/dr (moving atom 0 right shortens r)
    assert abs(f[0, 0] - de) < 1e-5 * max(1.0, abs(de))
    assert abs(f[0, 0] + f[1, 0]) < 1e-12
    # virial (reference convention -f (x) dr): a11 = -f0x * (x1 - x0)
    assert abs(vir[0, 0] - (-f[0, 0] * rr)) < 1e-9 * max(1, abs(vir[0, 0]))


def test_pair_cutoff_excludes_far_pairs():
    crd, q, t, A, B, cell, rcell = _two_atoms(10.5)
    f, elj, ec, _ = oracle.pair_f64(crd, q, t, A, B, cell, rcell, 10.0, 0.3,
                                    np.array([1, 0], np.int32), np.array([[1], [0]], np.int32))
    assert np.all(f == 0) and elj[0] == 0 and ec[0] == 0


def test_pair_minimum_image_across_boundary():
    L = 30.0
    crd = np.array([[0.5, 1.0, 1.0], [29.0, 1.0, 1.0]], np.float32)   # 1.5 A apart via the wall
    _, q, t, A, B, cell, rcell = _two_atoms(1.0, L=L)
    f, elj, _, _ = oracle.pair_f64(crd, q, t, A, B, cell, rcell, 10.0, 0.3,
                                   np.array([1, 0], np.int32), np.array([[1], [0]], np.int32))
    crd2 = np.array([[0.5, 1.0, 1.0], [-1.0, 1.0, 1.0]], np.float32)
    f2, elj2, _, _ = oracle.pair_f64(crd2, q, t, A, B, cell, rcell, 10.0, 0.3,
                                     np.array([1, 0], np.int32), np.array([[1], [0]], np.int32))
    np.testing.assert_allclose(f, f2, rtol=1e-6)
    assert abs(elj[0] - elj2[0]) < 1e-6 * abs(elj[0])


# --------------------------------------------------------- exclusion term
def test_exclusion_plus_direct_equals_bare_coulomb():
    """E_excl(pair) + E_direct(pair, no cutoff) = -qq erf/r + qq erfc/r: for an
    excluded pair the correction removes exactly the erf part, so
    E_excl + E_direct + qq/r * ... : (-qq erf(x) + qq erfc(x))/r + qq erf(x)/r*2 = qq/r."""
    crd, q, t, A, B, cell, rcell = _two_atoms(1.6)
    A[:] = 0
    B[:] = 0
    beta = 0.3
    excl = (np.array([0, 1], np.int32), np.array([1, 0], np.int32), np.array([1], np.int32))
    f_ex, e_ex, _ = oracle.excl_f64(crd, q, cell, rcell, beta, excl)
    f_d, _, ec, _ = oracle.pair_f64(crd, q, t, A, B, cell, rcell, 10.0, beta,
                                    np.array([1, 0], np.int32), np.array([[1], [0]], np.int32))
    r = float(np.float32(6.6) - np.float32(5.0))
    qq = float(q[0]) * float(q[1])
    beta = float(np.float32(beta))   # the oracle takes beta as float
    assert abs((e_ex[0] + ec[0]) - qq * (math.erfc(beta * r) - math.erf(beta * r)) / r) < 1e-10
    # forces on atom 0 along x: the direct (erfc) part plus the erf part that the
    # reciprocal sum carries is the bare Coulomb force, and the exclusion
    # correction removes exactly that erf part
    x = beta * r
    f_erf = -qq * (math.erf(x) - 2 / math.sqrt(math.pi) * x * math.exp(-x * x)) / r ** 2
    f_bare = -qq / r ** 2
    assert abs((f_d[0, 0] + f_erf) - f_bare) < 1e-9 * abs(f_bare)
    assert abs(f_ex[0, 0] + f_erf) < 1e-9 * abs(f_bare)


# ------------------------------------------------------- neighbor list
def test_nl_brute_force_invariants(small_water):
    s = small_water
    excl = (s["excl_start"], s["excl_numbers"], s["excl_list"])
    counts, rows, ov = oracle.nl_build(s["crd"], s["cell"], s["rcell"], 8.0, 2.0, excl)
    assert not ov
    L = s["box"][0]
    for i in range(s["n"]):
        r = rows[i, :counts[i]]
        assert np.all(r > i) and len(np.unique(r)) == len(r)
        d = s["crd"][r] - s["crd"][i]
        d -= np.round(d / L) * L
        assert np.all(np.einsum("ij,ij->i", d, d) < 100.0 + 1e-3)
    # exclusions removed: water i's H's are never in the O row
    for w in range(0, s["n"], 3):
        assert not np.isin([w + 1, w + 2], rows[w, :counts[w]]).any()


def test_nl_overflow_truncates(small_water):
    s = small_water
    counts, rows, ov = oracle.nl_build(s["crd"], s["cell"], s["rcell"], 8.0, 2.0, None, max_nb=8)
    assert ov and counts.max() == 8


# ---------------------------------------------------- twin vs FP64 layer
def test_twin_math_error_bounds():
    r2 = np.geomspace(0.5, 200.0, 400).astype(np.float32)
    gr = np.array([lib_call(oracle.lib().orc_twin_rsqrtf, v) for v in r2])
    assert np.max(np.abs(gr * np.sqrt(r2.astype(np.float64)) - 1)) < 3e-7
    # FP64 pieces of the energy / exclusion path
    import ctypes
    import mpmath
    L = oracle.lib()
    mpmath.mp.dps = 30
    for xv in np.concatenate([np.linspace(0.0, 6.5, 521), [0.5, 1.0, 5.999999]]):
        c_, g_ = ctypes.c_double(), ctypes.c_double()
        L.orc_twin_erfc_parts(float(xv), ctypes.byref(c_), ctypes.byref(g_))
        ref_c = float(mpmath.erfc(mpmath.mpf(float(xv))))
        ref_g = float(2 / mpmath.sqrt(mpmath.pi) * xv * mpmath.exp(-mpmath.mpf(float(xv)) ** 2))
        if xv < 6.0:
            # the rounding of x^2 alone moves e^{-x^2} by up to x^2 ulps
            assert abs(c_.value - ref_c) <= (3e-15 + 2.5e-16 * xv * xv) * ref_c, (xv, c_.value)
        else:
            assert c_.value == 0.0 and ref_c < 2.2e-17
        assert abs(g_.value - ref_g) <= (2e-15 + 2.5e-16 * xv * xv) * ref_g + 1e-300
        # erf - g of the exclusion correction, formed from the two
        if 0.3 <= xv < 6.0:
            ref_h = float(mpmath.erf(mpmath.mpf(float(xv)))) - ref_g
            h = (1.0 - c_.value) - g_.value
            # 1 - erfc carries an absolute rounding of an ulp of 1
            assert abs(h - ref_h) <= 2e-15 * ref_h + 4.5e-16, (xv, h, ref_h)
    for yv in np.linspace(-40.0, 0.0, 161):
        assert abs(L.orc_twin_exp_d(float(yv)) - np.exp(yv)) <= 3e-16 * np.exp(yv)


@pytest.mark.parametrize("beta,rc", [(0.3123, 10.0), (0.3472, 8.0), (0.26, 12.0), (0.45, 10.0)])
def test_twin_coulomb_table_against_erfc(beta, rc):
    # h(x) = erfc(x) + 2/sqrt(pi) x e^{-x^2} of the pair kernels' Hermite
    # table against scipy in double, over the whole tabulated range (x a bit
    # past beta rc included): the float coefficients bound the error near
    # h ~ 1 (absolute 1e-7); the relative error in the tail stays small too
    L = oracle.lib()
    x = np.linspace(0.0, beta * rc * 1.0078, 4001).astype(np.float32)
    got = np.array([L.orc_twin_coul_h(float(v), beta, rc) for v in x])
    xd = x.astype(np.float64)
    ref = sp_erfc(xd) + 2.0 / np.sqrt(np.pi) * xd * np.exp(-xd * xd)
    assert np.max(np.abs(got - ref)) < 1.5e-7
    assert np.max(np.abs(got - ref) / ref) < 1e-5
    # nodes reproduce h itself; a planted sign error in h' (the Hermite
    # slope) would show as a bias of order dx * h' mid-piece
    mid = (np.arange(256) + 0.5) * (beta * rc * (1 + 2.0 ** -7)) / 256
    gm = np.array([L.orc_twin_coul_h(float(v), beta, rc) for v in mid.astype(np.float32)])
    md = mid.astype(np.float32).astype(np.float64)
    rm = sp_erfc(md) + 2.0 / np.sqrt(np.pi) * md * np.exp(-md * md)
    assert np.max(np.abs(gm - rm)) < 1.5e-7
    # beta = 0 (plain cutoff Coulomb): h = 1 everywhere
    assert L.orc_twin_coul_h(1.7, 0.0, rc) == 1.0


def lib_call(fn, v):
    return float(fn(float(v)))


def test_fixed_point_rounding():
    L = oracle.lib()
    assert L.orc_to_fixed(1.0) == 2 ** 32
    assert L.orc_to_fixed(-1.5) == -3 * 2 ** 31
    assert L.orc_to_fixed(2.0 ** -33) == 0            # half an LSB rounds to even (0)
    assert L.orc_to_fixed(3 * 2.0 ** -33) == 2        # 1.5 LSB rounds to even (2)


def test_twin_pair_and_exclusion_match_fp64(water_9k):
    s = water_9k
    excl = (s["excl_start"], s["excl_numbers"], s["excl_list"])
    beta = oracle.get_beta(10.0, 1e-5)
    counts, rows, _ = oracle.nl_build(s["crd"], s["cell"], s["rcell"], 10.0, 2.0, excl)
    tf, telj, tec, tv = oracle.pair_twin(s["crd"], s["charge"], s["lj_type"], s["lj_a"],
                                         s["lj_b"], s["cell"], s["rcell"], 10.0, beta, counts,
                                         rows)
    df, delj, dec, dv = oracle.pair_f64(s["crd"], s["charge"], s["lj_type"], s["lj_a"],
                                        s["lj_b"], s["cell"], s["rcell"], 10.0, beta, counts,
                                        rows)
    assert rel_rms(tf.T / 2 ** 32, df) < 2e-6
    assert abs(telj.sum() / 2 ** 32 - delj.sum()) < 1e-6 * abs(delj.sum())
    # the direct Coulomb sum of neutral waters cancels ~1e3-fold: float per-pair
    # rounding (~1e-7 of each term, random) leaves ~1e-5 of the net sum
    assert abs(tec.sum() / 2 ** 32 - dec.sum()) < 3e-5 * abs(dec.sum())
    assert rel_rms(tv.T / 2 ** 32, dv) < 2e-6
    assert np.all(tf.sum(axis=1) == 0)
    ef, ee, ev = oracle.excl_twin(s["crd"], s["charge"], s["cell"], s["rcell"], beta, excl)
    df2, de2, dv2 = oracle.excl_f64(s["crd"], s["charge"], s["cell"], s["rcell"], beta, excl)
    assert rel_rms(ef.T / 2 ** 32, df2) < 2e-6
    assert abs(ee.sum() / 2 ** 32 - de2.sum()) < 1e-6 * abs(de2.sum())


def test_twin_spread_gather_match_fp64(water_9k):
    s = water_9k
    beta = oracle.get_beta(10.0, 1e-5)
    L = float(s["box"][0])
    vol = float(np.float32(L) ** 3)
    f, e, _, Q, phi = oracle.pme_f64(s["crd"], s["charge"], s["rcell"], (48, 48, 48), beta, vol,
                                     want_grids=True)
    qt = oracle.pme_spread_twin(s["crd"], s["charge"], s["rcell"], (48, 48, 48))
    assert np.max(np.abs(qt / 2 ** 32 - Q)) < 1e-5 * np.max(np.abs(Q))
    ft = oracle.pme_gather_twin(s["crd"], s["charge"], s["rcell"], (48, 48, 48),
                                phi.astype(np.float32))
    assert rel_rms(ft.T / 2 ** 32, f) < 1e-6


# ----------------------------------------------- the reference's fixtures
def test_oracle_reproduces_reference_cpu_output(gromacs_iter0, tip3p_1011):
    """Numbers the reference's own CPU build printed for its committed inputs:
    GROMACS CHARMM TIP3P iter0 (rc 12): LJ_short 1237.962, PM -7437.531;
    TIP3P water 1011 atoms (rc 8): LJ_short 508.1525, PM -3712.9453.  The
    reference runs in FP32 with MKL FFT, so agreement is ~1e-6 relative."""
    r = oracle.nonbonded_terms_f64(gromacs_iter0, cutoff=12.0)
    assert r["grid"] == (72, 72, 72) and abs(r["beta"] - 0.226635) < 2e-6
    assert abs(r["LJ_short"] - 1237.962) < 1e-3
    assert abs(r["PM"] - (-7437.531)) < 7.5e-3
    r = oracle.nonbonded_terms_f64(tip3p_1011, cutoff=8.0)
    assert r["grid"] == (24, 24, 20)
    assert abs(r["LJ_short"] - 508.1525) < 1e-3
    assert abs(r["PM"] - (-3712.9453)) < 3.7e-3


@pytest.mark.parametrize("it", [0, 1, 2])
def test_oracle_within_reference_gromacs_tolerances(it):
    """benchmarks/comparison/tests/gromacs/tests/comp_alanine_dipeptide.py:111-112."""
    s = gromacs_case(it)
    ref = json.load(open(os.path.join(GOLDEN, "gromacs_charmm_tip3p", "reference_terms.json")))
    entry = [x for x in ref["entries"] if x["iteration"] == it][0]
    r = oracle.nonbonded_terms_f64(s, cutoff=12.0)
    assert abs(r["LJ_short"] - entry["lj_sr"]) <= 25.0
    assert abs(r["PM"] - (entry["coulomb_sr"] + entry["coulomb_recip"])) <= 2.0


def test_lj_long_range_factor_histogram_equals_double_loop(tip3p_1011):
    s = tip3p_1011
    t = s["lj_type"]
    b = s["lj_b"]
    tot = 0.0
    for i in range(s["lj_ntypes"]):
        for j in range(s["lj_ntypes"]):
            hi, lo = max(i, j), min(i, j)
            tot += np.sum(t == i) * np.sum(t == j) * float(b[hi * (hi + 1) // 2 + lo])
    assert abs(oracle.lj_c6_total(t, s["lj_ntypes"], b) - tot) < 1e-9 * abs(tot)


# --------------------------------------------------------------- 1-4 pairs
def test_nb14_reader_resolves_fixture_files(gromacs_iter0):
    """xponge/load/native/nb14.hpp:58-95: the extra pairs come first with A x12
    and B x6; the scaled pairs take lj_scale times the pair's x12 / x6 entry."""
    from sponge_amd import native_io
    s = gromacs_iter0
    a, b, A, B, cf = s["nb14"]
    assert len(a) == 22 + 27
    d = os.path.join(GOLDEN, "gromacs_charmm_tip3p", "iter0")
    first = open(os.path.join(d, "sys_flexible_nb14_extra.txt")).read().split()[1:6]
    assert (a[0], b[0]) == (int(first[0]), int(first[1]))
    assert A[0] == np.float32(first[2]) * np.float32(12.0)
    assert B[0] == np.float32(first[3]) * np.float32(6.0)
    row = open(os.path.join(d, "sys_flexible_nb14.txt")).read().split()[1:5]
    ta, tb = s["lj_type"][int(row[0])], s["lj_type"][int(row[1])]
    hi, lo = max(ta, tb), min(ta, tb)
    assert A[22] == np.float32(row[2]) * s["lj_a"][hi * (hi + 1) // 2 + lo]
    assert cf[22] == np.float32(row[3])
    # round trip through the writer (extra-pair form)
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        native_io.write_system(tmp, "t", s)
        back = native_io.system_from_prefix(tmp, "t")["nb14"]
    np.testing.assert_array_equal(back[0], a)
    np.testing.assert_allclose(back[2], A, rtol=1e-6)


def test_nb14_gromacs_fixture_terms(gromacs_iter0):
    """GROMACS 2025.4 reference terms of the fixture: coulomb14 193.9056 within
    the reference's nb14_ee tolerance of 0.2 kcal/mol
    (benchmarks/comparison/tests/gromacs/tests/comp_alanine_dipeptide.py:77,
    :110), and lj14 2.8505 (no tolerance there; 1e-3 relative here)."""
    s = gromacs_iter0
    f, ep, ea, v = oracle.nb14_f64(s["crd"], s["charge"], s["cell"], s["rcell"], s["nb14"])
    assert abs(ep[1].sum() - 193.90558413001912) <= 0.2
    assert abs(ep[0].sum() - 2.850482791586998) <= 1e-3 * 2.850482791586998
    assert abs(ea.sum() - ep.sum()) < 1e-9 * abs(ep.sum())
    # momentum: every pair adds +f and -f
    assert np.abs(f.sum(0)).max() < 1e-9 * np.abs(f).max()


def test_nb14_force_is_minus_energy_gradient():
    """Central differences of the pair energy against the force on atom a,
    for pairs with LJ only, Coulomb only, both, and a zero charge (raised to
    1e-10 by the reference's TINY rule)."""
    rng = np.random.default_rng(4)
    L = 30.0
    cell = np.array([L, 0, L, 0, 0, L], np.float32)
    rcell = np.array([1 / L, 0, 1 / L, 0, 0, 1 / L], np.float32)
    for A, B, qa, qb, cf in ((6.0e5, 700.0, 0.0, 0.0, 0.0), (0.0, 0.0, 7.0, -5.0, 0.833),
                             (2.0e5, 365.0, -6.0, 3.0, 0.5), (2.0e5, 365.0, 0.0, 3.0, 1.0)):
        d = rng.normal(size=3)
        d *= 3.3 / np.linalg.norm(d)
        base = np.array([[10.0, 11.0, 12.0], [10.0, 11.0, 12.0] + d], np.float64)
        q = np.array([qa, qb], np.float32)
        nb14 = (np.array([0]), np.array([1]), np.array([A], np.float32),
                np.array([B], np.float32), np.array([cf], np.float32))

        def energy(crd):
            _, ep, _, _ = oracle.nb14_f64(crd.astype(np.float32), q, cell, rcell, nb14,
                                          virial=False)
            return ep.sum()

        f, _, _, _ = oracle.nb14_f64(base.astype(np.float32), q, cell, rcell, nb14, virial=False)
        h = 1e-2
        g = np.zeros(3)
        for c in range(3):
            p, m = base.copy(), base.copy()
            p[0, c] += h
            m[0, c] -= h
            g[c] = -(energy(p) - energy(m)) / (2 * h)
        scale = np.abs(f[0]).max() + 1e-12
        assert np.abs(g - f[0]).max() < 2e-3 * scale, (A, B, qa, qb, g, f[0])
        np.testing.assert_allclose(f[1], -f[0], rtol=1e-12, atol=1e-300)


def test_nb14_twin_matches_fp64(mixed_small):
    s = mixed_small
    f64, ep64, ea64, v64 = oracle.nb14_f64(s["crd"], s["charge"], s["cell"], s["rcell"],
                                           s["nb14"])
    ft, ept, eat, vt = oracle.nb14_twin(s["crd"], s["charge"], s["cell"], s["rcell"], s["nb14"])
    assert rel_rms(ft.T * 2.0 ** -32, f64) < 1e-6
    # energies: double arithmetic on the float displacement (its rounding,
    # ~1e-7 relative, is amplified by the r^-12 term)
    assert abs(ept.sum() * 2.0 ** -32 - ep64.sum()) < 1e-6 * np.abs(ep64).sum()
    assert rel_rms(vt.T * 2.0 ** -32, v64) < 1e-6


# ------------------------------------------------------------ harmonic bonds
def test_bond_reader_and_gromacs_fixture_terms():
    """xponge/load/native/bond.hpp:8-40 reads the fixture's 1,426 bonds; their
    energy on iterations 0-2 matches the GROMACS 2025.4 bond term within the
    reference's tolerance of 0.5 kcal/mol
    (benchmarks/comparison/tests/gromacs/tests/comp_alanine_dipeptide.py:69,
    :107)."""
    from sponge_amd import native_io
    ref = json.load(open(os.path.join(GOLDEN, "gromacs_charmm_tip3p", "reference_terms.json")))
    for it in (0, 1, 2):
        s = gromacs_case(it)
        a, b, k, r0 = s["bonds"]
        assert len(a) == 1426
        if it == 0:
            row = open(os.path.join(GOLDEN, "gromacs_charmm_tip3p", "iter0",
                                    "sys_flexible_bond.txt")).read().split()[1:5]
            assert (a[0], b[0], k[0], r0[0]) == (int(row[0]), int(row[1]), np.float32(row[2]),
                                                 np.float32(row[3]))
        f, ep, ea, v = oracle.bond_f64(s["crd"], s["cell"], s["rcell"], s["bonds"])
        want = [x for x in ref["entries"] if x["iteration"] == it][0]["bond"]
        assert abs(ep.sum() - want) <= 0.5, (it, ep.sum(), want)
        assert abs(ea.sum() - ep.sum()) < 1e-9 * ep.sum()
        assert np.abs(f.sum(0)).max() < 1e-9 * np.abs(f).max()
        # the virial trace of a pair force is f . dr = -2 k (r - r0) r
        d = s["crd"][a].astype(np.float64) - s["crd"][b]
        L = np.asarray(s["box"], np.float64)
        d -= np.rint(d / L) * L
        r = np.linalg.norm(d, axis=1)
        tr = v[:, 0] + v[:, 2] + v[:, 5]
        np.testing.assert_allclose(tr.sum(), np.sum(-2.0 * k * (r - r0) * r), rtol=1e-9)
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        native_io.write_system(tmp, "t", gromacs_case(0))
        back = native_io.system_from_prefix(tmp, "t")["bonds"]
    for x, y in zip(back, gromacs_case(0)["bonds"]):
        np.testing.assert_array_equal(x, y)


def test_bond_closed_form_and_gradient():
    """One bond of length r along a random direction (also across the periodic
    boundary): E = k (r - r0)^2, |f| = 2 k |r - r0| pointing back to r0, and
    central differences of E against the force on atom a."""
    rng = np.random.default_rng(11)
    L = 25.0
    cell = np.array([L, 0, L, 0, 0, L], np.float32)
    rcell = np.array([1 / L, 0, 1 / L, 0, 0, 1 / L], np.float32)
    for r, r0, kk, at in ((1.6, 1.53, 310.0, 5.0), (1.4, 1.53, 310.0, 24.5),
                          (0.95, 0.9572, 450.0, 0.2)):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        base = np.array([[at, at, at], np.array([at, at, at]) - r * u], np.float64) % L
        bonds = (np.array([0]), np.array([1]), np.array([kk], np.float32),
                 np.array([r0], np.float32))
        f, ep, ea, _ = oracle.bond_f64(base.astype(np.float32), cell, rcell, bonds)
        x = float(np.float32(r0))
        assert abs(ep[0] - kk * (r - x) ** 2) < 1e-5 * max(1.0, ep[0])
        np.testing.assert_allclose(f[0], -2 * kk * (r - x) * u, atol=2e-4 * kk)
        np.testing.assert_allclose(f[1], -f[0], atol=1e-12)

        def energy(c):
            return oracle.bond_f64(c.astype(np.float32), cell, rcell, bonds, virial=False)[1][0]

        h = 1e-3
        g = np.zeros(3)
        for c in range(3):
            p, m = base.copy(), base.copy()
            p[0, c] += h
            m[0, c] -= h
            g[c] = (energy(p) - energy(m)) / (2 * h)
        np.testing.assert_allclose(f[0], -g, atol=2e-3 * max(1.0, np.abs(f[0]).max()))


def test_bond_twin_matches_fp64(mixed_small):
    """The float twin against the FP64 layer on the chain bonds, stretched
    off r0 by N(0, 0.03) A: the force carries the float r - r0 cancellation
    (~1e-7 A on a ~0.03 A stretch), the energy is double on the float
    displacement."""
    s = mixed_small
    rng = np.random.default_rng(12)
    crd = (s["crd"] + rng.normal(0, 0.03, s["crd"].shape)).astype(np.float32)
    f64, ep64, ea64, v64 = oracle.bond_f64(crd, s["cell"], s["rcell"], s["bonds"])
    ft, ept, eat, vt = oracle.bond_twin(crd, s["cell"], s["rcell"], s["bonds"])
    assert rel_rms(ft.T * 2.0 ** -32, f64) < 5e-5
    assert abs(ept.sum() * 2.0 ** -32 - ep64.sum()) < 1e-6 * ep64.sum()
    np.testing.assert_array_equal(np.bincount(s["bonds"][0], ept, s["n"]).astype(np.int64), eat)
    assert rel_rms(vt.T * 2.0 ** -32, v64) < 5e-5


def test_pme_reciprocal_virial_is_half_the_strain_derivative(gromacs_iter0):
    """The reciprocal virial as the reference forms it (PM_force.cpp:575-607,
    PME_Sum_Virial :1021-1065) against central differences of E_rec when the box
    and the coordinates are stretched along one axis (fractional coordinates
    fixed, grid and beta fixed): each diagonal entry equals +1/2 dE_rec/dlambda.
    (-dE/dlambda is the strain virial in the convention of the atom virials
    f (x) dr, so the reference's reciprocal term carries the opposite sign and
    half the size.)"""
    s = gromacs_iter0
    crd = np.asarray(s["crd"], np.float64)
    L = np.asarray(s["box"], np.float64)
    grid = (72, 72, 72)
    beta = oracle.get_beta(12.0, 1e-5)

    def erec(lam):
        lam = np.asarray(lam, np.float64)
        c = (crd * lam).astype(np.float32)
        rc = np.array([1 / (L[0] * lam[0]), 0, 1 / (L[1] * lam[1]), 0, 0,
                       1 / (L[2] * lam[2])], np.float32)
        vol = float(np.float32(L[0] * lam[0]) * np.float32(L[1] * lam[1]) *
                    np.float32(L[2] * lam[2]))
        _, e, v, _, _ = oracle.pme_f64(c, s["charge"], rc, grid, beta, vol)
        return e, v

    _, vir = erec([1.0, 1.0, 1.0])
    h = 1e-2
    for axis, entry in ((0, 0), (1, 2), (2, 5)):
        lp, lm = [1.0] * 3, [1.0] * 3
        lp[axis] += h
        lm[axis] -= h
        d = (erec(lp)[0] - erec(lm)[0]) / (2 * h)
        assert abs(vir[entry] - 0.5 * d) < 1e-3 * abs(vir[entry]), (axis, vir[entry], d)


# ------------------------------------------- per-atom forces against GROMACS
@pytest.mark.parametrize("it", [0, 1, 2])
def test_gromacs_fixture_forces_on_waters(it):
    """GROMACS 2025.4 per-atom forces of the fixture (float64, kcal/mol/A,
    benchmarks/comparison/reference/gromacs/forces/...) on the 2,106 water
    atoms, whose forces hold the nonbonded path, the O-H bonds and the H-O-H
    angles: the oracle's path + bonds + angle/Urey-Bradley terms agree within
    the reference's own tolerances (max |dF| <= 0.07, RMS <= 0.01, cosine >=
    0.999; comp_alanine_dipeptide.py:113-115).  The fixture's angle +
    Urey-Bradley energy matches GROMACS's angle + urey_bradley within the
    reference's 0.5 kcal/mol (:108)."""
    from conftest import force_stats, gromacs_water_atoms
    s = gromacs_case(it)
    ref = json.load(open(os.path.join(GOLDEN, "gromacs_charmm_tip3p", "reference_terms.json")))
    entry = [x for x in ref["entries"] if x["iteration"] == it][0]
    r = oracle.nonbonded_terms_f64(s, cutoff=12.0)
    fa, ea = oracle.angle_f64(s["crd"], s["cell"], s["rcell"], s["urey_bradley"][0])
    fu, eu, _, _ = oracle.bond_f64(s["crd"], s["cell"], s["rcell"], s["urey_bradley"][1])
    assert abs(ea.sum() + eu.sum() - (entry["angle"] + entry["urey_bradley"])) <= 0.5
    gmx = np.load(os.path.join(GOLDEN, "gromacs_charmm_tip3p", "iter%d" % it, "gromacs_forces.npy"),
                  allow_pickle=False)
    w = gromacs_water_atoms(s)
    assert w.sum() == 3 * 702
    mx, rms, cos = force_stats(gmx[w], (r["frc"] + fa + fu)[w])
    print("water atoms: max", mx, "rms", rms, "cos", cos)
    assert mx <= 0.07 and rms <= 0.01 and cos >= 0.999


def test_angle_force_is_minus_energy_gradient():
    """Central differences of the angle energy against the forces on all three
    atoms, for a bent and a nearly linear angle."""
    rng = np.random.default_rng(5)
    L = 30.0
    cell = np.array([L, 0, L, 0, 0, L], np.float32)
    rcell = np.array([1 / L, 0, 1 / L, 0, 0, 1 / L], np.float32)
    for th in (1.2, 2.9):
        b = np.array([15.0, 15.0, 15.0])
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        v = rng.normal(size=3)
        v -= v.dot(u) * u
        v /= np.linalg.norm(v)
        a = b + 1.3 * u
        c = b + 1.1 * (np.cos(th) * u + np.sin(th) * v)
        base = np.array([a, b, c], np.float64)
        ang = (np.array([0]), np.array([1]), np.array([2]), np.array([60.0], np.float32),
               np.array([1.9], np.float32))
        f, e = oracle.angle_f64(base.astype(np.float32), cell, rcell, ang)
        h = 1e-3
        for atom in range(3):
            for d in range(3):
                p, m = base.copy(), base.copy()
                p[atom, d] += h
                m[atom, d] -= h
                g = (oracle.angle_f64(p.astype(np.float32), cell, rcell, ang)[1][0] -
                     oracle.angle_f64(m.astype(np.float32), cell, rcell, ang)[1][0]) / (2 * h)
                assert abs(f[atom, d] + g) < 2e-2 * max(1.0, np.abs(f).max()), (th, atom, d)
        assert np.abs(f.sum(0)).max() < 1e-9 * max(1.0, np.abs(f).max())


def _pair_case(m=64, seed=5, r0=1.0, spread=0.03):
    """Rigid pairs at distance r0 in random directions, remembered, then each
    atom moved at random (a box large enough for no wrap)."""
    import oracle
    rng = np.random.default_rng(seed)
    cell = np.array([60.0, 0, 60.0, 0, 0, 60.0], np.float32)
    rcell = np.array([1 / 60.0, 0, 1 / 60.0, 0, 0, 1 / 60.0], np.float32)
    u = rng.normal(size=(m, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    xi = rng.uniform(10, 50, (m, 3))
    crd = np.empty((2 * m, 3), np.float32)
    crd[0::2] = xi
    crd[1::2] = xi + r0 * u
    pairs = np.stack([np.arange(0, 2 * m, 2), np.arange(1, 2 * m, 2)], 1).astype(np.int32)
    mass = rng.uniform(1.0, 16.0, 2 * m).astype(np.float32)
    rk = np.full(m, r0, np.float32)
    ck = (np.float32(1.0) / (mass[0::2] + mass[1::2])).astype(np.float32)
    ab = oracle.settle_pair_remember(pairs, crd, cell, rcell)
    moved = (crd + rng.normal(0, spread, crd.shape)).astype(np.float32)
    return oracle, pairs, rk, ck, mass, moved, cell, rcell, ab


def test_settle_pair_restores_distance_momentum_and_direction():
    """settle_pair (settle.cpp:258-307): the pair distance is r0 again, the
    mass-weighted displacement (and so the momentum change) is zero, and each
    correction lies along the remembered bond j - i."""
    oracle, pairs, rk, ck, mass, moved, cell, rcell, ab = _pair_case()
    before = moved.copy()
    vel = np.zeros_like(moved)
    oracle.settle_pair_twin(pairs, rk, ck, mass, moved, cell, rcell, ab, np.float32(0.04),
                            np.float32(1.0), np.float32(1.0), vel)
    d = moved[pairs[:, 1]].astype(np.float64) - moved[pairs[:, 0]]
    # coordinates up to 50 A: a float ulp there is 3.8e-6 A
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, rtol=0, atol=4 * 3.8e-6)
    # the velocity change is the displacement over the step, free of the
    # coordinates' rounding: momentum and direction are checked on it
    v = vel.astype(np.float64)
    p = mass[pairs[:, 0], None] * v[pairs[:, 0]] + mass[pairs[:, 1], None] * v[pairs[:, 1]]
    assert np.abs(p).max() <= 1e-6 * np.abs(mass[:, None] * v).max()
    cr = np.cross(v[pairs[:, 1]], ab.astype(np.float64))
    assert np.abs(cr).max() <= 1e-6 * np.abs(v).max() * np.abs(ab).max()
    # and the coordinates moved by it times dt (v_factor = x_factor = 1), to
    # their float resolution
    np.testing.assert_allclose(moved - before, v * 0.04, rtol=0, atol=4 * 3.8e-6)


def test_settle_pair_stretch_along_bond_and_virial():
    """A bond stretched along itself comes back to r0 along the bond; the
    virial sums equal (m_j dx_j / dt^2) (x) r2 summed over pairs (Get_Virial_From_Force_Dis
    layout: xx, xy+yx, yy, xz+zx, yz+zy, zz) at v_factor = x_factor = 1."""
    import oracle
    cell = np.array([60.0, 0, 60.0, 0, 0, 60.0], np.float32)
    rcell = np.array([1 / 60.0, 0, 1 / 60.0, 0, 0, 1 / 60.0], np.float32)
    crd = np.array([[20, 20, 20], [21.2, 20.0, 20.0], [30, 30, 30], [30.6, 30.8, 30.0]],
                   np.float32)
    r0 = np.array([1.0, 1.0], np.float32)
    crd[3] = crd[2] + np.array([0.6, 0.8, 0.0], np.float32)
    pairs = np.array([[0, 1], [2, 3]], np.int32)
    mass = np.array([16.0, 1.0, 12.0, 12.0], np.float32)
    ck = (np.float32(1.0) / (mass[[0, 2]] + mass[[1, 3]])).astype(np.float32)
    ref = crd.copy()
    ref[1] = ref[0] + np.array([1.0, 0, 0], np.float32)      # pair 0 remembered at r0
    ab = oracle.settle_pair_remember(pairs, ref, cell, rcell)
    moved = crd.copy()
    moved[3] = moved[2] + np.float32(1.3) * np.array([0.6, 0.8, 0.0], np.float32)  # stretched
    vel = np.zeros_like(moved)
    dt = np.float32(0.05)
    vir = oracle.settle_pair_twin(pairs, r0, ck, mass, moved, cell, rcell, ab, dt,
                                  np.float32(1.0), np.float32(1.0), vel)
    np.testing.assert_allclose(np.linalg.norm(moved[1] - moved[0]), 1.0, rtol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(moved[3] - moved[2]), 1.0, rtol=1e-6)
    assert abs(moved[1, 1] - 20.0) < 1e-6 and abs(moved[1, 2] - 20.0) < 1e-6
    # m_j dx_j / dt^2 = m_j dv_j / dt (dv = dx / dt here), from the velocities,
    # which carry no coordinate rounding
    v = vel.astype(np.float64)
    want = np.zeros((3, 3))
    for t, (i, j) in enumerate(pairs):
        f = mass[j] * v[j] / float(dt)
        want += np.outer(f, ab[t].astype(np.float64))
    got = np.asarray(vir, np.float64) * 2.0 ** -32
    lay = np.array([want[0, 0], want[0, 1] + want[1, 0], want[1, 1], want[0, 2] + want[2, 0],
                    want[1, 2] + want[2, 1], want[2, 2]])
    np.testing.assert_allclose(got, lay, rtol=1e-5, atol=1e-6 * np.abs(lay).max())


def test_grid_rule_keeps_every_pair_of_an_orthorhombic_cell():
    """The reference's neighbour-grid rule (oracle.nl_cell_rule) reduces to the
    +-2-grid stencil for an orthorhombic cell, which holds every pair within
    rc + skin: no brute-force pair of the 9k water box falls outside it.
    With tilts below a grid width (angles 85/87/84) the rule leaves some out."""
    from sponge_amd import native_io, systems
    s = systems.config_cpu_9k()
    excl = (s["excl_start"], s["excl_numbers"], s["excl_list"])
    out = {}
    for angle in ((90.0, 90.0, 90.0), (85.0, 87.0, 84.0)):
        cell, rcell = native_io.box_to_cell(s["box"], angle)
        oc, orows, ov = oracle.nl_build(s["crd"], cell, rcell, 10.0, 2.0, excl)
        assert not ov
        rule = oracle.nl_cell_rule(s["crd"], cell, 10.0, 2.0)
        out[angle] = sum(int(oc[i]) - int(rule(np.full(oc[i], i), orows[i, :oc[i]]).sum())
                         for i in range(s["n"]))
    assert out[(90.0, 90.0, 90.0)] == 0
    assert out[(85.0, 87.0, 84.0)] > 0
