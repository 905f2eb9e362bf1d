"""Synthetic systems of the benchmark configurations, as arrays in the SPONGE
native conventions (charges x18.2223, LJ tables x12 / x6 in memory, exclusion
CSR listing j > i).

* ``tip3p_box``      rigid-geometry TIP3P waters on a jittered cubic lattice
                     with uniformly random orientations (Shoemake quaternions);
                     3,000 waters in a 44.8 A box is the 9,000-atom CPU case.
* ``mixed_system``   TIP3P waters plus bead-spring chains (bond 1.53 A, one LJ
                     type per bead, Lorentz-Berthelot mixing, exclusions
                     i -> i+1..i+3 along a chain); 30,000 waters + 3 x 1,000
                     beads in a 96.5 A box is the 93,000-atom GPU case.
"""
import numpy as np

from .native_io import CHARGE_SCALE, box_to_cell

# TIP3P (O-H 0.9572 A, HOH 104.52 deg; O -0.834 / H +0.417; LJ on O only)
TIP3P_ROH = 0.9572
TIP3P_THETA = np.deg2rad(104.52)
TIP3P_SIGMA = 3.15061
TIP3P_EPS = 0.1521
TIP3P_QO = -0.834
TIP3P_QH = 0.417
MASS_O, MASS_H = 15.999, 1.008


def _random_rotations(rng, n):
    """Uniform random rotation matrices (Shoemake, Graphics Gems III, 1992)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2),
                  np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3),
                  np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _water_geometry():
    half = TIP3P_THETA / 2
    return np.array([[0.0, 0.0, 0.0],
                     [TIP3P_ROH * np.sin(half), TIP3P_ROH * np.cos(half), 0.0],
                     [-TIP3P_ROH * np.sin(half), TIP3P_ROH * np.cos(half), 0.0]])


def _place_waters(rng, o_sites):
    R = _random_rotations(rng, len(o_sites))
    geo = _water_geometry()
    return (o_sites[:, None, :] + np.einsum("nij,kj->nki", R, geo)).reshape(-1, 3)


def _lattice_sites(box, per_axis, jitter, rng):
    g = (np.arange(per_axis) + 0.5) * (box / per_axis)
    s = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    return s + rng.uniform(-jitter, jitter, s.shape)


def _lj_tables(sigma, eps):
    """Per-type sigma/eps -> in-memory A (x12) and B (x6) lower-triangle tables
    with Lorentz-Berthelot mixing; a type with eps = 0 has no LJ."""
    t = len(sigma)
    ii, jj = np.tril_indices(t)
    order = np.argsort(ii * t + jj)  # row-major lower triangle: (i, j<=i)
    ii, jj = ii[order], jj[order]
    s = 0.5 * (sigma[ii] + sigma[jj])
    e = np.sqrt(eps[ii] * eps[jj])
    a_file = 4.0 * e * s ** 12
    b_file = 4.0 * e * s ** 6
    return (a_file.astype(np.float32) * np.float32(12.0),
            b_file.astype(np.float32) * np.float32(6.0))


def _csr(n, pairs):
    lists = [[] for _ in range(n)]
    for i, j in pairs:
        a, b = (i, j) if i < j else (j, i)
        lists[a].append(b)
    num = np.array([len(l) for l in lists], np.int32)
    start = np.zeros(n, np.int32)
    start[1:] = np.cumsum(num)[:-1]
    flat = np.array([j for l in lists for j in sorted(l)], np.int32)
    return start, num, flat


def _finish(crd, charge, ltype, sigma, eps, mass, pairs, box, cutoff, **extra):
    n = len(crd)
    a, b = _lj_tables(np.asarray(sigma, np.float64), np.asarray(eps, np.float64))
    start, num, flat = _csr(n, pairs)
    boxv = np.array([box, box, box], np.float32)
    cell, rcell = box_to_cell(boxv)
    d = dict(n=n, crd=np.asarray(crd, np.float32), charge=np.asarray(charge, np.float32),
             lj_ntypes=len(sigma), lj_a=a, lj_b=b, lj_type=np.asarray(ltype, np.int32),
             mass=np.asarray(mass, np.float32), box=boxv,
             angle=np.array([90.0, 90.0, 90.0], np.float32), cell=cell, rcell=rcell,
             excl_start=start, excl_numbers=num, excl_list=flat, cutoff=float(cutoff))
    d.update(extra)
    return d


def tip3p_box(n_waters=3000, box=44.8, seed=20260217, jitter=0.2, cutoff=10.0):
    """n_waters TIP3P waters on a jittered lattice (random subset of sites)."""
    rng = np.random.default_rng(seed)
    per = int(np.ceil(n_waters ** (1.0 / 3.0) - 1e-9))
    sites = _lattice_sites(box, per, jitter, rng)
    pick = np.sort(rng.choice(len(sites), n_waters, replace=False))
    crd = _place_waters(rng, sites[pick]) % box
    n = 3 * n_waters
    charge = np.tile([TIP3P_QO, TIP3P_QH, TIP3P_QH], n_waters) * CHARGE_SCALE
    ltype = np.tile([0, 1, 1], n_waters)
    mass = np.tile([MASS_O, MASS_H, MASS_H], n_waters)
    pairs = []
    for w in range(n_waters):
        o = 3 * w
        pairs += [(o, o + 1), (o, o + 2), (o + 1, o + 2)]
    return _finish(crd, charge, ltype, [TIP3P_SIGMA, 0.0], [TIP3P_EPS, 0.0], mass, pairs,
                   box, cutoff, n_waters=n_waters)


def _chain_walk(rng, n_beads, box, bond, placed_tree_pts, min_dist=3.0, angle_deg=111.0):
    """Self-avoiding periodic walk with a fixed bond angle and random
    dihedrals; a new bead keeps >= min_dist from every bead placed earlier
    except its three predecessors."""
    from scipy.spatial import cKDTree

    cos_a = np.cos(np.deg2rad(180.0 - angle_deg))
    others = list(placed_tree_pts)
    tree = cKDTree(np.array(others) % box, boxsize=box) if others else None

    def clear_of_others(p):
        return tree is None or tree.query(p % box)[0] >= min_dist

    while True:
        start = rng.uniform(0, box, 3)
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        if clear_of_others(start) and clear_of_others(start + bond * d):
            break
    pts = [start, start + bond * d]
    failures = 0
    while len(pts) < n_beads:
        placed = False
        for _attempt in range(100):
            prev = pts[-1] - pts[-2]
            prev /= np.linalg.norm(prev)
            # random direction at the fixed angle to the previous bond
            perp = np.cross(prev, rng.normal(size=3))
            perp /= np.linalg.norm(perp)
            new_dir = cos_a * prev + np.sqrt(1 - cos_a ** 2) * perp
            cand = pts[-1] + bond * new_dir
            if len(pts) > 3:
                dd = (np.array(pts[:-3]) - cand + box / 2) % box - box / 2
                if np.min(np.einsum("ij,ij->i", dd, dd)) < min_dist ** 2:
                    continue
            if not clear_of_others(cand):
                continue
            pts.append(cand)
            placed = True
            break
        if not placed:
            # trapped: back off a few beads and regrow
            failures += 1
            if failures > 1000:
                raise RuntimeError("chain growth failed")
            del pts[max(2, len(pts) - 5):]
    return np.array(pts)


def mixed_system(n_waters=30000, n_chains=3, beads_per_chain=1000, box=96.5,
                 seed=20260217, bond=1.53, cutoff=10.0, water_clearance=2.5):
    """Waters + bead chains; bead charges N(0, 0.3) shifted to zero sum,
    sigma U(3, 4) A, eps U(0.02, 0.25) kcal/mol, one LJ type per bead."""
    from scipy.spatial import cKDTree

    rng = np.random.default_rng(seed)
    beads = []
    for _c in range(n_chains):
        beads.extend(_chain_walk(rng, beads_per_chain, box, bond, beads))
    beads = np.array(beads) % box
    nb = len(beads)
    # water O sites: a lattice fine enough to leave n_waters free sites
    tree = cKDTree(beads, boxsize=box)
    per = int(np.ceil(n_waters ** (1.0 / 3.0)))
    while True:
        sites = _lattice_sites(box, per, 0.15, rng) % box
        free = np.array([len(x) == 0 for x in tree.query_ball_point(sites, water_clearance)])
        if free.sum() >= n_waters:
            break
        per += 1
    cand = np.nonzero(free)[0]
    pick = np.sort(rng.choice(cand, n_waters, replace=False))
    wcrd = _place_waters(rng, sites[pick]) % box
    crd = np.concatenate([wcrd, beads])
    nw = 3 * n_waters
    qb = rng.normal(0.0, 0.3, nb)
    qb -= qb.mean()
    charge = np.concatenate([np.tile([TIP3P_QO, TIP3P_QH, TIP3P_QH], n_waters), qb]) * CHARGE_SCALE
    sig = np.concatenate([[TIP3P_SIGMA, 0.0], rng.uniform(3.0, 4.0, nb)])
    eps = np.concatenate([[TIP3P_EPS, 0.0], rng.uniform(0.02, 0.25, nb)])
    ltype = np.concatenate([np.tile([0, 1, 1], n_waters), 2 + np.arange(nb)])
    mass = np.concatenate([np.tile([MASS_O, MASS_H, MASS_H], n_waters), np.full(nb, 14.027)])
    pairs = []
    for w in range(n_waters):
        o = 3 * w
        pairs += [(o, o + 1), (o, o + 2), (o + 1, o + 2)]
    for c in range(n_chains):
        b0 = nw + c * beads_per_chain
        for k in range(beads_per_chain):
            for d in (1, 2, 3):
                if k + d < beads_per_chain:
                    pairs.append((b0 + k, b0 + k + d))
    out = _finish(crd, charge, ltype, sig, eps, mass, pairs, box, cutoff,
                  n_waters=n_waters, n_beads=nb)
    # 1-4 pairs along each chain: LJ scaled by 0.5, Coulomb by 0.833 (the
    # form the nb14 loader resolves from the LJ tables)
    a14, b14 = [], []
    for c in range(n_chains):
        b0 = nw + c * beads_per_chain
        for k in range(beads_per_chain - 3):
            a14.append(b0 + k)
            b14.append(b0 + k + 3)
    a14 = np.array(a14, np.int32)
    b14 = np.array(b14, np.int32)
    ta, tb = out["lj_type"][a14], out["lj_type"][b14]
    hi, lo = np.maximum(ta, tb), np.minimum(ta, tb)
    p14 = hi * (hi + 1) // 2 + lo
    out["nb14"] = (a14, b14, np.float32(0.5) * out["lj_a"][p14],
                   np.float32(0.5) * out["lj_b"][p14], np.full(len(a14), 0.833, np.float32))
    # chain bonds: harmonic, r0 = the walk's bond length, k 310 kcal/mol/A^2
    # (E = k (r - r0)^2, the C-C stretch of the united-atom chains)
    ba = np.concatenate([nw + c * beads_per_chain + np.arange(beads_per_chain - 1)
                         for c in range(n_chains)]).astype(np.int32)
    out["bonds"] = (ba, (ba + 1).astype(np.int32), np.full(len(ba), 310.0, np.float32),
                    np.full(len(ba), bond, np.float32))
    # No LJ site (water O, bead) within 2.2 A of another non-excluded LJ site,
    # and no two atoms of different molecules within 0.6 A (random water
    # orientations on the lattice can bring two H's to ~0.8 A; H carries no LJ).
    mol = np.concatenate([np.repeat(np.arange(n_waters), 3),
                          n_waters + np.arange(nb) // beads_per_chain])
    pos = np.concatenate([np.zeros(nw, int), np.arange(nb) % beads_per_chain])
    lj_site = np.concatenate([np.tile([True, False, False], n_waters), np.ones(nb, bool)])
    c = out["crd"] % box
    t = cKDTree(c, boxsize=box)
    close = t.query_pairs(2.2, output_type="ndarray")
    a, b = close[:, 0], close[:, 1]
    excluded = (mol[a] == mol[b]) & ((a < nw) | (np.abs(pos[a] - pos[b]) <= 3))
    d = (c[a] - c[b] + box / 2) % box - box / 2
    r = np.sqrt(np.einsum("ij,ij->i", d, d))
    bad = ~excluded & ((lj_site[a] & lj_site[b]) | (r < 0.6))
    if bad.any():
        raise RuntimeError("generator produced overlapping atoms")
    return out


def config_cpu_9k(seed=20260217):
    """3,000 waters, 44.8 A box, cutoff 10 A, PME grid 48^3 (automatic)."""
    return tip3p_box(3000, 44.8, seed)


def config_gpu_93k(seed=20260217):
    """30,000 waters + 3 chains x 1,000 beads, 96.5 A box, PME grid 96^3."""
    d = mixed_system(seed=seed)
    d["pme_grid"] = (96, 96, 96)
    return d


def config_water_1m(seed=20260217):
    """333,334 TIP3P waters (1,000,002 atoms) at 1.0 g/cm^3 (215.24 A box),
    cutoff 10 A, PME grid 216^3 (automatic)."""
    return tip3p_box(333334, 215.24, seed)


KB = 0.00198716  # kcal/mol/K (common.h:122)
TIME_UNIT = 20.455  # internal time unit is 1/20.455 ps (common.h:136)


def water_triangles(sysd):
    """Rigid TIP3P triangles (O, H1, H2) of a generated system (waters come
    first, three atoms each): atom arrays and the constraint lengths O-H,
    O-H and H-H (the SETTLE triangle the reference finds among the constraint
    pairs of a rigid water)."""
    nw = int(sysd.get("n_waters", sysd["n"] // 3))
    o = 3 * np.arange(nw, dtype=np.int32)
    r_oh = np.float32(TIP3P_ROH)
    r_hh = np.float32(2.0 * TIP3P_ROH * np.sin(TIP3P_THETA / 2))
    return (o, o + 1, o + 2, np.full(nw, r_oh, np.float32), np.full(nw, r_oh, np.float32),
            np.full(nw, r_hh, np.float32))


def maxwell_velocities(mass, temperature=300.0, seed=20260217, remove_com=True):
    """Maxwell-Boltzmann velocities in the internal unit (A per 1/20.455 ps):
    sigma = sqrt(kB T / m) per component; zero-mass atoms get none; the
    centre-of-mass velocity is removed."""
    rng = np.random.default_rng(seed)
    m = np.asarray(mass, np.float64)
    sig = np.where(m > 0, np.sqrt(KB * temperature / np.where(m > 0, m, 1.0)), 0.0)
    v = rng.normal(size=(len(m), 3)) * sig[:, None]
    if remove_com and m.sum() > 0:
        v -= (m[:, None] * v).sum(0) / m.sum()
        v[m == 0] = 0.0
    return v.astype(np.float32)
