"""SPONGE native text formats: readers and writers.

The formats are the ones Xponge's native loaders read
(xponge/load/native/lj.hpp:8-69, xponge/load/native/md_core.hpp:8-284):

  LJ        ``N T`` / A lower triangle (T(T+1)/2) / B lower triangle / type per atom.
            The loader multiplies A by 12 and B by 6 (lj.hpp:47, :57).
  charge    ``N`` / one charge per atom (already x18.2223).
  mass      ``N`` / one mass per atom.
  coordinate ``N [t0]`` / xyz per atom / ``Lx Ly Lz`` / ``alpha beta gamma``.
  exclude   ``N total`` / per atom ``count j1 j2 ...``.
  residue   ``N R`` / atoms per residue.
  nb14      ``M`` / per pair ``a b lj_scale cf_scale`` (xponge/load/native/nb14.hpp:
            A = lj_scale * the pair's x12 A, B likewise); nb14_extra ``M`` / per
            pair ``a b A B cf_scale`` (A x12 and B x6 on load); extra pairs first.

``system_from_prefix`` assembles the arrays the nonbonded path needs, with the
box converted to the reference's LTMatrix3 cell/rcell (MD_core/pbc.hpp:66-101).
"""
import os

import numpy as np

CHARGE_SCALE = 18.2223  # common.h:132


def _tokens(path):
    with open(path) as f:
        return f.read().split()


def read_lj(path):
    t = _tokens(path)
    n, nt = int(t[0]), int(t[1])
    npair = nt * (nt + 1) // 2
    a = np.array(t[2:2 + npair], dtype=np.float32) * np.float32(12.0)
    b = np.array(t[2 + npair:2 + 2 * npair], dtype=np.float32) * np.float32(6.0)
    types = np.array(t[2 + 2 * npair:2 + 2 * npair + n], dtype=np.int32)
    return n, nt, a, b, types


def read_scalar_list(path, dtype=np.float32):
    t = _tokens(path)
    n = int(t[0])
    return np.array(t[1:1 + n], dtype=dtype)


def read_coordinate(path):
    with open(path) as f:
        first = f.readline().split()
        n = int(first[0])
        rest = f.read().split()
    xyz = np.array(rest[:3 * n], dtype=np.float32).reshape(n, 3)
    box = np.array(rest[3 * n:3 * n + 3], dtype=np.float32)
    ang = np.array(rest[3 * n + 3:3 * n + 6], dtype=np.float32)
    return xyz, box, ang


def read_exclude(path):
    t = _tokens(path)
    n, total = int(t[0]), int(t[1])
    pos = 2
    numbers = np.zeros(n, np.int32)
    start = np.zeros(n, np.int32)
    lst = []
    for i in range(n):
        c = int(t[pos])
        pos += 1
        start[i] = len(lst)
        numbers[i] = c
        lst.extend(int(x) for x in t[pos:pos + c])
        pos += c
    if len(lst) != total:
        raise ValueError("exclude_in_file: total does not match the lists")
    return start, numbers, np.array(lst, dtype=np.int32)


def read_nb14(path, extra_path, lj_type, lj_a, lj_b):
    """1-4 pair list as Xponge::Native_Load_NB14 builds it (nb14.hpp:8-95):
    returns (atom_a, atom_b, A, B, cf) with the extra pairs first.  Either
    path may be None."""
    rows_x, rows_m = [], []
    if extra_path is not None:
        t = _tokens(extra_path)
        rows_x = [t[1 + 5 * k:6 + 5 * k] for k in range(int(t[0]))]
    if path is not None:
        t = _tokens(path)
        rows_m = [t[1 + 4 * k:5 + 4 * k] for k in range(int(t[0]))]
    m = len(rows_x) + len(rows_m)
    a = np.zeros(m, np.int32)
    b = np.zeros(m, np.int32)
    A = np.zeros(m, np.float32)
    B = np.zeros(m, np.float32)
    cf = np.zeros(m, np.float32)
    for k, r in enumerate(rows_x):
        a[k], b[k] = int(r[0]), int(r[1])
        A[k] = np.float32(r[2]) * np.float32(12.0)
        B[k] = np.float32(r[3]) * np.float32(6.0)
        cf[k] = np.float32(r[4])
    for q, r in enumerate(rows_m):
        k = len(rows_x) + q
        a[k], b[k] = int(r[0]), int(r[1])
        ta, tb = int(lj_type[a[k]]), int(lj_type[b[k]])
        hi, lo = max(ta, tb), min(ta, tb)
        p = hi * (hi + 1) // 2 + lo
        A[k] = np.float32(r[2]) * lj_a[p]
        B[k] = np.float32(r[2]) * lj_b[p]
        cf[k] = np.float32(r[3])
    return a, b, A, B, cf


def write_nb14(path, atom_a, atom_b, lj_scale, cf_scale):
    """nb14 file of pairs with an LJ scale (the form read_nb14 resolves)."""
    with open(path, "w") as f:
        f.write("%d\n" % len(atom_a))
        for a, b, s, c in zip(atom_a, atom_b, lj_scale, cf_scale):
            f.write("%d %d %.6f %.6f\n" % (int(a), int(b), float(s), float(c)))


def read_bond(path):
    """Harmonic bond list as Xponge::Native_Load_Bonds reads it
    (xponge/load/native/bond.hpp:8-40): a count, then "atom_a atom_b k r0"
    rows (E = k (r - r0)^2).  Returns (atom_a, atom_b, k, r0)."""
    t = _tokens(path)
    m = int(t[0])
    if len(t) < 1 + 4 * m:
        raise ValueError("%s: the format of bond_in_file is not right" % path)
    rows = np.array(t[1:1 + 4 * m], dtype=object).reshape(m, 4)
    return (rows[:, 0].astype(np.int32), rows[:, 1].astype(np.int32),
            rows[:, 2].astype(np.float32), rows[:, 3].astype(np.float32))


def read_urey_bradley(path):
    """Urey-Bradley terms as Xponge::Native_Load_Urey_Bradley reads them
    (xponge/load/native/urey_bradley.hpp:8-52): a count, then
    "a b c angle_k angle_theta0 bond_k bond_r0" rows; the module is an angle
    a-b-c plus a bond a-c (Urey_Bradley_force.cpp:56-72).  Returns
    ((a, b, c, angle_k, theta0), (a, c, bond_k, bond_r0))."""
    t = _tokens(path)
    m = int(t[0])
    if len(t) < 1 + 7 * m:
        raise ValueError("%s: the format of urey_bradley_in_file is not right" % path)
    rows = np.array(t[1:1 + 7 * m], dtype=object).reshape(m, 7)
    a, b, c = (rows[:, i].astype(np.int32) for i in range(3))
    ak, at0, bk, br0 = (rows[:, i].astype(np.float32) for i in range(3, 7))
    return (a, b, c, ak, at0), (a, c, bk, br0)


def write_bond(path, atom_a, atom_b, k, r0):
    with open(path, "w") as f:
        f.write("%d\n" % len(atom_a))
        for a, b, kk, r in zip(atom_a, atom_b, k, r0):
            f.write("%d %d %.9g %.9g\n" % (int(a), int(b), float(kk), float(r)))


def box_to_cell(box, angle=(90.0, 90.0, 90.0)):
    """MD_core/pbc.hpp:66-101 in float32: returns (cell, rcell) LTMatrix3."""
    f = np.float32
    d2r = f(0.0174532925)
    a, b, c = (f(x) for x in box)
    al, be, ga = (f(x) * d2r for x in angle)
    cal, cbe, cga = np.cos(al), np.cos(be), np.cos(ga)
    sga = np.sin(ga)
    za = np.sqrt(f(1) - cal * cal - cbe * cbe - cga * cga + f(2) * cal * cbe * cga)
    cell = np.array([a, b * cga, b * sga, c * cbe, c / sga * (cal - cbe * cga),
                     c / sga * za], dtype=np.float32)
    rcell = np.zeros(6, np.float32)
    rcell[0] = f(1) / cell[0]
    rcell[2] = f(1) / cell[2]
    rcell[5] = f(1) / cell[5]
    rcell[1] = -rcell[0] / np.tan(ga)
    rcell[3] = (cal / np.tan(ga) - cbe / sga) / za / a
    rcell[4] = (cbe / np.tan(ga) - cal / sga) / za / b
    for m in (cell, rcell):
        for k in (1, 3, 4):
            if abs(m[k]) < 1e-3:
                m[k] = 0.0
    return cell, rcell


def system_from_prefix(directory, prefix):
    """Loads ``<prefix>_{LJ,charge,coordinate,exclude}.txt`` (and mass when
    present) into the dict the path uses."""
    p = lambda name: os.path.join(directory, "%s_%s.txt" % (prefix, name))
    n, nt, a, b, types = read_lj(p("LJ"))
    crd, box, ang = read_coordinate(p("coordinate"))
    charge = read_scalar_list(p("charge"))
    es, en, el = read_exclude(p("exclude"))
    cell, rcell = box_to_cell(box, ang)
    d = dict(n=n, crd=crd, charge=charge, lj_ntypes=nt, lj_a=a, lj_b=b, lj_type=types,
             box=box, angle=ang, cell=cell, rcell=rcell, excl_start=es,
             excl_numbers=en, excl_list=el)
    if os.path.exists(p("mass")):
        d["mass"] = read_scalar_list(p("mass"))
    if os.path.exists(p("nb14")) or os.path.exists(p("nb14_extra")):
        d["nb14"] = read_nb14(p("nb14") if os.path.exists(p("nb14")) else None,
                              p("nb14_extra") if os.path.exists(p("nb14_extra")) else None,
                              types, a, b)
    if os.path.exists(p("bond")):
        d["bonds"] = read_bond(p("bond"))
    if os.path.exists(p("urey_bradley")):
        d["urey_bradley"] = read_urey_bradley(p("urey_bradley"))
    return d


# ---------------------------------------------------------------------------
# writers (used by the synthetic-system generators)
# ---------------------------------------------------------------------------
def write_system(directory, prefix, sysd):
    os.makedirs(directory, exist_ok=True)
    p = lambda name: os.path.join(directory, "%s_%s.txt" % (prefix, name))
    n = int(sysd["n"])
    nt = int(sysd["lj_ntypes"])
    with open(p("LJ"), "w") as f:
        f.write("%d %d\n" % (n, nt))
        np.savetxt(f, (sysd["lj_a"] / np.float32(12.0)).reshape(-1, 1), fmt="%.9e")
        np.savetxt(f, (sysd["lj_b"] / np.float32(6.0)).reshape(-1, 1), fmt="%.9e")
        np.savetxt(f, sysd["lj_type"].reshape(-1, 1), fmt="%d")
    with open(p("charge"), "w") as f:
        f.write("%d\n" % n)
        np.savetxt(f, sysd["charge"].reshape(-1, 1), fmt="%.9e")
    if "mass" in sysd:
        with open(p("mass"), "w") as f:
            f.write("%d\n" % n)
            np.savetxt(f, sysd["mass"].reshape(-1, 1), fmt="%.6f")
    with open(p("coordinate"), "w") as f:
        f.write("%d\n" % n)
        np.savetxt(f, sysd["crd"], fmt="%.7f")
        f.write("%.7f %.7f %.7f\n" % tuple(sysd["box"]))
        f.write("%.7f %.7f %.7f\n" % tuple(sysd.get("angle", (90.0, 90.0, 90.0))))
    if "bonds" in sysd:
        write_bond(p("bond"), *sysd["bonds"])
    if "nb14" in sysd:
        a14, b14, A14, B14, cf14 = sysd["nb14"]
        with open(p("nb14_extra"), "w") as f:
            f.write("%d\n" % len(a14))
            for k in range(len(a14)):
                f.write("%d %d %.9e %.9e %.6f\n" % (int(a14[k]), int(b14[k]),
                                                    float(A14[k] / np.float32(12.0)),
                                                    float(B14[k] / np.float32(6.0)),
                                                    float(cf14[k])))
    with open(p("exclude"), "w") as f:
        es, en, el = sysd["excl_start"], sysd["excl_numbers"], sysd["excl_list"]
        f.write("%d %d\n" % (n, int(en.sum())))
        for i in range(n):
            js = el[es[i]:es[i] + en[i]]
            f.write(" ".join([str(int(en[i]))] + [str(int(j)) for j in js]) + "\n")
