"""CPU-side checks of the product: the C-ABI library loads and exports every
symbol the header declares, the host-side generators and the native-format
I/O (no GPU calls)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT


def test_library_exports_every_declared_symbol():
    from sponge_amd import lib
    assert os.path.exists(lib.LIB_PATH), "build the library first (__graft_entry__.build())"
    L = lib.load()
    declared = lib.header_symbols()
    assert len(declared) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", lib.LIB_PATH], check=True,
                         capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    assert [s for s in declared if s not in exported] == []
    for s in declared:
        assert hasattr(L, s)
    assert L.snb_abi_version() == 1
    assert L.snb_status_string(3) == b"the box is too small for the neighbor-list cell grid"


def test_binding_signatures_cover_the_header():
    from sponge_amd import lib
    assert sorted(lib._SIGS) == lib.header_symbols()


def test_native_io_round_trip(tmp_path, tip3p_1011):
    from sponge_amd import native_io
    native_io.write_system(str(tmp_path), "rt", tip3p_1011)
    back = native_io.system_from_prefix(str(tmp_path), "rt")
    for k in ("crd", "charge", "lj_a", "lj_b"):
        np.testing.assert_allclose(back[k], tip3p_1011[k], rtol=2e-7, atol=1e-6)
    for k in ("lj_type", "excl_start", "excl_numbers", "excl_list"):
        np.testing.assert_array_equal(back[k], tip3p_1011[k])
    np.testing.assert_array_equal(back["cell"], tip3p_1011["cell"])


def test_box_to_cell_orthorhombic_is_diagonal():
    from sponge_amd import native_io
    cell, rcell = native_io.box_to_cell([96.5, 96.5, 96.5])
    np.testing.assert_array_equal(cell, np.array([96.5, 0, 96.5, 0, 0, 96.5], np.float32))
    assert rcell[0] == np.float32(1) / np.float32(96.5)
    assert rcell[1] == rcell[3] == rcell[4] == 0


def test_fixture_loader_matches_reference_counts(gromacs_iter0):
    s = gromacs_iter0
    assert s["n"] == 2129 and s["lj_ntypes"] == 10
    assert int(s["excl_numbers"].sum()) == 2216          # header of sys_flexible_exclude.txt
    assert s["excl_numbers"][0] == 12 and list(s["excl_list"][:3]) == [1, 2, 3]
    np.testing.assert_array_equal(s["box"], np.float32([70, 70, 70]))


def test_tip3p_generator_geometry(water_9k):
    s = water_9k
    assert s["n"] == 9000
    crd = s["crd"].astype(np.float64).reshape(-1, 3, 3)
    L = 44.8
    d1 = crd[:, 1] - crd[:, 0]
    d2 = crd[:, 2] - crd[:, 0]
    d1 -= np.round(d1 / L) * L
    d2 -= np.round(d2 / L) * L
    np.testing.assert_allclose(np.linalg.norm(d1, axis=1), 0.9572, atol=2e-5)
    cosang = np.einsum("ij,ij->i", d1, d2) / (np.linalg.norm(d1, axis=1) * np.linalg.norm(d2, axis=1))
    np.testing.assert_allclose(np.degrees(np.arccos(cosang)), 104.52, atol=1e-2)
    assert abs(s["charge"].sum()) < 1e-2
    assert int(s["excl_numbers"].sum()) == 3 * 3000


def test_mixed_system_generator_small():
    from sponge_amd import systems
    s = systems.mixed_system(n_waters=900, n_chains=2, beads_per_chain=60, box=32.0, seed=3)
    assert s["n"] == 2700 + 120
    assert s["lj_ntypes"] == 2 + 120
    assert abs(s["charge"][2700:].sum()) < 1e-3
    assert s["lj_a"].size == s["lj_ntypes"] * (s["lj_ntypes"] + 1) // 2
    b = s["crd"][2700:2760].astype(np.float64)
    d = np.diff(b, axis=0)
    d -= np.round(d / 32.0) * 32.0
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.53, atol=2e-4)
    # bead i excludes i+1..i+3 within its chain
    i = 2700
    st, nm = s["excl_start"][i], s["excl_numbers"][i]
    assert list(s["excl_list"][st:st + nm]) == [2701, 2702, 2703]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run(["python3", os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
