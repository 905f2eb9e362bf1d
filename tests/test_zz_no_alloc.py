"""The allocation-free exclusion correction (snb_pme_excluded_force_noalloc)
against the context entry point (snb_pme_excluded_force), through the C ABI:
forces, per-atom energies and per-atom virials bit for bit, on two calls in a
row, the second with a smaller system; sizes above the device array's
capacity are refused before any GPU work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _context_result(sysd):
    from sponge_amd.lib import Nonbonded
    nb = Nonbonded(sysd)
    nb.zero_outputs()
    nb.excluded(need_energy=True, need_virial=True, frc=True)
    nb.sync()
    out = (nb.frc.cpu().numpy().copy(), nb.atom_energy.cpu().numpy().copy(),
           nb.atom_virial.cpu().numpy().copy(), nb.beta)
    nb.close()
    return out


def _noalloc_result(sysd, beta):
    import torch
    from sponge_amd import lib
    L = lib.load()
    dev = torch.device("cuda", 0)
    n = int(sysd["n"])
    crd = torch.as_tensor(np.asarray(sysd["crd"], np.float32)).to(dev).contiguous()
    q = torch.as_tensor(np.asarray(sysd["charge"], np.float32)).to(dev)
    es = torch.as_tensor(np.asarray(sysd["excl_start"], np.int32)).to(dev)
    el = torch.as_tensor(np.asarray(sysd["excl_list"], np.int32)).to(dev)
    en = torch.as_tensor(np.asarray(sysd["excl_numbers"], np.int32)).to(dev)
    frc = torch.zeros((n, 3), dtype=torch.float32, device=dev)
    ae = torch.zeros(n, dtype=torch.float32, device=dev)
    av = torch.zeros((n, 6), dtype=torch.float32, device=dev)
    cell = lib.LTMatrix3.of(sysd["cell"])
    rcell = lib.LTMatrix3.of(sysd["rcell"])
    stream = torch.cuda.current_stream(dev)
    p = lib._ptr
    lib.check(L.snb_pme_excluded_force_noalloc(n, p(crd), cell, rcell, p(q), beta, p(es), p(el),
                                               p(en), p(frc), 1, p(ae), 1, p(av),
                                               ctypes.c_void_p(stream.cuda_stream)),
              "snb_pme_excluded_force_noalloc")
    torch.cuda.synchronize()
    return frc.cpu().numpy(), ae.cpu().numpy(), av.cpu().numpy()


def test_noalloc_exclusion_matches_context_entry_point(water_9k, small_water):
    results = []
    for name, s in (("water_9k", water_9k), ("small_water", small_water)):
        print("case", name, s["n"], "atoms", flush=True)
        f, e, v, beta = _context_result(s)
        g, ge, gv = _noalloc_result(s, beta)
        np.testing.assert_array_equal(g, f)
        np.testing.assert_array_equal(ge, e)
        np.testing.assert_array_equal(gv, v)
        assert np.any(f != 0) and np.any(e != 0)
        results.append(g)
    assert results[1].shape[0] < results[0].shape[0]


def test_noalloc_rejects_sizes_above_capacity():
    import torch
    from sponge_amd import lib
    L = lib.load()
    big = 131073
    rc = L.snb_pme_excluded_force_noalloc(big, None, lib.LTMatrix3(), lib.LTMatrix3(), None,
                                          0.3, None, None, None, None, 0, None, 0, None, None)
    assert rc == 8
    assert b"SNB_NOALLOC_MAX_ATOMS" in L.snb_status_string(rc)
    torch.cuda.synchronize()
