"""ctypes binding of libsponge_nb.so (C ABI: include/sponge_nb.h) plus a
small host driver of the nonbonded path on torch-allocated device memory.

torch is used only for device allocation and the stream handle; all compute
runs in the library's HIP kernels.  If the library is missing, loading fails
with an error -- there is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.abspath(os.path.join(_HERE, "..", ".."))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "libsponge_nb.so")
HEADER_PATH = os.path.abspath(os.path.join(PKG_ROOT, "..", "include", "sponge_nb.h"))
FIXED_SCALE = 2.0 ** 32

SNB_OK = 0
STATUS_NAMES = {1: "invalid argument", 2: "HIP runtime error", 3: "box too small",
                4: "unsupported configuration", 5: "module not initialised",
                6: "device allocation failed", 7: "device status flag set",
                8: "above the allocation-free entry point's capacity"}
FLAG_NL_OVERFLOW, FLAG_CELL_OVERFLOW, FLAG_NONFINITE_CRD = 1, 2, 4


class SnbError(RuntimeError):
    def __init__(self, code, where):
        super().__init__("%s failed: %s (%d)" % (where, STATUS_NAMES.get(code, "?"), code))
        self.code = code


class LTMatrix3(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("a11", "a21", "a22", "a31", "a32", "a33")]

    @classmethod
    def of(cls, a6):
        return cls(*[float(x) for x in np.asarray(a6, np.float32)])


class Vector(ctypes.Structure):
    _fields_ = [("x", ctypes.c_float), ("y", ctypes.c_float), ("z", ctypes.c_float)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_F = ctypes.c_float
_SIGS = {
    "snb_abi_version": ([], _I),
    "snb_status_string": ([_I], ctypes.c_char_p),
    "snb_context_create": ([ctypes.POINTER(_P), _I, _P, _I], _I),
    "snb_context_destroy": ([_P], _I),
    "snb_set_stream": ([_P, _P], _I),
    "snb_set_pme_overlap": ([_P, _I], _I),
    "snb_synchronize": ([_P], _I),
    "snb_read_status": ([_P, _I, ctypes.POINTER(_I)], _I),
    "snb_set_exclusions": ([_P, _P, _P, _P, _I], _I),
    "snb_nl_initial": ([_P, _F, _F, LTMatrix3, LTMatrix3, _I, _I, _I, _F], _I),
    "snb_nl_update": ([_P, _P, _I, _I, _P, LTMatrix3, LTMatrix3, _I, _I, _P, _P, _P], _I),
    "snb_nl_device_list": ([_P, ctypes.POINTER(_P), ctypes.POINTER(_P), ctypes.POINTER(_I)], _I),
    "snb_nl_copy_rows": ([_P, _P, _P], _I),
    "snb_nl_set_export": ([_P, _I], _I),
    "snb_lj_set_owner_attribution": ([_P, _I], _I),
    "snb_nl_copy_internal": ([_P, _P, _P], _I),
    "snb_nl_window_info": ([_P] + [ctypes.POINTER(_I)] * 6, _I),
    "snb_nl_set_window_capacity": ([_P, _I], _I),
    "snb_nl_set_window_shape": ([_P, _I, _I, _I], _I),
    "snb_lj_set_direct_j": ([_P, _I], _I),
    "snb_lj_set_row_stream": ([_P, _I], _I),
    "snb_lj_set_row_lds": ([_P, _I], _I),
    "snb_nl_grid_shape": ([_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "snb_lj_initial": ([_P, _I, _P, _P, _P, _F], _I),
    "snb_lj_pme_direct_force": ([_P, _I, _I, _I, _I, _P, _P, _P, LTMatrix3, LTMatrix3, _P, _F,
                                 _I, _P, _I, _P, _P], _I),
    "snb_lj_long_range_correction": ([_P, _I, _P, _I, _P, _F], _I),
    "snb_lj_energy": ([_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)], _I),
    "snb_nb14_initial": ([_P, _I, _P, _P, _P, _P, _P], _I),
    "snb_nb14_force": ([_P, _P, _P, LTMatrix3, LTMatrix3, _P, _I, _P, _I, _P], _I),
    "snb_nb14_energy": ([_P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)],
                        _I),
    "snb_bond_initial": ([_P, _I, _P, _P, _P, _P], _I),
    "snb_bond_force": ([_P, _P, LTMatrix3, LTMatrix3, _P, _I, _P, _I, _P], _I),
    "snb_bond_energy": ([_P, ctypes.POINTER(ctypes.c_double)], _I),
    "snb_leap_frog": ([_P, _P, _P, _P, _P, _F, _F], _I),
    "snb_langevin_initial": ([_P, _P, _F, _F, _I, _F, _F], _I),
    "snb_langevin_iterate": ([_P, _P, _P, _P, _P], _I),
    "snb_langevin_parameters": ([_P] + [ctypes.POINTER(_F)] * 4, _I),
    "snb_settle_initial": ([_P, _I, _P, _P, _P, _P, _P, _P, _P], _I),
    "snb_settle_pair_initial": ([_P, _I, _P, _P, _P, _P], _I),
    "snb_settle_remember": ([_P, _P, LTMatrix3, LTMatrix3], _I),
    "snb_settle_do": ([_P, _P, _P, LTMatrix3, LTMatrix3, _P, _I, _P, _F, _F, _F], _I),
    "snb_md_iterate_fused": ([_P, _P, _P, LTMatrix3, LTMatrix3, _P, _I, _P], _I),
    "snb_get_potential_to_stress": ([_P, _I, _P, LTMatrix3, _P, _P], _I),
    "snb_get_kinetic_to_stress": ([_P, _I, _P, _P, LTMatrix3, _P], _I),
    "snb_get_pressure": ([_P, _I, _P, _P, _P, LTMatrix3, _P, _P, ctypes.POINTER(LTMatrix3),
                          ctypes.POINTER(_F)], _I),
    "snb_pme_initial": ([_P, LTMatrix3, LTMatrix3, Vector, _F, _I, _F, _I, _I, _I, _F, _I], _I),
    "snb_pme_update_box": ([_P, LTMatrix3, LTMatrix3, LTMatrix3, _F], _I),
    "snb_pme_parameters": ([_P, ctypes.POINTER(_F), ctypes.POINTER(_I), ctypes.POINTER(_I),
                            ctypes.POINTER(_I)], _I),
    "snb_pme_excluded_force": ([_P, _I, _P, _P, _P, LTMatrix3, LTMatrix3, _P, _P, _P, _P, _P,
                                _I, _P, _I, _P], _I),
    "snb_pme_excluded_force_noalloc": ([_I, _P, LTMatrix3, LTMatrix3, _P, _F, _P, _P, _P, _P,
                                        _I, _P, _I, _P, _P], _I),
    "snb_pme_reciprocal_force": ([_P, _P, LTMatrix3, LTMatrix3, _P, _P, _I, _I, _P, _P, _I], _I),
    "snb_pme_energy": ([_P] + [ctypes.POINTER(ctypes.c_double)] * 4, _I),
    "snb_pme_reciprocal_virial": ([_P, _P], _I),
    "snb_flush_forces": ([_P, _P], _I),
    "snb_flush_atom_energy": ([_P, _P, _P], _I),
    "snb_flush_atom_virial": ([_P, _P], _I),
    "snb_copy_fixed_forces": ([_P, _P], _I),
    "snb_clear_fixed_forces": ([_P], _I),
    "snb_copy_fixed_atom_terms": ([_P, _I, _P], _I),
    "snb_copy_pme_grid": ([_P, _I, _P], _I),
    "snb_fft_r2c": ([_P, _P, _P], _I),
    "snb_fft_c2r": ([_P, _P, _P], _I),
    "snb_f64_force": ([_P, _P, _P, LTMatrix3, LTMatrix3, _I, _I, _P, _I, _I, _P, _P], _I),
    "snb_f64_nve_iterate": ([_P, _P, _P, _P, _P, _P, LTMatrix3, LTMatrix3, ctypes.c_double],
                            _I),
    "snb_batch_create": ([ctypes.POINTER(_P), _I], _I),
    "snb_batch_destroy": ([_P], _I),
    "snb_batch_record_begin": ([_P, _P], _I),
    "snb_batch_record_end": ([_P], _I),
    "snb_batch_plan": ([_P, ctypes.POINTER(_I), ctypes.POINTER(_I)], _I),
    "snb_batch_launch": ([_P, _P], _I),
    "snb_batch_info": ([_P] + [ctypes.POINTER(_I)] * 4, _I),
    "snb_profile_enable": ([_P, _I], _I),
    "snb_profile_read": ([_P, _I, _I, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(_I)], _I),
}
PROF_PAIR, PROF_NL, PROF_PME, PROF_EXCL = 0, 1, 2, 3

_lib = None


def load():
    """Loads the in-tree library; raises if it has not been built."""
    global _lib
    if _lib is None:
        # torch's HIP runtime (SONAME libamdhip64.so.7) must be the one the
        # library binds to: import torch before dlopen so there is one runtime.
        import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libsponge_nb.so is not built (%s); run "
                               "__graft_entry__.build() or make -C sponge2.0_amd" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(rc, where):
    if rc != SNB_OK:
        raise SnbError(rc, where)


def header_symbols():
    """Names of the functions declared in include/sponge_nb.h."""
    import re
    txt = open(HEADER_PATH).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(snb_\w+)\s*\(", txt, re.M)))


def _ptr(t):
    if t is None:
        return None
    return ctypes.c_void_p(t.data_ptr())


class Nonbonded:
    """One system on one GPU: the three reference modules behind one context.

    sysd is a system dict (sponge_amd.systems / native_io.system_from_prefix).
    Device arrays are torch tensors on `device`; the caller-owned buffers of
    the reference (crd, charge, frc, per-atom energy and virial) are held here
    so tests and the bench can drive the path like the reference's main loop.
    """

    def __init__(self, sysd, cutoff=None, skin=2.0, tolerance=1e-5, grid=None,
                 max_neighbor_numbers=1200, max_atom_in_grid_numbers=150,
                 refresh_interval=0, skin_permit=0.5, update_interval=1, device=0,
                 stream=None):
        import torch

        self.torch = torch
        L = load()
        self.L = L
        self.n = int(sysd["n"])
        self.dev = torch.device("cuda", device)
        self.stream = stream if stream is not None else torch.cuda.current_stream(self.dev)
        self.cutoff = float(sysd.get("cutoff", 10.0) if cutoff is None else cutoff)
        self.cell = LTMatrix3.of(sysd["cell"])
        self.rcell = LTMatrix3.of(sysd["rcell"])
        self.box = np.asarray(sysd["box"], np.float32)
        self.angle = tuple(float(x) for x in sysd.get("angle", (90.0, 90.0, 90.0)))
        ctx = ctypes.c_void_p()
        check(L.snb_context_create(ctypes.byref(ctx), device,
                                   ctypes.c_void_p(self.stream.cuda_stream), self.n),
              "snb_context_create")
        self.ctx = ctx
        es = np.ascontiguousarray(sysd["excl_start"], np.int32)
        en = np.ascontiguousarray(sysd["excl_numbers"], np.int32)
        el = np.ascontiguousarray(sysd["excl_list"], np.int32)
        if el.size == 0:
            el = np.zeros(1, np.int32)
        check(L.snb_set_exclusions(ctx, es.ctypes.data, el.ctypes.data, en.ctypes.data,
                                   int(en.sum())), "snb_set_exclusions")
        check(L.snb_nl_initial(ctx, self.cutoff, skin, self.cell, self.rcell,
                               max_neighbor_numbers, max_atom_in_grid_numbers,
                               refresh_interval, skin_permit), "snb_nl_initial")
        a = np.ascontiguousarray(sysd["lj_a"], np.float32)
        b = np.ascontiguousarray(sysd["lj_b"], np.float32)
        t = np.ascontiguousarray(sysd["lj_type"], np.int32)
        check(L.snb_lj_initial(ctx, int(sysd["lj_ntypes"]), a.ctypes.data, b.ctypes.data,
                               t.ctypes.data, self.cutoff), "snb_lj_initial")
        self.nb14 = None
        if "nb14" in sysd and len(sysd["nb14"][0]) > 0:
            a14, b14, A14, B14, cf14 = (np.ascontiguousarray(x, dt) for x, dt in
                                        zip(sysd["nb14"], (np.int32, np.int32, np.float32,
                                                           np.float32, np.float32)))
            check(L.snb_nb14_initial(ctx, len(a14), a14.ctypes.data, b14.ctypes.data,
                                     A14.ctypes.data, B14.ctypes.data, cf14.ctypes.data),
                  "snb_nb14_initial")
            self.nb14 = (a14, b14, A14, B14, cf14)
        self.bonds = None
        if "bonds" in sysd and len(sysd["bonds"][0]) > 0:
            ba, bb, bk, br = (np.ascontiguousarray(x, dt) for x, dt in
                              zip(sysd["bonds"], (np.int32, np.int32, np.float32, np.float32)))
            check(L.snb_bond_initial(ctx, len(ba), ba.ctypes.data, bb.ctypes.data,
                                     bk.ctypes.data, br.ctypes.data), "snb_bond_initial")
            self.bonds = (ba, bb, bk, br)
        g = grid if grid is not None else sysd.get("pme_grid", (0, 0, 0))
        check(L.snb_pme_initial(ctx, self.cell, self.rcell, Vector(*[float(x) for x in self.box]),
                                self.cutoff, 0, tolerance, int(g[0]), int(g[1]), int(g[2]), 1.0,
                                update_interval), "snb_pme_initial")
        beta = ctypes.c_float()
        fx, fy, fz = _I(), _I(), _I()
        check(L.snb_pme_parameters(ctx, ctypes.byref(beta), ctypes.byref(fx), ctypes.byref(fy),
                                   ctypes.byref(fz)), "snb_pme_parameters")
        self.beta = beta.value
        self.grid = (fx.value, fy.value, fz.value)
        self.volume = float(np.float32(sysd["cell"][0]) * np.float32(sysd["cell"][2]) *
                            np.float32(sysd["cell"][5]))
        nl = ctypes.c_void_p()
        ser = ctypes.c_void_p()
        mx = _I()
        check(L.snb_nl_device_list(ctx, ctypes.byref(nl), ctypes.byref(ser), ctypes.byref(mx)),
              "snb_nl_device_list")
        self.d_nl, self.d_serial, self.max_nb = nl, ser, mx.value
        n = self.n
        f32 = torch.float32
        self.crd = torch.as_tensor(np.asarray(sysd["crd"], np.float32)).to(self.dev).contiguous()
        self.charge = torch.as_tensor(np.asarray(sysd["charge"], np.float32)).to(self.dev)
        self.frc = torch.zeros((n, 3), dtype=f32, device=self.dev)
        self.atom_energy = torch.zeros(n, dtype=f32, device=self.dev)
        self.direct_energy = torch.zeros(n, dtype=f32, device=self.dev)
        self.atom_virial = torch.zeros((n, 6), dtype=f32, device=self.dev)
        self.virial = torch.zeros(6, dtype=f32, device=self.dev)   # d_virial[0]
        self.potential = torch.zeros(1, dtype=f32, device=self.dev)

    def close(self):
        if self.ctx is not None:
            self.L.snb_context_destroy(self.ctx)
            self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference module calls --------------------------------------------
    def nl_update(self, step=0, forced=True):
        check(self.L.snb_nl_update(self.ctx, None, self.n, 0, _ptr(self.crd), self.cell,
                                   self.rcell, step, 1 if forced else 0, None, None, None),
              "snb_nl_update")

    def lj_direct(self, need_energy=False, need_virial=False, frc=True):
        check(self.L.snb_lj_pme_direct_force(
            self.ctx, self.n, self.n, 0, 0, _ptr(self.crd), _ptr(self.charge),
            _ptr(self.frc) if frc else None, self.cell, self.rcell, self.d_nl, self.beta,
            1 if need_energy else 0, _ptr(self.atom_energy) if frc else None,
            1 if need_virial else 0, _ptr(self.atom_virial) if frc else None,
            _ptr(self.direct_energy) if frc else None), "snb_lj_pme_direct_force")

    def excluded(self, need_energy=False, need_virial=False, frc=True):
        check(self.L.snb_pme_excluded_force(
            self.ctx, self.n, None, None, _ptr(self.crd), self.cell, self.rcell,
            _ptr(self.charge), None, None, None, _ptr(self.frc) if frc else None,
            1 if need_energy else 0, _ptr(self.atom_energy) if frc else None,
            1 if need_virial else 0, _ptr(self.atom_virial) if frc else None),
            "snb_pme_excluded_force")

    def nb14_force(self, need_energy=False, need_virial=False, frc=True):
        check(self.L.snb_nb14_force(
            self.ctx, _ptr(self.crd), _ptr(self.charge), self.cell, self.rcell,
            _ptr(self.frc) if frc else None, 1 if need_energy else 0,
            _ptr(self.atom_energy) if frc else None, 1 if need_virial else 0,
            _ptr(self.atom_virial) if frc else None), "snb_nb14_force")

    def nb14_energy(self):
        lj, ee = ctypes.c_double(), ctypes.c_double()
        check(self.L.snb_nb14_energy(self.ctx, ctypes.byref(lj), ctypes.byref(ee)),
              "snb_nb14_energy")
        return lj.value, ee.value

    def bond_force(self, need_energy=False, need_virial=False, frc=True):
        check(self.L.snb_bond_force(
            self.ctx, _ptr(self.crd), self.cell, self.rcell, _ptr(self.frc) if frc else None,
            1 if need_energy else 0, _ptr(self.atom_energy) if frc else None,
            1 if need_virial else 0, _ptr(self.atom_virial) if frc else None),
            "snb_bond_force")

    def bond_energy(self):
        e = ctypes.c_double()
        check(self.L.snb_bond_energy(self.ctx, ctypes.byref(e)), "snb_bond_energy")
        return e.value

    def reciprocal(self, step=0, need_energy=False, need_virial=False, frc=True):
        check(self.L.snb_pme_reciprocal_force(
            self.ctx, _ptr(self.crd), self.cell, self.rcell, _ptr(self.charge),
            _ptr(self.frc) if frc else None, 1 if need_virial else 0, 1 if need_energy else 0,
            _ptr(self.virial), _ptr(self.potential), step), "snb_pme_reciprocal_force")

    def long_range(self, need_energy=False, need_virial=False):
        check(self.L.snb_lj_long_range_correction(
            self.ctx, 1 if need_virial else 0, _ptr(self.virial), 1 if need_energy else 0,
            _ptr(self.potential), self.volume), "snb_lj_long_range_correction")

    def flush(self):
        check(self.L.snb_flush_forces(self.ctx, _ptr(self.frc)), "snb_flush_forces")

    def zero_outputs(self):
        self.frc.zero_()
        self.atom_energy.zero_()
        self.atom_virial.zero_()
        self.virial.zero_()
        self.potential.zero_()

    def set_pme_overlap(self, enable):
        """PME reciprocal pipeline on the library's second stream (default on)."""
        check(self.L.snb_set_pme_overlap(self.ctx, 1 if enable else 0), "snb_set_pme_overlap")

    def set_stream(self, stream):
        """Points the library context at another torch stream."""
        check(self.L.snb_set_stream(self.ctx, ctypes.c_void_p(stream.cuda_stream)),
              "snb_set_stream")
        self.stream = stream

    def capture_step(self, **step_kwargs):
        """Records one step() (same keyword arguments) as a HIP graph.

        Capture runs on a side stream that the context is pointed at for the
        duration; the graph is replayed with graph.replay() on torch's current
        stream, which is the context's stream.  Host-side decisions inside the
        step (neighbour-list rebuild or not, energy or not) are fixed at
        capture: record one graph per kind of step.  The library issues no
        memset nodes (its clears are kernels)."""
        torch = self.torch
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=self.dev)
        home = self.stream
        self.sync()  # joins the reciprocal stream: nothing outside the capture is pending
        torch.cuda.synchronize()
        self.set_stream(side)
        try:
            with torch.cuda.graph(graph, stream=side):
                self.step(**step_kwargs)
        finally:
            self.set_stream(home)
        torch.cuda.synchronize()
        return graph

    def step(self, step=0, need_energy=False, need_virial=False, forced_nl=False, fused=True,
             flush=True):
        """One force evaluation of the path in the reference's order
        (main.cpp:307-334, :371-396, :439-441, :455-458, :474-478).  fused=True keeps every term
        in the fixed-point accumulators and converts once at the end."""
        check(self.L.snb_nl_update(self.ctx, None, self.n, 0, _ptr(self.crd), self.cell,
                                   self.rcell, step, 1 if forced_nl else 0, None, None, None),
              "snb_nl_update")
        self.excluded(need_energy, need_virial, frc=not fused)
        self.lj_direct(need_energy, need_virial, frc=not fused)
        if need_energy or need_virial:
            self.long_range(need_energy, need_virial)
        if self.nb14 is not None:
            self.nb14_force(need_energy, need_virial, frc=not fused)
        if self.bonds is not None:
            self.bond_force(need_energy, need_virial, frc=not fused)
        self.reciprocal(step, need_energy, need_virial, frc=not fused)
        if fused:
            if flush:
                self.flush()
            if need_energy:
                check(self.L.snb_flush_atom_energy(self.ctx, _ptr(self.atom_energy),
                                                   _ptr(self.direct_energy)),
                      "snb_flush_atom_energy")
            if need_virial:
                check(self.L.snb_flush_atom_virial(self.ctx, _ptr(self.atom_virial)),
                      "snb_flush_atom_virial")

    # -- integration (leap-frog / middle Langevin + SETTLE) --------------------
    def md_initial(self, sysd, temperature=300.0, tau=1.0, seed=20260217, dt=0.002,
                   velocities=None, settle=True):
        """Sets up the device integrator of Main_Iteration (main.cpp:622-712):
        masses, Maxwell velocities (or the given ones), the middle Langevin
        thermostat (tau in ps, dt in ps) and SETTLE on the rigid waters."""
        from . import systems
        torch = self.torch
        mass = np.ascontiguousarray(sysd["mass"], np.float32)
        self.mass = torch.as_tensor(mass).to(self.dev)
        inv = np.where(mass > 0, 1.0 / np.where(mass > 0, mass, 1.0), 0.0).astype(np.float32)
        self.inv_mass = torch.as_tensor(inv).to(self.dev)
        v = systems.maxwell_velocities(mass, temperature, seed) if velocities is None \
            else np.asarray(velocities, np.float32)
        self.vel = torch.as_tensor(v).to(self.dev).contiguous()
        self.stress = torch.zeros(6, dtype=torch.float32, device=self.dev)
        self.virial_tensor = torch.zeros(6, dtype=torch.float32, device=self.dev)
        check(self.L.snb_langevin_initial(self.ctx, mass.ctypes.data, temperature, tau, seed, dt,
                                          0.0), "snb_langevin_initial")
        d, e, vf, xf = _F(), _F(), _F(), _F()
        check(self.L.snb_langevin_parameters(self.ctx, ctypes.byref(d), ctypes.byref(e),
                                             ctypes.byref(vf), ctypes.byref(xf)),
              "snb_langevin_parameters")
        self.dt_internal, self.exp_gamma = d.value, e.value
        self.v_factor, self.x_factor = vf.value, xf.value
        self.tri = None
        if settle and "n_waters" in sysd:
            a, b, c, rab, rac, rbc = systems.water_triangles(sysd)
            check(self.L.snb_settle_initial(self.ctx, len(a), a.ctypes.data, b.ctypes.data,
                                            c.ctypes.data, rab.ctypes.data, rac.ctypes.data,
                                            rbc.ctypes.data, mass.ctypes.data),
                  "snb_settle_initial")
            self.tri = (a, b, c, rab, rac, rbc)

    def md_iterate(self, need_pressure=False):
        """Remember (SETTLE), Langevin leap-frog, SETTLE: the iteration half
        of a step, after the forces (main.cpp:634-712)."""
        check(self.L.snb_settle_remember(self.ctx, _ptr(self.crd), self.cell, self.rcell),
              "snb_settle_remember")
        check(self.L.snb_langevin_iterate(self.ctx, _ptr(self.frc), _ptr(self.vel), None,
                                          _ptr(self.crd)), "snb_langevin_iterate")
        check(self.L.snb_settle_do(self.ctx, _ptr(self.mass), _ptr(self.crd), self.cell,
                                   self.rcell, _ptr(self.vel), 1 if need_pressure else 0,
                                   _ptr(self.stress), self.dt_internal, self.v_factor,
                                   self.x_factor), "snb_settle_do")

    def md_step(self, step=0, need_energy=False, need_virial=False, forced_nl=False,
                fused_iterate=True):
        """One MD step on the device: zero the forces (and, on a pressure step,
        the virials and the stress: MD_Reset_Atom_Energy_And_Virial_And_Force,
        MD_core.cpp:654-676), the force path, then the iteration
        (Main_Calculate_Force + Main_Iteration); a pressure step ends with the
        stress assembly of Get_pressure (main.cpp:713-717), left on the device.
        fused_iterate takes the forces straight from the fixed-point
        accumulators into one remember + Langevin + SETTLE pass
        (snb_md_iterate_fused; bit-identical, self.frc is not written)."""
        if need_virial:
            self.atom_virial.zero_()
            self.virial.zero_()
            self.stress.zero_()
            self.virial_tensor.zero_()
        if fused_iterate:
            self.step(step=step, need_energy=need_energy, need_virial=need_virial,
                      forced_nl=forced_nl, fused=True, flush=False)
            check(self.L.snb_md_iterate_fused(self.ctx, _ptr(self.mass), _ptr(self.crd),
                                              self.cell, self.rcell, _ptr(self.vel),
                                              1 if need_virial else 0, _ptr(self.stress)),
                  "snb_md_iterate_fused")
        else:
            self.frc.zero_()
            self.step(step=step, need_energy=need_energy, need_virial=need_virial,
                      forced_nl=forced_nl, fused=True)
            self.md_iterate(need_pressure=need_virial)
        if need_virial:
            self.get_pressure(copy=False)

    def update_box(self, box):
        """A barostat's small box change (Main_Box_Change, main.cpp:860-897):
        new cell / rcell / volume for every later call and the PME influence
        function for the new box (Particle_Mesh::Update_Box).  Coordinates are
        the caller's to scale."""
        from . import native_io
        cell, rcell = native_io.box_to_cell(box, self.angle)
        self.cell, self.rcell = LTMatrix3.of(cell), LTMatrix3.of(rcell)
        self.box = np.asarray(box, np.float32)
        self.volume = float(np.float32(cell[0]) * np.float32(cell[2]) * np.float32(cell[5]))
        check(self.L.snb_pme_update_box(self.ctx, self.cell, self.rcell, LTMatrix3(), 0.0),
              "snb_pme_update_box")

    def nve_step(self, step=0, need_energy=False, forced_nl=False):
        """One NVE step (mode NVE of Main_Iteration, main.cpp:634-712): the
        force path, SETTLE's remember, the leap-frog update
        (NVE_iteration::Leap_Frog, MD_core/nve.hpp:3-48) with the time step of
        md_initial, then SETTLE with the constraint factors of a plain
        leap-frog (v_factor = x_factor = 1, constrain/constrain.h:25-28)."""
        self.frc.zero_()
        self.step(step=step, need_energy=need_energy, forced_nl=forced_nl, fused=True)
        check(self.L.snb_settle_remember(self.ctx, _ptr(self.crd), self.cell, self.rcell),
              "snb_settle_remember")
        check(self.L.snb_leap_frog(self.ctx, _ptr(self.inv_mass), _ptr(self.frc), _ptr(self.vel),
                                   _ptr(self.crd), self.dt_internal, 0.0), "snb_leap_frog")
        check(self.L.snb_settle_do(self.ctx, _ptr(self.mass), _ptr(self.crd), self.cell,
                                   self.rcell, _ptr(self.vel), 0, _ptr(self.stress),
                                   self.dt_internal, 1.0, 1.0), "snb_settle_do")

    def potential_energy(self):
        """LJ (short + long) + PM + 1-4 (LJ + Coulomb) + bond energies of the
        last energy step (kcal/mol; synchronises)."""
        e = self.energies()
        u = e["LJ"] + e["PM"]
        if self.nb14 is not None:
            u += sum(self.nb14_energy())
        if self.bonds is not None:
            u += self.bond_energy()
        return u

    def kinetic_energy(self):
        """0.5 sum m v^2 of the current velocities, in FP64 on the host."""
        v = self.vel.cpu().numpy().astype(np.float64)
        return 0.5 * float((self.mass.cpu().numpy().astype(np.float64)[:, None] * v * v).sum())

    def get_pressure(self, copy=True):
        """MD_INFORMATION::Get_pressure (MD_core.cpp:948-975) on the current
        velocities: the reciprocal and long-range virial (this wrapper keeps it
        in self.virial) goes into atom 0's entry, where the reference's
        d_virial[0] holds it, then stress += (sum of atom virials + sum m v v) / V.
        With copy, returns (stress[6], pressure) in internal units."""
        self.atom_virial[0].add_(self.virial)
        self.virial.zero_()
        hs, hp = LTMatrix3(), _F()
        check(self.L.snb_get_pressure(self.ctx, self.n, _ptr(self.vel), _ptr(self.mass),
                                      _ptr(self.atom_virial), self.cell,
                                      _ptr(self.virial_tensor), _ptr(self.stress),
                                      ctypes.byref(hs) if copy else None,
                                      ctypes.byref(hp) if copy else None), "snb_get_pressure")
        if copy:
            return np.array([hs.a11, hs.a21, hs.a22, hs.a31, hs.a32, hs.a33], np.float32), hp.value
        return None

    # -- read-back helpers ----------------------------------------------------
    def sync(self):
        check(self.L.snb_synchronize(self.ctx), "snb_synchronize")

    def status(self, clear=False):
        f = _I()
        check(self.L.snb_read_status(self.ctx, 1 if clear else 0, ctypes.byref(f)),
              "snb_read_status")
        return f.value

    def fixed_forces(self):
        t = self.torch.empty(3 * self.n, dtype=self.torch.int64, device=self.dev)
        check(self.L.snb_copy_fixed_forces(self.ctx, _ptr(t)), "snb_copy_fixed_forces")
        self.sync()
        return t.view(3, self.n).cpu().numpy()

    def clear_fixed_forces(self):
        check(self.L.snb_clear_fixed_forces(self.ctx), "snb_clear_fixed_forces")

    def fixed_atom_terms(self, which):
        """which: 0 LJ, 1 direct, 2 exclusion correction, 4 1-4 energy, 5 bond
        energy (n each); 3 virial (6 x n)."""
        cnt = 6 * self.n if which == 3 else self.n
        t = self.torch.empty(cnt, dtype=self.torch.int64, device=self.dev)
        check(self.L.snb_copy_fixed_atom_terms(self.ctx, which, _ptr(t)),
              "snb_copy_fixed_atom_terms")
        self.sync()
        a = t.cpu().numpy()
        return a.reshape(6, self.n) if which == 3 else a

    def pme_grid(self, which):
        t = self.torch.empty(self.grid, dtype=self.torch.float32, device=self.dev)
        check(self.L.snb_copy_pme_grid(self.ctx, which, _ptr(t)), "snb_copy_pme_grid")
        self.sync()
        return t.cpu().numpy()

    def nl_rows(self):
        """(counts int32[n], rows int32[n, max_nb]) of the public half list."""
        torch = self.torch
        counts = torch.empty(self.n, dtype=torch.int32, device=self.dev)
        rows = torch.empty((self.n, self.max_nb), dtype=torch.int32, device=self.dev)
        check(self.L.snb_nl_copy_rows(self.ctx, _ptr(counts), _ptr(rows)), "snb_nl_copy_rows")
        self.sync()
        return counts.cpu().numpy(), rows.cpu().numpy()

    def internal_rows(self):
        """(counts, rows) of the list the pair kernel reads, by row-owner atom id."""
        torch = self.torch
        counts = torch.empty(self.n, dtype=torch.int32, device=self.dev)
        rows = torch.empty((self.n, self.max_nb), dtype=torch.int32, device=self.dev)
        check(self.L.snb_nl_copy_internal(self.ctx, _ptr(counts), _ptr(rows)),
              "snb_nl_copy_internal")
        self.sync()
        return counts.cpu().numpy(), rows.cpu().numpy()

    def internal_counts(self):
        """Row lengths of the internal list, by row-owner atom id."""
        torch = self.torch
        counts = torch.empty(self.n, dtype=torch.int32, device=self.dev)
        rows = torch.empty((self.n, self.max_nb), dtype=torch.int32, device=self.dev)
        check(self.L.snb_nl_copy_internal(self.ctx, _ptr(counts), _ptr(rows)),
              "snb_nl_copy_internal")
        self.sync()
        return counts.cpu().numpy()

    def set_export(self, enable):
        """Whether rebuilds also write the reference-layout public list: 0 / False
        no, 1 / True yes (entries of a row in append order), 2 yes with every
        row sorted ascending."""
        mode = int(enable) if not isinstance(enable, bool) else (1 if enable else 0)
        check(self.L.snb_nl_set_export(self.ctx, mode), "snb_nl_set_export")

    def set_owner_attribution(self, enable):
        """Per-atom pair energy/virial to the internal row owner (True) or to
        the lower atom id as the reference does (False, default)."""
        check(self.L.snb_lj_set_owner_attribution(self.ctx, 1 if enable else 0),
              "snb_lj_set_owner_attribution")

    def set_window_capacity(self, slots):
        check(self.L.snb_nl_set_window_capacity(self.ctx, int(slots)),
              "snb_nl_set_window_capacity")

    def set_window_shape(self, a, b, s):
        """Block shape of the windowed pair path, (0, 0, 0) = plain path;
        returns False when the grid or the LDS cannot take it."""
        rc = self.L.snb_nl_set_window_shape(self.ctx, int(a), int(b), int(s))
        if rc == 1:      # SNB_ERR_ARGUMENT: shape not possible here
            return False
        check(rc, "snb_nl_set_window_shape")
        return True

    def set_direct_j(self, mode):
        """Window forces of force-only steps: 0 through the stage and its
        gather, 1 by atomic adds per block, -1 by system size (default)."""
        check(self.L.snb_lj_set_direct_j(self.ctx, int(mode)), "snb_lj_set_direct_j")

    def set_row_stream(self, mode):
        """Force-only pair loop: 0 pipelined within each row, 1 across row
        ends, -1 by system size (default)."""
        check(self.L.snb_lj_set_row_stream(self.ctx, int(mode)), "snb_lj_set_row_stream")

    def set_row_lds(self, mode):
        """Row kernel of small systems on force-only steps: 1 LDS accumulators,
        0 global atomics, -1 LDS only inside recorded step batches (snb_lj_set_row_lds)."""
        check(self.L.snb_lj_set_row_lds(self.ctx, int(mode)), "snb_lj_set_row_lds")

    def window_info(self):
        v = [_I() for _ in range(6)]
        check(self.L.snb_nl_window_info(self.ctx, *[ctypes.byref(x) for x in v]),
              "snb_nl_window_info")
        keys = ("windowed", "a", "b", "s", "wcap", "nblock")
        return dict(zip(keys, [x.value for x in v]))

    def profile(self, enable=True):
        check(self.L.snb_profile_enable(self.ctx, 1 if enable else 0), "snb_profile_enable")

    def profile_read(self, group, reset=True):
        ms = ctypes.c_double()
        cnt = _I()
        check(self.L.snb_profile_read(self.ctx, group, 1 if reset else 0, ctypes.byref(ms),
                                      ctypes.byref(cnt)), "snb_profile_read")
        return ms.value, cnt.value

    # -- FP64 reference evaluation ---------------------------------------------
    F64_TERMS = ("LJ_short", "LJ_long", "PM_direct", "PM_reciprocal", "PM_self",
                 "PM_correction", "nb14_LJ", "nb14_EE", "bond")

    def f64_force(self, crd64, frc64, step=0, forced_nl=False, need_energy=False,
                  need_virial=False):
        """The whole path in double on float64 device tensors crd64 / frc64
        [n, 3] (frc64 accumulates).  Returns (terms dict or None, virial[6] or
        None); both synchronise."""
        terms = (ctypes.c_double * 9)() if need_energy else None
        vir = (ctypes.c_double * 6)() if need_virial else None
        check(self.L.snb_f64_force(self.ctx, _ptr(crd64), _ptr(self.charge), self.cell,
                                   self.rcell, step, 1 if forced_nl else 0, _ptr(frc64),
                                   1 if need_energy else 0, 1 if need_virial else 0,
                                   ctypes.cast(terms, ctypes.c_void_p) if terms else None,
                                   ctypes.cast(vir, ctypes.c_void_p) if vir else None),
              "snb_f64_force")
        out = None
        if need_energy:
            out = dict(zip(self.F64_TERMS, terms[:]))
            out["LJ"] = out["LJ_short"] + out["LJ_long"]
            out["PM"] = (out["PM_direct"] + out["PM_reciprocal"] + out["PM_self"]
                         + out["PM_correction"])
        return out, (np.array(vir[:]) if need_virial else None)

    def f64_nve_iterate(self, mass64, inv_mass64, frc64, vel64, crd64, dt_internal):
        """Remember, leap-frog and SETTLE in double (after f64_force)."""
        check(self.L.snb_f64_nve_iterate(self.ctx, _ptr(mass64), _ptr(inv_mass64), _ptr(frc64),
                                         _ptr(vel64), _ptr(crd64), self.cell, self.rcell,
                                         float(dt_internal)), "snb_f64_nve_iterate")

    def energies(self):
        d = [ctypes.c_double() for _ in range(4)]
        check(self.L.snb_pme_energy(self.ctx, *[ctypes.byref(x) for x in d]), "snb_pme_energy")
        ls, ll = ctypes.c_double(), ctypes.c_double()
        check(self.L.snb_lj_energy(self.ctx, ctypes.byref(ls), ctypes.byref(ll)), "snb_lj_energy")
        out = dict(LJ_short=ls.value, LJ_long=ll.value, PM_direct=d[0].value,
                   PM_reciprocal=d[1].value, PM_self=d[2].value, PM_correction=d[3].value)
        out["LJ"] = out["LJ_short"] + out["LJ_long"]
        out["PM"] = out["PM_direct"] + out["PM_reciprocal"] + out["PM_self"] + out["PM_correction"]
        return out

    def reciprocal_virial(self):
        v = (ctypes.c_double * 6)()
        check(self.L.snb_pme_reciprocal_virial(self.ctx, ctypes.cast(v, ctypes.c_void_p)),
              "snb_pme_reciprocal_virial")
        return np.array(v[:])


class StepBatch:
    """Several systems' MD steps in lockstep, one dispatch per kernel
    (snb_batch_*).  record(nb, first_step, steps) records nb.md_step for
    steps first_step .. first_step + steps - 1 (force-only, fused iteration)
    without launching; plan() merges the records of all members; launch(stream)
    issues them, as often as wanted (it can be captured in a HIP graph).
    Every member's results are bit-identical to its own md_step calls.  The
    members must stay alive while a launch of this batch may run."""

    def __init__(self, device=0):
        self.L = load()
        b = ctypes.c_void_p()
        check(self.L.snb_batch_create(ctypes.byref(b), device), "snb_batch_create")
        self.b = b
        self.members = []
        self.launches = self.batched = 0

    def record(self, nb, first_step, steps):
        check(self.L.snb_batch_record_begin(self.b, nb.ctx), "snb_batch_record_begin")
        try:
            for k in range(steps):
                nb.md_step(step=first_step + k)
        except BaseException:
            # the member's record is incomplete: the batch cannot be planned
            self.L.snb_batch_record_end(self.b)
            self.close()
            raise
        check(self.L.snb_batch_record_end(self.b), "snb_batch_record_end")
        self.members.append(nb)

    def plan(self):
        if self.b is None:
            raise RuntimeError("the step batch was closed")
        a, b = _I(), _I()
        check(self.L.snb_batch_plan(self.b, ctypes.byref(a), ctypes.byref(b)), "snb_batch_plan")
        self.launches, self.batched = a.value, b.value
        return self.launches, self.batched

    def info(self):
        v = [_I() for _ in range(4)]
        check(self.L.snb_batch_info(self.b, *[ctypes.byref(x) for x in v]), "snb_batch_info")
        return dict(zip(("members", "launches", "batched", "merged"), (x.value for x in v)))

    def launch(self, stream):
        check(self.L.snb_batch_launch(self.b, ctypes.c_void_p(stream.cuda_stream)),
              "snb_batch_launch")

    def close(self):
        if self.b is not None:
            self.L.snb_batch_destroy(self.b)
            self.b = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
