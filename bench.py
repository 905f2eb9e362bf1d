#!/usr/bin/env python3
"""bench.py -- ns/day of SPONGE's nonbonded force path on one MI355X.

Workload: the 93,000-atom mixed system (30,000 TIP3P waters + 3 bead chains
of 1,000 beads, 96.5 A cubic box, cutoff 10 A, skin 2 A, PME grid 96^3,
order 4, Ewald tolerance 1e-5), synthetic coordinates from the seeded
generator.  One step = one evaluation of the path in the reference's order:
neighbor-list update (rebuilt every --refresh steps, the reference's
interval policy), PME exclusion correction, LJ + direct Coulomb, PME
reciprocal, one fixed-point -> float force conversion; energies and virials
every --energy-interval steps.  ns/day = steps/s * dt(2 fs) * 86400 / 1000.

Prints one JSON line (rank 0).  The roofline object is for the pair kernel
(algorithmic bytes 4 * E_list + 44 * N per launch, timed with HIP events on
its stream); cpu_baseline times the CPU port (oracle/) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "sponge2.0_amd", "python"))

HBM_PEAK_GBS = 8000.0
DT_PS = 0.002
NS_PER_DAY_PER_STEP_S = DT_PS * 86400.0 / 1000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", choices=("mixed93k", "water1m"), default="mixed93k",
                    help="mixed93k: the 93,000-atom mixed system (the metric's workload); "
                         "water1m: 333,334 TIP3P waters, 1,000,002 atoms, PME 216^3, list "
                         "rebuilt every 20 steps (timed force path only unless other phases "
                         "are asked for)")
    ap.add_argument("--refresh", type=int, default=None,
                    help="neighbor_list_refresh_interval (rebuild every N steps)")
    ap.add_argument("--energy-interval", type=int, default=100)
    ap.add_argument("--cpu-seconds", type=float, default=None,
                    help="CPU-baseline sample budget (0 disables; default 15 s, 0 for water1m)")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-overlap", action="store_true",
                    help="run the reciprocal pipeline on the context stream (diagnostics)")
    ap.add_argument("--no-graph", action="store_true",
                    help="time eager launches instead of HIP-graph replays of each kind of step")
    ap.add_argument("--profile-steps", type=int, default=200,
                    help="eager steps, before the timed region, whose kernel-group times "
                         "give kernel_avg_us and the roofline")
    ap.add_argument("--md-steps", type=int, default=None,
                    help="also time this many whole MD steps (forces + middle Langevin + "
                         "SETTLE) after the force-path measurement; 0 skips (default 500, 0 "
                         "for water1m)")
    ap.add_argument("--nve-steps", type=int, default=None,
                    help="precision study: NVE steps at 1 fs (SETTLE-rigid waters) after a "
                         "Langevin equilibration, total-energy drift per ns per degree of "
                         "freedom; 0 skips (default 20,000, 0 for water1m; the full study is "
                         "100,000)")
    ap.add_argument("--nve-equil", type=int, default=2000,
                    help="Langevin steps (300 K, tau 0.1 ps, 1 fs) before the NVE run")
    ap.add_argument("--nve-f64-steps", type=int, default=None,
                    help="NVE steps of the all-FP64 trajectory from the same state (default "
                         "5,000, 0 for water1m; the full study is 100,000)")
    ap.add_argument("--nve-compare-every", type=int, default=1000,
                    help="FP64-trajectory steps between frames on which the FP32 path and the "
                         "FP64 evaluation are compared")
    ap.add_argument("--serve-requests", type=int, default=None,
                    help="serving workload: this many water-box segment requests (N log-uniform "
                         "1,500-50,000 atoms, Poisson arrivals at 40/s), served one at a time and "
                         "with continuous batching; 0 skips (default 256, the full workload, 0 "
                         "for water1m)")
    ap.add_argument("--serve-steps", type=int, default=1000,
                    help="MD steps per serving request (2 fs)")
    ap.add_argument("--serve-slots", type=int, default=8,
                    help="segments in flight in the per-segment streams run")
    ap.add_argument("--serve-batch-slots", type=int, default=12,
                    help="segments in flight in the step-batch run")
    ap.add_argument("--export-steps", type=int, default=None,
                    help="also time this many steps with the reference-layout list written at "
                         "every rebuild (the drop-in default); 0 skips (default 200, 0 for "
                         "water1m)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run one 256 MiB device copy (scales the HBM byte counters "
                         "of a profiler run)")
    a = ap.parse_args()
    big = a.workload == "water1m"
    defaults = {"refresh": 20 if big else 10, "cpu_seconds": 0.0 if big else 15.0,
                "md_steps": 0 if big else 500, "nve_steps": 0 if big else 20000,
                "nve_f64_steps": 0 if big else 5000,
                "serve_requests": 0 if big else 256, "export_steps": 0 if big else 200}
    for k, v in defaults.items():
        if getattr(a, k) is None:
            setattr(a, k, v)
    if big:
        a.no_accuracy = True   # the FP64 oracle over 210M pairs is not a bench leg
    return a


def cpu_baseline(sysd, nb, counts, rows, budget_s):
    """The CPU port (oracle/, OpenMP over the host's threads) on the same 93k
    workload: FP32 twin pair kernel over the GPU-built list + FP32 twin
    exclusion correction + FP32 twin 1-4 pairs and bonds + FP64 PME reciprocal (spread,
    mixed-radix FFT, gather).  The list rebuild is left out (it favours the CPU).  Runs whole
    steps until budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    s = sysd
    excl = (s["excl_start"], s["excl_numbers"], s["excl_list"])
    steps = 0
    t0 = time.perf_counter()
    pme = None
    while True:
        oracle.pair_twin(s["crd"], s["charge"], s["lj_type"], s["lj_a"], s["lj_b"], s["cell"],
                         s["rcell"], nb.cutoff, nb.beta, counts, rows, virial=False)
        oracle.excl_twin(s["crd"], s["charge"], s["cell"], s["rcell"], nb.beta, excl,
                         virial=False)
        if "nb14" in s:
            oracle.nb14_twin(s["crd"], s["charge"], s["cell"], s["rcell"], s["nb14"],
                             virial=False)
        if "bonds" in s:
            oracle.bond_twin(s["crd"], s["cell"], s["rcell"], s["bonds"], virial=False)
        pme = oracle.pme_f64(s["crd"], s["charge"], s["rcell"], nb.grid, nb.beta, nb.volume)
        steps += 1
        el = time.perf_counter() - t0
        if el >= budget_s or steps >= 500:
            break
    sps = steps / el
    nt = oracle.num_threads()
    return {"value": sps * NS_PER_DAY_PER_STEP_S, "unit": "ns/day", "cores": nt, "kind": "port",
            "sample": "%d whole steps of the 93k workload (pair+exclusion+1-4+bond FP32 "
                      "fixed-point twin, FP64 PME), NL rebuild excluded, %.1f s, "
                      "%d OpenMP threads"
                      % (steps, el, nt)}, pme


def frame_accuracy(nb, sysd, frame):
    """One energy/virial step of the library on its current coordinates
    against the FP64 oracle on the same coordinates and list."""
    import numpy as np
    s = dict(sysd)
    s["crd"] = nb.crd.cpu().numpy().copy()
    nb.zero_outputs()
    nb.step(step=0, need_energy=True, need_virial=True, forced_nl=True, fused=True)
    nb.sync()
    gpu = {"frc": nb.frc.cpu().numpy(), "energy": nb.energies()}
    if nb.nb14 is not None:
        gpu["energy"]["nb14_LJ"], gpu["energy"]["nb14_EE"] = nb.nb14_energy()
    if nb.bonds is not None:
        gpu["energy"]["bond"] = nb.bond_energy()
    # total virial: per-atom virials (pair, exclusion, 1-4, bond; float, as
    # the reference's atom_virial) summed in FP64, plus the reciprocal tensor
    gpu["virial"] = (nb.atom_virial.cpu().numpy().astype(np.float64).sum(axis=0)
                     + nb.reciprocal_virial())
    counts, rows = nb.internal_rows()
    acc = accuracy(s, nb, counts, rows, None, gpu)
    return {"force_rel_rms_vs_fp64": acc["force_rel_rms"],
            "energy_rel_err_vs_fp64": acc["energy_rel_err"],
            "virial_rel_err_vs_fp64": acc["virial_rel_err"],
            "accuracy_frame": frame}


def accuracy(sysd, nb, counts, rows, pme, gpu):
    """Errors of the GPU energy/virial step against the FP64 oracle on the
    same list: per-atom force relative RMS, each energy term's relative error
    and the relative error (max norm) of the total virial tensor."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    s = sysd
    excl = (s["excl_start"], s["excl_numbers"], s["excl_list"])
    f1, e_lj, e_c, v1 = oracle.pair_f64(s["crd"], s["charge"], s["lj_type"], s["lj_a"],
                                        s["lj_b"], s["cell"], s["rcell"], nb.cutoff, nb.beta,
                                        counts, rows, virial=True)
    f2, e_ex, v2 = oracle.excl_f64(s["crd"], s["charge"], s["cell"], s["rcell"], nb.beta, excl,
                                   virial=True)
    if pme is None:
        pme = oracle.pme_f64(s["crd"], s["charge"], s["rcell"], nb.grid, nb.beta, nb.volume)
    ref = f1 + f2 + pme[0]
    vir_atoms = v1 + v2
    terms = {"LJ_short": float(np.sum(e_lj)), "PM_direct": float(np.sum(e_c)),
             "PM_correction": float(np.sum(e_ex)), "PM_reciprocal": pme[1],
             "PM_self": oracle.self_energy(s["charge"], nb.beta, nb.volume)}
    terms["PM"] = (terms["PM_direct"] + terms["PM_correction"] + terms["PM_reciprocal"] +
                   terms["PM_self"])
    if "nb14" in s:
        f14, ep14, _, v14 = oracle.nb14_f64(s["crd"], s["charge"], s["cell"], s["rcell"],
                                            s["nb14"], virial=True)
        ref, vir_atoms = ref + f14, vir_atoms + v14
        terms["nb14_LJ"], terms["nb14_EE"] = float(ep14[0].sum()), float(ep14[1].sum())
    if "bonds" in s:
        fb, epb, _, vb = oracle.bond_f64(s["crd"], s["cell"], s["rcell"], s["bonds"],
                                         virial=True)
        ref, vir_atoms = ref + fb, vir_atoms + vb
        terms["bond"] = float(epb.sum())
    g = gpu["frc"].astype(np.float64)
    out = {"force_rel_rms": float(np.sqrt(((g - ref) ** 2).sum() / (ref ** 2).sum()))}
    # the four PM terms cancel heavily (self ~ -correction): each is taken
    # relative to the PM total, the other terms relative to themselves
    out["energy_rel_err"] = {k: abs(gpu["energy"][k] - v) / abs(terms["PM"] if k.startswith("PM_")
                                                                else v)
                             for k, v in terms.items() if k in gpu["energy"]}
    vref = np.asarray(vir_atoms, np.float64).sum(axis=0) + np.asarray(pme[2])
    vgot = gpu["virial"]
    out["virial_rel_err"] = float(np.max(np.abs(vgot - vref)) / np.max(np.abs(vref)))
    return out


def serving_phase(args):
    """The serving workload: the same seeded request stream served twice, one
    segment at a time (sequential plain runs) and with continuous batching;
    window shapes tuned once per grid class beforehand (untimed, both runs
    use them).  Every request's final coordinates and velocities are compared
    between the two runs (alone vs in the mix)."""
    import numpy as np
    from sponge_amd import serving
    t = time.perf_counter()
    tuned = serving.pretune(serving.make_requests(count=args.serve_requests,
                                                  steps=args.serve_steps))
    tune_s = time.perf_counter() - t
    runs = {}
    reqs = {}
    # The sequential runs use the plain path's default (reciprocal pipeline
    # on its own stream beside the direct-space work).  The batched run
    # advances the admitted segments as one step batch (one dispatch per
    # kernel over all segments, one captured graph per period); the streams
    # run gives each segment its own stream and period graph instead.
    # "plain": the sequential runs with the library's defaults, as a drop-in
    # user gets them (default window shape, no tuner, reference-layout list
    # written at every rebuild); "sequential": the same with the tuned shapes
    # and the list export off.
    for name, slots, overlap, launch, plain in (
            ("plain", 1, True, "streams", True),
            ("sequential", 1, True, "streams", False),
            ("batched", args.serve_batch_slots, False, "batched", False),
            ("streams", args.serve_slots, False, "streams", False)):
        rq = serving.make_requests(count=args.serve_requests, steps=args.serve_steps)
        runs[name] = serving.Server(slots=slots, overlap=overlap, launch=launch,
                                    plain=plain).run(rq, realtime=True)
        reqs[name] = rq
        print("serving %s: %.1f ns/day, makespan %.2f s" % (
            name, runs[name]["aggregate_ns_per_day"], runs[name]["makespan_s"]),
            file=sys.stderr, flush=True)
    same = sum(int(all(np.array_equal(a.result["crd"], reqs[o][k].result["crd"]) and
                       np.array_equal(a.result["vel"], reqs[o][k].result["vel"])
                       for o in ("sequential", "batched", "streams")))
               for k, a in enumerate(reqs["plain"]))
    status = sum(r.result["status"] for v in reqs.values() for r in v)
    pl, seq, bat, strm = runs["plain"], runs["sequential"], runs["batched"], runs["streams"]
    return {"requests": args.serve_requests, "steps_per_request": args.serve_steps,
            "atoms_total": bat["atoms_total"], "arrival": "Poisson, 40/s",
            "batched_aggregate_ns_per_day": bat["aggregate_ns_per_day"],
            "batched_p50_latency_s": bat["p50_latency_s"],
            "batched_p99_latency_s": bat["p99_latency_s"],
            "sequential_aggregate_ns_per_day": seq["aggregate_ns_per_day"],
            "sequential_p50_latency_s": seq["p50_latency_s"],
            "sequential_p99_latency_s": seq["p99_latency_s"],
            "plain_aggregate_ns_per_day": pl["aggregate_ns_per_day"],
            "batched_over_sequential": bat["aggregate_ns_per_day"] / pl["aggregate_ns_per_day"],
            "batched_over_sequential_tuned": (bat["aggregate_ns_per_day"] /
                                              seq["aggregate_ns_per_day"]),
            "sequential_baselines": {"plain": "library defaults (default window shape, no "
                                              "tuner, reference-layout list written at every "
                                              "rebuild), reciprocal pipeline on its own stream",
                                     "sequential": "tuned window shapes, list export off"},
            "streams_aggregate_ns_per_day": strm["aggregate_ns_per_day"],
            "streams_over_sequential_tuned": (strm["aggregate_ns_per_day"] /
                                              seq["aggregate_ns_per_day"]),
            "slots": {"batched": args.serve_batch_slots, "streams": args.serve_slots},
            "makespan_s": {"plain": pl["makespan_s"],
                                                      "sequential": seq["makespan_s"],
                                                      "batched": bat["makespan_s"],
                                                      "streams": strm["makespan_s"]},
            "batched_plans": bat["plans"], "batched_last_plan": bat["last_plan"],
            "same_alone_and_in_mix": "%d of %d requests bit-identical in all four runs"
                                     % (same, len(reqs["batched"])),
            "status_flags": status, "tune_s": tune_s,
            "shapes": {str(k): list(v[0]) + [v[2]] for k, v in tuned.items()},
            "launch": {"batched": "one step batch over the admitted segments: one dispatch per "
                                  "kernel of a step, reciprocal pipeline on a second stream, one "
                                  "captured graph per 10-step period, re-planned when a segment "
                                  "joins or leaves, periods queued 2 ahead",
                       "streams": "per-segment HIP stream, one captured graph per 10-step "
                                  "period, periods queued 2 ahead"}}


def _drift(t_ps, etot, dof):
    """Linear fit of the total energy over the whole run and over its part
    after the first 10% (the switch from the thermostat leaves a transient),
    plus the total energy at a few times."""
    import numpy as np
    t, E = np.asarray(t_ps), np.asarray(etot)
    kt = 0.00198716 * 300.0

    def slope(a, b):
        return float(np.polyfit(a, b, 1)[0]) if len(a) > 1 else 0.0   # kcal/mol per ps
    late = t >= 0.1 * t[-1]
    s_all, s_late = slope(t, E), slope(t[late], E[late])
    marks = sorted(set(int(i) for i in np.linspace(0, len(E) - 1, 6)) | {1, 2, 10})
    return {"samples": len(E), "E_total_first": float(E[0]), "E_total_last": float(E[-1]),
            "E_total_std": float(E.std()), "E_total_std_after_10pct": float(E[late].std()),
            "E_total_at_ps": {"%g" % t[i]: float(E[i]) for i in marks if i < len(E)},
            "drift_kcal_per_mol_per_ns_per_dof": s_all * 1000.0 / dof,
            "drift_kT_per_ns_per_dof": s_all * 1000.0 / dof / kt,
            "drift_after_10pct_kcal_per_mol_per_ns_per_dof": s_late * 1000.0 / dof,
            "drift_after_10pct_kT_per_ns_per_dof": s_late * 1000.0 / dof / kt}


def nve_phase(nb, sysd, args):
    """Precision study: Maxwell velocities at 300 K and a Langevin
    equilibration of the FP32 path, then from that one state two NVE
    leap-frog trajectories at 1 fs with SETTLE: the FP32 / int64 path and the
    all-FP64 evaluation of the same path (snb_f64_force / snb_f64_nve_iterate).
    Total energy (path potential + kinetic as the mean of the two half-step
    values) every 100 steps, linear drift per ns per degree of freedom for
    both; on frames of the FP64 trajectory (every --nve-compare-every steps)
    the FP32 path and the FP64 evaluation on the same float coordinates:
    per-atom force relative RMS, each energy term's relative error (PM terms
    relative to the PM total) and the total virial's."""
    import numpy as np
    import torch
    from sponge_amd import lib
    nb.md_initial(sysd, temperature=300.0, tau=0.1, seed=20260218, dt=0.001)
    for k in range(args.nve_equil):
        nb.md_step(step=k, forced_nl=(k == 0))
    nb.sync()
    crd0, vel0 = nb.crd.clone(), nb.vel.clone()
    nw = int(sysd["n_waters"])
    dof = 6 * nw + 3 * (sysd["n"] - 3 * nw) - 3
    sample = 100
    t_ps, etot = [], []
    t0 = time.perf_counter()
    for k in range(args.nve_steps):
        e = (k % sample) == 0
        if e:
            ke0 = nb.kinetic_energy()
        nb.nve_step(step=k, need_energy=e, forced_nl=(k == 0))
        if e:
            u = nb.potential_energy()
            ke1 = nb.kinetic_energy()
            t_ps.append(k * 0.001)
            etot.append(u + 0.5 * (ke0 + ke1))
        if (k + 1) % 10000 == 0:
            print("nve fp32: %d steps, E_total %.3f" % (k + 1, etot[-1]), file=sys.stderr,
                  flush=True)
    nb.sync()
    wall = time.perf_counter() - t0
    out = {"steps": args.nve_steps, "dt_fs": 1.0, "equilibration_steps": args.nve_equil,
           "dof": dof}
    out.update(_drift(t_ps, etot, dof))
    out.update({"wall_s": wall, "status_flags": nb.status(),
                "arithmetic": "FP32 pair/spline/integrator math, int64 fixed-point forces, "
                              "FP64 energy reductions"})
    if args.nve_f64_steps <= 0:
        return out
    # the FP64 trajectory from the same state
    nb64 = lib.Nonbonded(sysd, grid=nb.grid, refresh_interval=10)
    nb64.md_initial(sysd, temperature=300.0, tau=0.1, seed=20260218, dt=0.001)  # SETTLE
    crd, vel = crd0.double().contiguous(), vel0.double().contiguous()
    mass = torch.as_tensor(np.asarray(sysd["mass"], np.float64)).to(nb.dev)
    inv = torch.where(mass > 0, 1.0 / torch.where(mass > 0, mass, torch.ones_like(mass)),
                      torch.zeros_like(mass))
    frc = torch.zeros_like(crd)
    probe64 = torch.zeros_like(crd)

    def ke64():
        return 0.5 * float((mass[:, None] * vel * vel).sum())

    def potential(t):
        return t["LJ"] + t["PM"] + t["nb14_LJ"] + t["nb14_EE"] + t["bond"]

    t64, e64, frames = [], [], []
    dt = nb.dt_internal
    t0 = time.perf_counter()
    for k in range(args.nve_f64_steps):
        e = (k % sample) == 0
        if e:
            ke0 = ke64()
        frc.zero_()
        terms, _ = nb64.f64_force(crd, frc, step=k, forced_nl=(k == 0), need_energy=e)
        if args.nve_compare_every > 0 and k % args.nve_compare_every == 0:
            frames.append(_compare_frame(nb, nb64, crd, probe64))
        nb64.f64_nve_iterate(mass, inv, frc, vel, crd, dt)
        if e:
            t64.append(k * 0.001)
            e64.append(potential(terms) + 0.5 * (ke0 + ke64()))
        if (k + 1) % 5000 == 0:
            print("nve fp64: %d steps, E_total %.3f" % (k + 1, e64[-1]), file=sys.stderr,
                  flush=True)
    torch.cuda.synchronize()
    f64 = {"steps": args.nve_f64_steps, "wall_s": time.perf_counter() - t0,
           "status_flags": nb64.status(),
           "arithmetic": "FP64 pair/exclusion/1-4/bond/PME (direct FP64 DFTs) and leap-frog + "
                         "SETTLE; forces summed in int64 at 2^40"}
    f64.update(_drift(t64, e64, dof))
    nb64.close()
    out["fp64"] = f64
    if frames:
        keys = sorted(frames[0]["energy_rel_err"])
        out["frames_compared"] = len(frames)
        out["force_rel_rms_vs_fp64"] = {"max": max(f["force"] for f in frames),
                                        "mean": float(np.mean([f["force"] for f in frames])),
                                        "last": frames[-1]["force"]}
        out["energy_rel_err_vs_fp64_max"] = {k: max(f["energy_rel_err"][k] for f in frames)
                                             for k in keys}
        out["potential_rel_err_vs_fp64_max"] = max(f["potential"] for f in frames)
        out["virial_rel_err_vs_fp64_max"] = max(f["virial"] for f in frames)
    return out


def _compare_frame(nb, nb64, crd64, probe64):
    """The FP32 path (context nb) and the FP64 evaluation (context nb64) on
    the float rounding of the FP64 trajectory's current frame."""
    import numpy as np
    x32 = crd64.float()
    nb.crd.copy_(x32)
    nb.zero_outputs()
    nb.step(step=0, need_energy=True, need_virial=True, forced_nl=True, fused=True)
    nb.sync()
    e32 = nb.energies()
    e32["nb14_LJ"], e32["nb14_EE"] = nb.nb14_energy() if nb.nb14 is not None else (0.0, 0.0)
    e32["bond"] = nb.bond_energy() if nb.bonds is not None else 0.0
    f32 = nb.frc.cpu().numpy().astype(np.float64)
    v32 = (nb.atom_virial.cpu().numpy().astype(np.float64).sum(axis=0) + nb.reciprocal_virial())
    probe64.zero_()
    t64, v64 = nb64.f64_force(x32.double().contiguous(), probe64, step=0, forced_nl=True,
                              need_energy=True, need_virial=True)
    f64 = probe64.cpu().numpy()
    err = {}
    for k, v in t64.items():
        if k in ("LJ", "PM") or v == 0.0:
            continue
        scale = abs(t64["PM"]) if k.startswith("PM_") else abs(v)
        err[k] = abs(e32[k] - v) / scale
    u32 = e32["LJ"] + e32["PM"] + e32["nb14_LJ"] + e32["nb14_EE"] + e32["bond"]
    u64 = t64["LJ"] + t64["PM"] + t64["nb14_LJ"] + t64["nb14_EE"] + t64["bond"]
    return {"force": float(np.sqrt(((f32 - f64) ** 2).sum() / (f64 ** 2).sum())),
            "energy_rel_err": err, "potential": abs(u32 - u64) / abs(u64),
            "virial": float(np.max(np.abs(v32 - v64)) / np.max(np.abs(v64)))}


class StepGraphs:
    """Captured HIP graphs of the steps of a run, made before the timed
    region: one graph per neighbour-list period (`period` steps from a
    multiple of it, so one rebuild each) and one per single step for the
    rest, keyed by the kinds of their steps (rebuild or not, energy or not).
    A graph captured with one run of step numbers is replayed for every run
    of the same kinds: the step number only decides those.  Replaying them in
    order issues the same work as step-by-step replays, with one graph launch
    per period instead of one per step."""

    def __init__(self, nb, run_step, kind, period):
        self.nb, self.run_step, self.kind, self.period = nb, run_step, kind, period
        self.graphs = {}

    def _capture(self, ks):
        import torch
        nb = self.nb
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=nb.dev)
        home = nb.stream
        nb.sync()
        torch.cuda.synchronize()
        nb.set_stream(side)
        try:
            with torch.cuda.graph(graph, stream=side):
                for k in ks:
                    self.run_step(k)
        finally:
            nb.set_stream(home)
        torch.cuda.synchronize()
        return graph

    def plan(self, k0, steps):
        """The graph sequence for steps k0 .. k0 + steps - 1 (captured now)."""
        seq, k, P = [], k0, self.period
        while k < k0 + steps:
            ks = list(range(k, k + P)) if (P > 1 and k % P == 0 and k + P <= k0 + steps) else [k]
            key = tuple(self.kind(j) for j in ks)
            if key not in self.graphs:
                self.graphs[key] = self._capture(ks)
            seq.append(self.graphs[key])
            k += len(ks)
        return seq


def md_phase(nb, sysd, args, kind):
    """Whole MD steps of the same workload: the force path, then the middle
    Langevin leap-frog (300 K, tau 0.1 ps) and SETTLE on the 30,000 waters,
    from Maxwell velocities; replayed as one HIP graph per kind of step like
    the force-path phase.  Runs last: it moves the atoms."""
    import numpy as np
    import torch
    # tau 0.1 ps: the generated lattice releases its relaxation heat quickly
    nb.md_initial(sysd, temperature=300.0, tau=0.1, seed=20260217, dt=0.002)
    k0 = 0
    for k in range(k0, k0 + 20):
        nb.md_step(step=k, need_energy=False, need_virial=False, forced_nl=(k == 0))
    k0 += 20
    def run_step(k):
        e = (k % args.energy_interval) == 0
        nb.md_step(step=k, need_energy=e, need_virial=e)
    seq = None
    if not args.no_graph:
        seq = StepGraphs(nb, run_step, kind, args.refresh).plan(k0, args.md_steps)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    ev0.record(nb.stream)
    if seq is not None:
        for g in seq:
            g.replay()
    else:
        for k in range(k0, k0 + args.md_steps):
            run_step(k)
    ev1.record(nb.stream)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / args.md_steps
    v = nb.vel.cpu().numpy().astype(np.float64)
    m = np.asarray(sysd["mass"], np.float64)
    nw = int(sysd["n_waters"])
    dof = 6 * nw + 3 * (sysd["n"] - 3 * nw) - 3
    t_kin = float((m[:, None] * v * v).sum() / (dof * 0.00198716))
    st = nb.stress.cpu().numpy().astype(np.float64)    # the last pressure step's stress
    p_bar = (st[0] + st[2] + st[5]) / 3.0 * 6.946827162543585e4
    return {"ns_per_day": 1000.0 / ms * NS_PER_DAY_PER_STEP_S, "ms_per_step": ms,
            "steps": args.md_steps,
            "launch": "hip-graph replay, one graph per neighbour-list period"
            if seq is not None else "eager", "integrator": "middle Langevin leap-frog, 300 K, tau 0.1 ps, 2 fs",
            "constraints": "SETTLE on %d rigid waters" % nw, "t_kin_end": t_kin,
            "pressure_bar_last": p_bar,
            "pressure_steps": "every %d steps (virials, SETTLE virial, Get_pressure on the "
                              "device)" % args.energy_interval,
            "status_flags": nb.status()}


def pairs_in_cutoff(sysd, counts, rows, cutoff):
    """Number of list pairs closer than the cutoff (numpy over the GPU-built
    list, minimum image in the diagonal box of this workload)."""
    import numpy as np
    crd = sysd["crd"].astype(np.float64)
    L = np.array([sysd["cell"][0], sysd["cell"][2], sysd["cell"][5]], dtype=np.float64)
    total = 0
    for i0 in range(0, len(counts), 8192):
        c = counts[i0:i0 + 8192]
        blk = rows[i0:i0 + 8192]
        mask = np.arange(blk.shape[1])[None, :] < c[:, None]
        ii = np.nonzero(mask)
        j = blk[ii]
        d = crd[j] - crd[i0 + ii[0]]
        d -= np.rint(d / L) * L
        total += int(np.count_nonzero((d * d).sum(axis=1) < cutoff * cutoff))
    return total


def pair_profile(kernel):
    """The committed rocprofv3 figures of the pair kernel on this workload: HBM
    bytes per launch from the FETCH_SIZE / WRITE_SIZE passes (scaled by what
    the counters read for a 256 MiB copy) and its VALU issue from the SQ
    counter passes; None when no such profile is present."""
    path = os.path.join(ROOT, "profiles", "k_pair_win_93k_mixed_traffic.json")
    if kernel != "k_pair_win" or not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box here has one GPU and the path has no "
              "multi-GPU mode (--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from sponge_amd import lib, systems

    if args.workload == "water1m":
        sysd = systems.config_water_1m()
        # corner cells take the remainder of 215.24 / 6 A, about 1.9 times a
        # cell width per axis, and hold about 145 atoms: the capacity key of
        # the reference's neighbour list is raised from 150 to 256, the most
        # the windowed list path takes
        nb = lib.Nonbonded(sysd, refresh_interval=args.refresh, max_atom_in_grid_numbers=256)
    else:
        sysd = systems.config_gpu_93k()
        nb = lib.Nonbonded(sysd, grid=(96, 96, 96), refresh_interval=args.refresh)
    # the pair kernel reads the library's internal cell-ordered list; nothing
    # here reads the reference-layout copy
    nb.set_export(False)
    if args.no_overlap:
        nb.set_pme_overlap(False)
    torch.cuda.synchronize()

    profiling = [False]

    def step(k):
        e = (k % args.energy_interval) == 0
        # kernel averages are over force-only steps (the steady state); the
        # energy/virial steps still count in ms_per_step
        if profiling[0] and e:
            nb.profile(False)
        nb.step(step=k, need_energy=e, need_virial=e, forced_nl=(k == 0), fused=True)
        if profiling[0] and e:
            nb.profile(True)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    # kernel-group times from eager launches (HIP events on the library stream)
    k0 = args.warmup
    nb.profile(True)
    profiling[0] = True
    for g in (lib.PROF_PAIR, lib.PROF_NL, lib.PROF_PME, lib.PROF_EXCL):
        nb.profile_read(g, reset=True)
    for k in range(k0, k0 + args.profile_steps):
        step(k)
    torch.cuda.synchronize()
    prof = {name: nb.profile_read(g) for name, g in
            (("pair", lib.PROF_PAIR), ("nl", lib.PROF_NL), ("pme", lib.PROF_PME),
             ("excl", lib.PROF_EXCL))}
    nb.profile(False)
    profiling[0] = False
    k0 += args.profile_steps

    def kind(k):
        return ((k + 1) % args.refresh == 0, (k % args.energy_interval) == 0)

    def run_step(k):
        e = (k % args.energy_interval) == 0
        nb.step(step=k, need_energy=e, need_virial=e)

    # one graph per neighbour-list period (and per single step for the rest),
    # keyed by the kinds of their steps, captured before the timed region
    seq = None
    if not args.no_graph:
        seq = StepGraphs(nb, run_step, kind, args.refresh).plan(k0, args.steps)
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(nb.stream)
    if seq is not None:
        for g in seq:
            g.replay()
    else:
        for k in range(k0, k0 + args.steps):
            step(k)
    ev1.record(nb.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = ev0.elapsed_time(ev1)
    status = nb.status()
    counts = nb.internal_counts()
    e_list = int(counts.sum())
    n = sysd["n"]
    ms_per_step = ms / args.steps
    sps = 1000.0 / ms_per_step
    export_leg = None
    if seq is not None and args.export_steps > 0:
        # the drop-in default: rebuilds also write the reference-layout list
        # (rows in append order); the same graphs, captured again
        nb.set_export(1)
        k1 = k0 + args.steps
        gx = StepGraphs(nb, run_step, kind, args.refresh).plan(k1, args.export_steps)
        ex0 = torch.cuda.Event(enable_timing=True)
        ex1 = torch.cuda.Event(enable_timing=True)
        ex0.record(nb.stream)
        for g in gx:
            g.replay()
        ex1.record(nb.stream)
        torch.cuda.synchronize()
        xms = ex0.elapsed_time(ex1) / args.export_steps
        export_leg = {"steps": args.export_steps, "ms_per_step": xms,
                      "ns_per_day": NS_PER_DAY_PER_STEP_S * 1000.0 / xms,
                      "export": "reference-layout list written at every rebuild (rows in "
                                "append order, snb_nl_set_export mode 1)"}
        del gx
        nb.set_export(False)
    serial = None
    if not args.no_overlap and args.workload == "mixed93k":
        # the same kernel groups with the reciprocal pipeline (and, on these
        # force-only steps, the exclusion, 1-4 and bond terms) on the context
        # stream: each group alone, without the other stream beside it
        nb.set_pme_overlap(False)
        k2 = k0 + args.steps + max(args.export_steps, 0)
        for k in range(k2, k2 + 10):
            if (k + 1) % args.refresh and k % args.energy_interval:
                nb.step(step=k)
        torch.cuda.synchronize()
        nb.profile(True)
        for g in (lib.PROF_PAIR, lib.PROF_PME):
            nb.profile_read(g, reset=True)
        done = 0
        for k in range(k2 + 10, k2 + 1000):
            if (k + 1) % args.refresh == 0 or k % args.energy_interval == 0:
                continue
            nb.step(step=k)
            done += 1
            if done == 50:
                break
        torch.cuda.synchronize()
        sp, sq = nb.profile_read(lib.PROF_PAIR), nb.profile_read(lib.PROF_PME)
        nb.profile(False)
        nb.set_pme_overlap(True)
        serial = {"pair": 1000.0 * sp[0] / max(sp[1], 1), "pme": 1000.0 * sq[0] / max(sq[1], 1),
                  "steps": done,
                  "note": "eager force-only steps with both chains on one stream; "
                          "kernel_avg_us has them overlapped"}
    pair_ms, pair_n = prof["pair"]
    pair_avg_s = pair_ms / max(pair_n, 1) / 1000.0
    alg_bytes = 4.0 * e_list + 44.0 * n
    achieved = alg_bytes / pair_avg_s / 1e9
    if args.calibrate_copy:
        src = torch.ones(64 << 20, dtype=torch.float32, device=nb.dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    counts, rows = nb.internal_rows()
    base = None
    if args.cpu_seconds > 0:
        base, _ = cpu_baseline(sysd, nb, counts, rows, args.cpu_seconds)
    npairs = None if counts is None else pairs_in_cutoff(sysd, counts, rows, nb.cutoff)
    per = {k: (v[0] / max(v[1], 1)) for k, v in prof.items()}
    pair_kernel = "k_pair_win" if nb.window_info()["windowed"] else "k_lj_direct"
    prof_pair = pair_profile(pair_kernel) if args.workload == "mixed93k" else None
    out = {
        "metric": ("ns/day at 93k atoms; batched agg ns/day + p99 latency at force err<5e-5"
                   if args.workload == "mixed93k" else
                   "ns/day at 1,000,002 atoms (TIP3P water box)"),
        "value": sps * NS_PER_DAY_PER_STEP_S,
        "unit": "ns/day",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair/spline math, int64 fixed-point accumulation, f64 reductions",
        "data": ("synthetic (seeded generator: 333,334 TIP3P waters, 1.0 g/cm^3)"
                 if args.workload == "water1m" else
                 "synthetic (seeded generator: 30,000 TIP3P + 3x1,000-bead chains)"),
        "config": {"workload": ("1,000,002-atom TIP3P water box, plain nonbonded path (NL every "
                                "%d steps + exclusion + LJ/direct + PME 216^3), fixed "
                                "coordinates, no integrator" % args.refresh
                                if args.workload == "water1m" else
                                "93,000-atom mixed system, plain nonbonded path "
                                "(NL every %d steps + exclusion + LJ/direct + 1-4 pairs + chain "
                                "bonds + PME 96^3), fixed coordinates, no integrator"
                                % args.refresh),
                   "atoms": n, "box_A": float(sysd["box"][0]), "cutoff_A": 10.0, "skin_A": 2.0,
                   "pme_grid": list(nb.grid), "dt_fs": 2.0, "nl_entries": e_list,
                   "energy_interval": args.energy_interval},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (prof_pair["bytes_per_launch"] if prof_pair else None),
                     "limiter": "VALU issue" if prof_pair else None,
                     "valu_issue": None if not prof_pair else {
                         "frac": prof_pair["valu_issue_frac"],
                         "valu_wave_instructions_per_launch": prof_pair["sq_insts_valu"],
                         "source": "profiles/k_pair_win_93k_mixed.md (SQ_INSTS_VALU over "
                                   "256 CUs x GRBM_GUI_ACTIVE cycles per XCD)"},
                     "kernel": pair_kernel, "alg_bytes_per_launch": alg_bytes,
                     "avg_launch_us": pair_avg_s * 1e6, "launches": pair_n,
                     "pairs_in_cutoff": npairs,
                     "pair_interactions_per_s": None if npairs is None else npairs / pair_avg_s},
        "kernel_avg_us": {k: v * 1000.0 for k, v in per.items()},
        "kernel_avg_us_alone": serial,
        "pair_window": nb.window_info(),
        "launch": ("hip-graph replay, one graph per neighbour-list period of %d steps"
                   % args.refresh) if seq is not None else "eager",
        "export_on": export_leg,
        "status_flags": status,
        "wall_s_timed": wall,
        "cpu_baseline": base,
    }
    if args.md_steps > 0:
        out["md"] = md_phase(nb, sysd, args, kind)
    if not args.no_accuracy:
        out.update(frame_accuracy(nb, sysd, "after the md leg's %d steps" % (args.md_steps + 20)
                                  if args.md_steps > 0 else "the initial coordinates"))
    if args.nve_steps > 0:
        out["nve"] = nve_phase(nb, sysd, args)
    nb.close()
    if args.serve_requests > 0:
        out["serving"] = serving_phase(args)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
