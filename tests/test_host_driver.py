"""The C++ host (sponge2.0_amd/host): SPONGE's module interfaces over the C
ABI and the MD driver bin/sponge_nb_md, run on native input files and an mdin
file, writing mdout and the force trajectory.

- CPU: the driver reads the mdin and every native file before it touches the
  device (a missing input is reported by name; with valid inputs it stops at
  device set-up on a machine without a GPU).
- GPU: the mdout columns (the reference's names and formats) and the
  frc.dat frames (float32 N x 3, no header) read back against the library
  driven from Python on the same files: step-0 forces bit for bit (the same
  module calls in the same order), step-0 energies to the printed precision,
  one frame per write interval.
"""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
DRIVER = os.path.join(ROOT, "sponge2.0_amd", "bin", "sponge_nb_md")

MDIN = """driver test: rigid TIP3P water, middle Langevin
mode = NVT
dt = 0.001
step_limit = 20
write_mdout_interval = 5
write_trajectory_interval = 5
frc = frc.dat
cutoff = 8.0
default_in_file_prefix = sys
thermostat = middle_langevin
thermostat_tau = 0.1
thermostat_seed = 7
target_temperature = 300
constrain_mode = SETTLE
neighbor_list
{
    refresh_interval = 10
}
PM
{
    print_detail = 1
}
"""


def _water_system(tmp):
    """648-atom TIP3P box with the rigid waters as bonds (O-H, O-H, H-H at
    the TIP3P lengths: the constraint pairs SETTLE's triangles come from),
    written as native files with prefix sys plus a velocity file."""
    from sponge_amd import native_io, systems
    s = systems.tip3p_box(216, 18.7, seed=7, cutoff=8.0)
    a, b, c, rab, rac, rbc = systems.water_triangles(s)
    s["bonds"] = (np.concatenate([a, a, b]).astype(np.int32),
                  np.concatenate([b, c, c]).astype(np.int32),
                  np.full(3 * len(a), 553.0, np.float32),
                  np.concatenate([rab, rac, rbc]).astype(np.float32))
    native_io.write_system(str(tmp), "sys", s)
    v = systems.maxwell_velocities(s["mass"], 300.0, 11)
    with open(os.path.join(str(tmp), "sys_velocity.txt"), "w") as f:
        f.write("%d\n" % s["n"])
        np.savetxt(f, v, fmt="%.9e")
    with open(os.path.join(str(tmp), "mdin.txt"), "w") as f:
        f.write(MDIN)
    return native_io.system_from_prefix(str(tmp), "sys")


def _run(tmp, *extra, timeout=120):
    return subprocess.run([DRIVER, "-mdin", "mdin.txt", *extra], cwd=str(tmp),
                          capture_output=True, text=True, timeout=timeout)


def test_driver_reads_inputs_before_the_device(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.skip("driver not built")
    _water_system(tmp_path)
    r = _run(tmp_path, "-dump_commands", "1")
    assert r.returncode == 0
    table = dict(line.split(" = ", 1) for line in r.stdout.splitlines())
    assert table["PM_print_detail"] == "1" and table["neighbor_list_refresh_interval"] == "10"
    assert table["mode"] == "NVT" and table["md_name"].startswith("driver test")
    r = _run(tmp_path, "-coordinate_in_file", "missing.txt")
    assert r.returncode != 0 and "missing.txt" in r.stderr
    import torch
    if torch.cuda.is_available():
        return
    r = _run(tmp_path)
    assert r.returncode != 0 and "no usable HIP device" in r.stderr, r.stderr


def _mdout(path):
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    return lines[0], [dict(zip(lines[0], row)) for row in lines[1:]]


@pytest.mark.gpu
def test_driver_mdout_and_force_trajectory(tmp_path):
    from sponge_amd.lib import Nonbonded
    s = _water_system(tmp_path)
    r = _run(tmp_path)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    keys, rows = _mdout(os.path.join(str(tmp_path), "mdout.txt"))
    assert keys == ["step", "time", "temperature", "potential", "bond", "LJ_short", "LJ_long",
                    "LJ", "PM", "PM_direct", "PM_reciprocal", "PM_self", "PM_correction"]
    assert [int(r["step"]) for r in rows] == [0, 5, 10, 15, 20]
    assert [r["time"] for r in rows] == ["0.000", "0.005", "0.010", "0.015", "0.020"]
    n = s["n"]
    frames = np.fromfile(os.path.join(str(tmp_path), "frc.dat"), dtype=np.float32)
    assert frames.size == 5 * n * 3
    frames = frames.reshape(5, n, 3)
    assert np.isfinite(frames).all()
    # step 0 from Python: the same calls (exclusion, LJ, long range, bonds,
    # reciprocal), each flushing into the float forces
    nb = Nonbonded(s, cutoff=8.0, refresh_interval=10)
    nb.zero_outputs()
    nb.step(step=0, need_energy=True, forced_nl=True, fused=False)
    nb.sync()
    np.testing.assert_array_equal(frames[0], nb.frc.cpu().numpy())
    e = nb.energies()
    e["bond"] = nb.bond_energy()
    for k in ("LJ_short", "LJ_long", "LJ", "PM", "PM_direct", "PM_reciprocal", "PM_self",
              "PM_correction", "bond"):
        assert abs(float(rows[0][k]) - e[k]) <= 0.0051 + 1e-6 * abs(e[k]), (k, rows[0][k], e[k])
    pot = e["LJ"] + e["PM"] + e["bond"]
    assert abs(float(rows[0]["potential"]) - pot) <= 0.02 + 1e-6 * abs(pot)
    # Maxwell velocities at 300 K; the lattice start then releases its
    # electrostatic strain as heat within the first tens of femtoseconds
    temps = [float(r["temperature"]) for r in rows]
    assert 200.0 < temps[0] < 400.0 and all(np.isfinite(temps)), temps
    assert not np.array_equal(frames[0], frames[1])
    nb.close()
