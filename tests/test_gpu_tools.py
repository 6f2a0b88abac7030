"""The C++ adapter and the precompute driver on the device."""
import os
import subprocess

import numpy as np
import pytest

from helpers import REF_G, REF_S, REF_SCALE, ROOT, bh_step_dynamic, rel_err_rows

pytestmark = pytest.mark.gpu

INC = os.path.join(ROOT, "include")
PKG = os.path.join(ROOT, "procedural-universe_amd")


def _build_example(tmp_path):
    exe = str(tmp_path / "adapter_example")
    src = os.path.join(ROOT, "tests", "cpp", "adapter_example.cpp")
    subprocess.run(["g++", "-std=c++17", "-I", INC, src, "-o", exe, "-L", PKG, "-lnbody", f"-Wl,-rpath,{PKG}"],
                   check=True)
    return exe


@pytest.mark.parametrize("kind", [0, 2])
def test_cpp_adapter_update_matches_oracle(gpu, oracle, tmp_path, kind):
    pu = gpu
    exe = _build_example(tmp_path)
    n, steps = 512, 3
    out = subprocess.run([exe, str(kind), str(n), str(steps)], capture_output=True, text=True, timeout=120, check=True)
    lines = out.stdout.strip().splitlines()
    assert lines[0] == ("Brute Force CPU" if kind == 0 else "Barnes-Hut")
    vals = np.array([[float(x) for x in l.split()] for l in lines[1:]])
    pos, vel, mass = pu.ic_uniform(n, 11, half=500.0, mass_lo=1e20, mass_hi=1e30)
    dt = np.float32(0.02 / 60)
    if kind == 0:
        p_ref, v_ref = oracle.ref_step(pos, np.zeros((n, 3)), mass, dt, steps)
        assert np.all(vals[:, 6] == 0)  # brute force leaves Forces zero
    else:
        p_ref, v_ref = oracle.bh_step(pos, np.zeros((n, 3)), mass, dt, steps, 2.0, [-4000] * 3, 8000.0)
        assert np.any(vals[:, 6] != 0)  # Barnes-Hut leaves the last forces
    cells = out.stderr.split()
    if kind == 0:
        assert cells[:2] == ["cells", "0"]
    else:
        assert int(cells[1]) > n // 8 and int(cells[3]) == n  # the root holds every particle
    assert rel_err_rows(vals[:, :3], p_ref).max() < 1e-4


def test_precompute_driver_writes_snapshot(gpu, oracle, tmp_path):
    """`-c` seeds with the StarSystem seeder as RunSimulation does
    (SimulationState.cpp:286-290) and steps Barnes-Hut with the reference
    constants; the snapshot matches the restated BarnesHut::Update run from the
    same initial state on the same fixed root with FP64 cell centres of mass,
    the device tree's arithmetic (positions and velocities 1e-4; with the
    reference's FP32 centres some opening decisions at theta 2 flip, and a
    flipped cell moves a body's velocity by up to ~3e-4)."""
    pu = gpu
    exe = os.path.join(PKG, "nbody_precompute")
    out = str(tmp_path / "run.nbody")
    n, steps, seed = 3000, 4, 0x5EED0006
    r = subprocess.run([exe, "-c", "-p", str(n), "--steps", str(steps), "--out", out, "--method", "bh",
                        "--seed", str(seed)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = pu.read_snapshot(out, recentre=False)
    # the seeder's initial state, drawn again on the device
    with pu.NBodySim(pu.reference_config(pu.METHOD_BARNES_HUT)) as sim:
        sim.init_seeded(pu.SEEDER_STAR_SYSTEM, n, seed)
        p0, v0, _ = sim.read_soa()
        init = np.zeros(n, dtype=pu.PARTICLE_DTYPE)
        sim.read_particles(init, pu.FIELD_MASS)
    mass = init["mass"]
    assert mass[0] == 1e30 and np.all(p0[1:, :2] == 0)  # the star, and bodies on the z axis
    dt = np.float32(0.02) * np.float32(1.0 / 60.0)
    p_ref, v_ref = bh_step_dynamic(oracle, p0, v0, mass, dt, steps, 2.0, G=REF_G, S=REF_S, scale=REF_SCALE,
                                   root=([-4000.0] * 3, 8000.0))
    assert np.array_equal(got["mass"], mass)
    perr = rel_err_rows(got["position"], p_ref)
    verr = rel_err_rows(got["velocity"], v_ref)
    print(f"precompute (star system, {steps} steps): position max rel {perr.max():.2e}, velocity {verr.max():.2e}")
    assert perr.max() < 1e-4 and verr.max() < 1e-4
    # resume from the snapshot (-f) for two more steps
    out2 = str(tmp_path / "run2.nbody")
    r = subprocess.run([exe, "-f", out, "--steps", "2", "--out", out2], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert len(pu.read_snapshot(out2)) == 3000


def test_benchmark_table(gpu):
    """nbody_benchmark: RunBenchmark's table, one line per ENBodySim."""
    exe = os.path.join(PKG, "nbody_benchmark")
    r = subprocess.run([exe, "-p", "4096", "--seeder", "galaxy", "--frames", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l for l in r.stdout.splitlines() if not l.startswith("#")]
    names = [l[:16].strip() for l in rows]
    assert names == ["Brute Force CPU", "Brute Force GPU", "Barnes-Hut"]
    # what ran: the library has no CPU simulator
    assert [l[17:35].strip() for l in rows] == ["device direct sum", "device direct sum", "device Barnes-Hut"]
    assert all(float(l.split()[-2]) > 0 for l in rows)


def test_adapter_thousand_frames_against_oracle(gpu, oracle):
    """Update() through the adapter 1,000 times (reference constants, a
    RandomSeeder-like set): the device keeps FP32 velocities where the
    reference keeps FP64 ones, and the caller's FP64 Velocity fields receive
    the FP32 values every frame (the device state never reads them back).
    After 1,000 frames the positions stay within 1e-4 of the restated
    BruteForceCPU::Update and the velocities within 1e-4."""
    pu = gpu
    n, frames = 512, 1000
    pos, _, mass = pu.ic_uniform(n, 0x5EED0400, half=500.0, mass_lo=1e20, mass_hi=1e30)
    vel = (pos / np.linalg.norm(pos, axis=1, keepdims=True) * 1e16).astype(np.float32)
    parts = pu.particles_from_soa(pos, vel.astype(np.float64), mass)
    dt = np.float32(0.02 / 60)
    sim = pu.create_nbody_sim(pu.ENBodySim.BruteForceCPU)
    sim.init(parts)
    for _ in range(frames):
        sim.update(dt)
    sim.close()
    p_ref, v_ref = oracle.ref_step(pos, vel.astype(np.float64), mass, dt, frames)
    perr = rel_err_rows(parts["position"], p_ref)
    verr = rel_err_rows(parts["velocity"], v_ref)
    print(f"1000 frames: position max rel {perr.max():.2e}, velocity max rel {verr.max():.2e}")
    assert perr.max() < 1e-4 and verr.max() < 1e-4
