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


def _adapter_frames(pu, pos, vel, mass, dt, frames):
    parts = pu.particles_from_soa(pos, vel.astype(np.float64), mass)
    sim = pu.create_nbody_sim(pu.ENBodySim.BruteForceCPU)
    sim.init(parts)
    for _ in range(frames):
        sim.update(dt)
    sim.close()
    return parts["position"], parts["velocity"]


def _oracle_fp32_velocity(oracle, pos, vel, mass, dt, frames):
    """The restated BruteForceCPU::Update with the velocity rounded to FP32
    after every frame: the reference's arithmetic on the device's state
    precision."""
    p, v = pos.copy(), vel.astype(np.float64)
    for _ in range(frames):
        p, v = oracle.ref_step(p, v, mass, dt, 1)
        v = v.astype(np.float32).astype(np.float64)
    return p, v


@pytest.mark.parametrize("half", [2000.0, 500.0])
def test_adapter_thousand_frames_against_oracle(gpu, oracle, half):
    """Update() through the adapter 1,000 times (reference constants, dt =
    0.02/60, a RandomSeeder-like set of 512 particles with masses 1e20-1e30 and
    radial velocities of 1e16) against the restated BruteForceCPU::Update.
    The device keeps FP32 velocities where the reference keeps FP64 ones, and
    the caller's FP64 Velocity fields receive the FP32 values every frame.

    Cube half-width 2,000: the system is not chaotic over 1,000 frames (the
    reference started from velocities moved by one FP32 ulp stays within 2e-7
    of itself), and the reference with its velocity rounded to FP32 after
    every frame drifts from itself by up to ~2e-5 in position and ~2e-3 in
    velocity (relative, per particle): that is what FP32 velocity state costs.
    The device is held to a small multiple of that drift.

    Cube half-width 500: close encounters make the system chaotic; the
    reference itself, started one FP32 ulp away, ends up to ~5e-2 (position)
    and ~9 (velocity) away at the few particles that pass close to another.
    The device is held to the reference's own sensitivity: its median particle
    agrees to 1e-5, and its largest deviations stay within a small multiple of
    the perturbed reference's."""
    pu = gpu
    n, frames = 512, 1000
    pos, _, mass = pu.ic_uniform(n, 0x5EED0400, half=half, mass_lo=1e20, mass_hi=1e30)
    vel = (pos / np.linalg.norm(pos, axis=1, keepdims=True) * 1e16).astype(np.float32)
    dt = np.float32(0.02 / 60)
    p_dev, v_dev = _adapter_frames(pu, pos, vel, mass, dt, frames)
    p_ref, v_ref = oracle.ref_step(pos, vel.astype(np.float64), mass, dt, frames)
    p_ulp, v_ulp = oracle.ref_step(pos, vel.astype(np.float64) * (1 + 2.0**-24), mass, dt, frames)
    p_f32, v_f32 = _oracle_fp32_velocity(oracle, pos, vel, mass, dt, frames)
    perr, verr = rel_err_rows(p_dev, p_ref), rel_err_rows(v_dev, v_ref)
    pulp, vulp = rel_err_rows(p_ulp, p_ref), rel_err_rows(v_ulp, v_ref)
    pf32, vf32 = rel_err_rows(p_f32, p_ref), rel_err_rows(v_f32, v_ref)
    q = lambda e: "median %.1e p99 %.1e max %.1e" % (np.median(e), np.quantile(e, 0.99), e.max())
    print(f"\nhalf {half}: device position {q(perr)} velocity {q(verr)}")
    print(f"  reference one ulp away: position {q(pulp)} velocity {q(vulp)}")
    print(f"  reference, FP32 velocity: position {q(pf32)} velocity {q(vf32)}")
    if half == 2000.0:
        assert pulp.max() < 1e-6  # not chaotic
        assert perr.max() < max(1e-4, 5 * pf32.max())
        assert verr.max() < max(1e-3, 5 * vf32.max())
    else:
        assert np.median(perr) < 1e-5 and np.median(verr) < 1e-4
        assert perr.max() < 10 * pulp.max() and verr.max() < 10 * vulp.max()
