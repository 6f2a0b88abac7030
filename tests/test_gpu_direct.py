"""Parity of the device direct sum with the CPU oracle (runs on the MI355X).

Tolerances (law R is the reference's law, src/Sim/Physics.hpp:25-35):
  * accelerations vs the restated BruteForceCPU (FP32 geometry, FP64 sums):
    max per-particle relative error 1e-5;
  * accelerations vs the FP64 direct-sum oracle: RMS relative error 1e-6;
  * positions after 10 kick-drift steps vs the restated BruteForceCPU::Update:
    max per-particle relative error 1e-4.
"""
import numpy as np
import pytest

from helpers import REF_G, REF_S, REF_SCALE, rel_err_rows

pytestmark = pytest.mark.gpu

ACC_TOL = 1e-5
POS_TOL = 1e-4
RMS_TOL = 1e-6


def _sim(pu, cfg, pos, vel, mass):
    sim = pu.NBodySim(cfg)
    sim.init_soa(pos, vel, mass)
    return sim


def test_reference_constants_accelerations(gpu, oracle):
    """Reference constants and a RandomSeeder-like input (uniform in +-500, masses
    U(1e20, 1e30), src/Sim/RandomSeeder.cpp:16-19)."""
    pu = gpu
    n = 4096
    pos, vel, mass = pu.ic_uniform(n, 0x5EED0001, half=500.0, mass_lo=1e20, mass_hi=1e30)
    with _sim(pu, pu.reference_config(pu.METHOD_DIRECT), pos, vel, mass) as sim:
        a = sim.accelerations()
    ref = oracle.ref_accel(pos, mass)
    err = rel_err_rows(a, ref)
    print(f"reference constants N={n}: max rel {err.max():.2e} rms {np.sqrt(np.mean(err**2)):.2e}")
    assert err.max() < ACC_TOL


def test_reference_constants_ten_steps(gpu, oracle):
    pu = gpu
    n = 2048
    pos, _, mass = pu.ic_uniform(n, 0x5EED0003, half=500.0, mass_lo=1e20, mass_hi=1e30)
    # radial velocities of magnitude 1e16 as RandomSeeder gives (RandomSeeder.cpp:27-33)
    vel = (pos / np.linalg.norm(pos, axis=1, keepdims=True) * 1e16).astype(np.float32)
    dt = np.float32(0.02 / 60)  # the -c CLI's default step (src/App/NBody.cpp:56,87)
    with _sim(pu, pu.reference_config(pu.METHOD_DIRECT), pos, vel, mass) as sim:
        sim.step(dt, 10)
        p, v, _ = sim.read_soa()
    p_ref, v_ref = oracle.ref_step(pos, vel.astype(np.float64), mass, dt, 10)
    perr = rel_err_rows(p, p_ref)
    verr = rel_err_rows(v, v_ref)
    print(f"10 steps: pos max rel {perr.max():.2e}, vel max rel {verr.max():.2e}")
    assert perr.max() < POS_TOL
    assert verr.max() < 1e-5


@pytest.mark.parametrize("law", [0, 1])
def test_plummer_config_accelerations_vs_fp64(gpu, oracle, law):
    """N = 4096 equal-mass Plummer sphere, eps = 0.01 (S = eps^2 for law R)."""
    pu = gpu
    n, soft = 4096, 1e-4
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0001)
    with _sim(pu, pu.default_config(law=law, softening=soft), pos, vel, mass) as sim:
        a = sim.accelerations()
    a64 = oracle.fp64_accel(law, pos, mass, 1.0, soft)
    err = rel_err_rows(a, a64)
    rms = np.sqrt(np.mean(err ** 2))
    print(f"law {law}: rms {rms:.2e} max {err.max():.2e}")
    assert rms < RMS_TOL
    assert err.max() < ACC_TOL
    if law == 0:
        ref = oracle.ref_accel(pos, mass, G=1.0, S=soft)
        assert rel_err_rows(a, ref).max() < ACC_TOL


def test_plummer_config_100_steps(gpu, oracle):
    """The CPU-path configuration: N = 4096 Plummer, eps = 0.01, dt = 1e-3, 100
    steps; positions at step 10 to 1e-4 and energy drift equal to the CPU path's."""
    pu = gpu
    n, soft, dt = 4096, 1e-4, np.float32(1e-3)
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0001)
    with _sim(pu, pu.default_config(law=pu.LAW_R, softening=soft), pos, vel, mass) as sim:
        d0 = sim.diagnostics()
        sim.step(dt, 10)
        p10, _, _ = sim.read_soa()
        sim.step(dt, 90)
        p100, v100, _ = sim.read_soa()
        d1 = sim.diagnostics()
    p_ref10, v_ref10 = oracle.ref_step(pos, vel.astype(np.float64), mass, dt, 10, G=1.0, S=soft, scale=1.0)
    assert rel_err_rows(p10, p_ref10).max() < POS_TOL
    p_ref, v_ref = oracle.ref_step(p_ref10, v_ref10, mass, dt, 90, G=1.0, S=soft, scale=1.0)
    e_ref0 = oracle.fp64_energy(0, pos, vel, mass, 1.0, soft)["total"]
    e_ref1 = oracle.fp64_energy(0, p_ref, v_ref.astype(np.float32), mass, 1.0, soft)["total"]
    drift_gpu = (d1["total"] - d0["total"]) / abs(d0["total"])
    drift_ref = (e_ref1 - e_ref0) / abs(e_ref0)
    print(f"100 steps: pos max rel {rel_err_rows(p100, p_ref).max():.2e}; drift gpu {drift_gpu:.3e} ref {drift_ref:.3e}")
    assert rel_err_rows(p100, p_ref).max() < 1e-3
    assert abs(drift_gpu - drift_ref) < 1e-4


def test_diagnostics_match_fp64_oracle(gpu, oracle):
    pu = gpu
    n = 2048
    pos, vel, mass = pu.ic_plummer(n, 77)
    for law, soft in ((0, 1e-4), (1, 1e-4), (0, 0.0)):
        if law == 1 and soft == 0:
            continue
        with _sim(pu, pu.default_config(law=law, softening=soft), pos, vel, mass) as sim:
            d = sim.diagnostics()
        ref = oracle.fp64_energy(law, pos, vel, mass, 1.0, soft)
        assert d["kinetic"] == pytest.approx(ref["kinetic"], rel=1e-12)
        assert d["potential"] == pytest.approx(ref["potential"], rel=1e-10)
        assert np.allclose(d["momentum"], ref["momentum"], atol=1e-12)
        assert np.allclose(d["angular_momentum"], ref["angular_momentum"], atol=1e-12)
        assert d["mass"] == pytest.approx(1.0, rel=1e-12)


@pytest.mark.parametrize("n", [1, 2, 3, 255, 257, 1000])
def test_ragged_and_tiny_sizes(gpu, oracle, n):
    pu = gpu
    pos, vel, mass = pu.ic_uniform(n, 100 + n, half=50.0, mass_lo=1e24, mass_hi=1e26)
    with _sim(pu, pu.reference_config(pu.METHOD_DIRECT), pos, vel, mass) as sim:
        a = sim.accelerations()
        sim.step(np.float32(0.02 / 60), 3)
        p, v, _ = sim.read_soa()
    ref = oracle.ref_accel(pos, mass)
    if n == 1:
        assert np.all(a == 0)
    else:
        assert rel_err_rows(a, ref).max() < ACC_TOL
    p_ref, _ = oracle.ref_step(pos, vel.astype(np.float64), mass, np.float32(0.02 / 60), 3)
    assert rel_err_rows(p, p_ref).max() < POS_TOL


def test_coincident_particles(gpu, oracle):
    """Coincident pairs contribute nothing (Normalize of a zero vector is zero)."""
    pu = gpu
    pos, vel, mass = pu.ic_uniform(64, 5, half=10.0, mass_lo=1e25, mass_hi=1e25)
    pos[10] = pos[11] = pos[12]
    pos[40:48] = pos[3]
    with _sim(pu, pu.reference_config(pu.METHOD_DIRECT), pos, vel, mass) as sim:
        a = sim.accelerations()
    assert np.all(np.isfinite(a))
    ref = oracle.ref_accel(pos, mass)
    assert rel_err_rows(a, ref).max() < ACC_TOL
    same = np.zeros((4, 3), np.float32)
    with _sim(pu, pu.reference_config(pu.METHOD_DIRECT), same, None, np.full(4, 1e30)) as sim:
        assert np.all(sim.accelerations() == 0)


def test_empty_state(gpu):
    pu = gpu
    with pu.NBodySim(pu.default_config()) as sim:
        sim.init_soa(np.zeros((0, 3), np.float32), None, np.zeros(0))
        sim.step(0.1, 5)
        p, v, a = sim.read_soa()
        assert p.shape == (0, 3)
        assert sim.diagnostics()["mass"] == 0


@pytest.mark.parametrize(
    "shape", [(1, 256, 1), (2, 256, 1), (4, 128, 1), (8, 64, 1), (2, 128, 4), (1, 64, 16), (4, 256, 3)]
)
def test_launch_shapes_agree(gpu, oracle, shape):
    pu = gpu
    k, b, split = shape
    n, soft = 3000, 1e-4
    pos, vel, mass = pu.ic_plummer(n, 31)
    cfg = pu.default_config(law=pu.LAW_R, softening=soft, tune_i_per_lane=k, tune_block=b, tune_split=split)
    with _sim(pu, cfg, pos, vel, mass) as sim:
        a = sim.accelerations()
        sim.step(1e-3, 2)
        p, _, _ = sim.read_soa()
    ref = oracle.ref_accel(pos, mass, G=1.0, S=soft)
    assert rel_err_rows(a, ref).max() < ACC_TOL
    p_ref, _ = oracle.ref_step(pos, vel.astype(np.float64), mass, 1e-3, 2, G=1.0, S=soft, scale=1.0)
    assert rel_err_rows(p, p_ref).max() < POS_TOL


def test_bitwise_deterministic(gpu):
    pu = gpu
    pos, vel, mass = pu.ic_plummer(5000, 8)
    out = []
    for _ in range(2):
        with _sim(pu, pu.default_config(law=pu.LAW_R, softening=1e-4), pos, vel, mass) as sim:
            sim.step(1e-3, 5)
            out.append(sim.read_soa())
    for x, y in zip(out[0], out[1]):
        assert np.array_equal(x, y)


def test_adapter_update_in_place(gpu, oracle):
    """INBodySim::Update semantics: the adopted records change in place; the
    direct sum leaves Forces zero (BruteForceCPU.cpp:72)."""
    pu = gpu
    n = 300
    pos, vel, mass = pu.ic_uniform(n, 9, half=200.0, mass_lo=1e22, mass_hi=1e26)
    parts = pu.particles_from_soa(pos, vel, mass)
    parts["colour"] = 0.5
    sim = pu.create_nbody_sim(pu.ENBodySim.BruteForceCPU)
    sim.init(parts)
    dt = np.float32(0.02 / 60)
    for _ in range(3):
        sim.update(dt)
    p_ref, v_ref = oracle.ref_step(pos, vel.astype(np.float64), mass, dt, 3)
    assert rel_err_rows(parts["position"], p_ref).max() < POS_TOL
    assert np.all(parts["forces"] == 0)
    assert np.all(parts["colour"] == 0.5)
    # FORCES read-back is m * a of the last force evaluation, in FP64
    _, _, acc_last = sim.read_soa()
    sim.read_particles(parts, pu.FIELD_FORCES)
    assert np.array_equal(parts["forces"], mass[:, None] * acc_last.astype(np.float64))
    a_now = sim.accelerations()
    ref_now = oracle.ref_accel(parts["position"].astype(np.float32), mass)
    assert rel_err_rows(a_now, ref_now).max() < ACC_TOL
    sim.close()
    assert pu.nbody_sim_get_name(pu.ENBodySim.BarnesHut) == "Barnes-Hut"


def test_large_n_properties(gpu, oracle):
    """N = 1,048,576 (the benchmark input): a 512-row FP64 sample and Newton's
    third law on the whole system (size-independent checks)."""
    pu = gpu
    n, soft = 1 << 20, 0.0025
    pos, vel, mass = pu.ic_disc_bulge(n, 0x5EED0002)
    with _sim(pu, pu.default_config(law=pu.LAW_R, softening=soft), pos, vel, mass) as sim:
        a = sim.accelerations()
    rows = np.linspace(0, n - 1, 512).astype(np.int64)
    a64 = oracle.fp64_accel(0, pos, mass, 1.0, soft, rows=rows)
    err = rel_err_rows(a[rows], a64)
    rms = np.sqrt(np.mean(err ** 2))
    print(f"N=1M: rms {rms:.2e} max {err.max():.2e}")
    assert rms < RMS_TOL
    assert err.max() < ACC_TOL
    f = a.astype(np.float64) * mass[:, None]
    assert np.all(np.abs(f.sum(0)) < 1e-5 * np.abs(f).sum(0))


def test_errors(gpu):
    pu = gpu
    sim = pu.NBodySim(pu.default_config())
    with pytest.raises(pu.NBodyError) as e:
        sim.step(0.1, 1)
    assert e.value.status == 4  # NBODY_ERR_NOT_INITIALISED
    sim.init_soa(np.zeros((4, 3), np.float32), None, np.ones(4))
    with pytest.raises(pu.NBodyError) as e:
        sim.step(float("nan"), 1)
    assert e.value.status == 1
    sim.close()
