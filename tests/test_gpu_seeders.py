"""The reference's particle seeders drawn on the device (nbody_init_seeded):
each distribution's defining properties (RandomSeeder.cpp:14-40,
GalaxySeeder.cpp:42-144, StarSystemSeeder.cpp:17-55), reproducibility, and a
seeded galaxy through the Barnes-Hut parity check."""
import numpy as np
import pytest

from helpers import rel_err_rows

pytestmark = pytest.mark.gpu


def _seeded(pu, kind, n, seed=5, scale=1.0):
    sim = pu.NBodySim(pu.reference_config(pu.METHOD_DIRECT))
    sim.init_seeded(kind, n, seed, scale)
    pos, vel, _ = sim.read_soa()
    return sim, pos.astype(np.float64), vel.astype(np.float64)


def test_random_seeder(gpu):
    pu = gpu
    n, scale = 20000, 2.0
    sim, pos, vel = _seeded(pu, pu.SEEDER_RANDOM, n, scale=scale)
    d = sim.diagnostics()
    sim.close()
    assert np.all(np.abs(pos) <= 500.0 / scale)
    speed = np.linalg.norm(vel, axis=1)
    assert np.allclose(speed, 1e16, rtol=1e-6)
    cosang = (pos * vel).sum(1) / (np.linalg.norm(pos, axis=1) * speed)
    assert cosang.min() > 1 - 1e-5  # radial
    assert d["mass"] / n == pytest.approx(0.5 * (1e20 + 1e30), rel=0.02)  # uniform in [1e20, 1e30]


def test_galaxy_seeder(gpu):
    pu = gpu
    n = 30000
    sim, pos, vel = _seeded(pu, pu.SEEDER_GALAXY, n)
    d = sim.diagnostics()
    sim.close()
    per_loop = int(np.floor(np.float32(n) * np.float32(0.4) / np.float32(60.0)))
    arms = 2 * 61 * per_loop  # the float angle loop runs 61 times
    disc = n - arms
    # masses: 1e20 in the arms, uniform in [1e28, 1e30] in the disc
    assert d["mass"] == pytest.approx(arms * 1e20 + disc * 0.5 * (1e28 + 1e30), rel=0.03)
    assert np.all(vel[:, 2] == 0.0)  # velocities are not rotated: in the seeding plane
    # the galaxy is flat: thickness of the normal(0, 16) heights along its thinnest principal axis
    c = pos - pos.mean(0)
    w, v = np.linalg.eigh(c.T @ c / n)
    assert 14.0 < np.sqrt(w[0]) < 18.0
    assert np.sqrt(w[1]) > 150.0
    # disc particles (the last ones) lie within 720 of the centre
    assert np.linalg.norm(pos[arms:], axis=1).max() <= 720.0 * (1 + 1e-6)
    # disc speeds: (y, -x, 0) * uniform(0.8, 1.2) * 1e14 with |(x, y)| <= 720
    assert np.linalg.norm(vel[arms:], axis=1).max() <= 1.2e14 * 720.0 * (1 + 1e-6)
    # arm speeds: 2e16 * 1000 / |spiral - spiraln| along the arm's normal
    assert np.linalg.norm(vel[:arms], axis=1).min() > 1e17


def test_star_system_seeder(gpu):
    pu = gpu
    n = 2000
    sim, pos, vel = _seeded(pu, pu.SEEDER_STAR_SYSTEM, n)
    d = sim.diagnostics()
    sim.close()
    assert np.all(pos[0] == 0) and np.all(vel[0] == 0)
    assert np.all(pos[1:, :2] == 0)
    au_m = 1.15e12 * 1000.0
    z = pos[1:, 2]
    assert z.min() >= 4 * au_m / (20 * 1.15e12) * (1 - 1e-6) and z.max() <= 7 * au_m / (20 * 1.15e12) * (1 + 1e-6)
    assert vel[1:, 0].min() >= 1 * au_m * (1 - 1e-6) and vel[1:, 0].max() <= 5 * au_m * (1 + 1e-6)
    assert np.allclose(vel[1:, 1].max(), 5 * au_m / 20, rtol=0.05) and np.all(vel[:, 2] == 0)
    assert d["mass"] == pytest.approx(1e30 + (n - 1) * 0.5 * (1e10 + 1e26), rel=0.05)


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_seeders_reproducible(gpu, kind):
    pu = gpu
    a = _seeded(pu, kind, 5000, seed=11)
    b = _seeded(pu, kind, 5000, seed=11)
    c = _seeded(pu, kind, 5000, seed=12)
    for s in (a[0], b[0], c[0]):
        s.close()
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])
    if kind != 2:
        assert not np.array_equal(a[1], c[1])


def test_seeded_galaxy_barnes_hut_parity(gpu, oracle):
    """A seeded galaxy on the reference's fixed root with theta 2.0 against the
    restated reference octree (FP64 centres of mass)."""
    pu = gpu
    n = 8192
    sim = pu.NBodySim(pu.reference_config(pu.METHOD_BARNES_HUT))
    sim.init_seeded(pu.SEEDER_GALAXY, n, 3)
    pos, _, _ = sim.read_soa()
    parts = np.zeros(n, dtype=pu.PARTICLE_DTYPE)
    a = sim.accelerations()
    sim.read_particles(parts, pu.FIELD_FORCES)
    sim.close()
    # the masses (drawn on the device) from the FORCES read-back m a
    mass = np.linalg.norm(parts["forces"], axis=1) / np.linalg.norm(a.astype(np.float64), axis=1)
    F, st = oracle.bh_forces(pos, mass, 2.0, [-4000.0] * 3, 8000.0, com_fp64=True)
    assert st["dropped"] == 0
    assert rel_err_rows(a, F / mass[:, None]).max() < 1e-5


def test_read_back_masses(gpu):
    """NBODY_FIELD_MASS returns the device-drawn masses (their sum is the FP64 diagnostic's)."""
    pu = gpu
    n = 3000
    sim, _, _ = _seeded(pu, pu.SEEDER_GALAXY, n)
    parts = np.zeros(n, dtype=pu.PARTICLE_DTYPE)
    sim.read_particles(parts, pu.FIELD_MASS)
    d = sim.diagnostics()
    sim.close()
    assert np.all(parts["position"] == 0)  # only the mass field was written
    assert parts["mass"].sum() == pytest.approx(d["mass"], rel=1e-12)
