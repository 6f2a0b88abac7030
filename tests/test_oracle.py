"""Pins of the CPU oracle (oracle/oracle.cpp) against what the mathematics fixes.

The reference's own tests hold nothing for the gravity path (test/MathsTests.cpp
and test/LogTests.cpp cover ClosestParticle and log formatting only), so every
check here is a closed form, an invariant, a special case or an independent
brute-force computation, chosen so that a dropped term, a wrong sign, a wrong
index or a transposed operand in the oracle fails one of them.
"""
from fractions import Fraction

import numpy as np
import pytest

from helpers import REF_G, REF_S, REF_SCALE, law_r_force, opening_boundary_case, rel_err_rows


def _two_body(r=3.0, m0=2.0e24, m1=5.0e25):
    pos = np.array([[0.5, -1.0, 2.0], [0.5 + r, -1.0, 2.0]], np.float32)
    mass = np.array([m0, m1])
    return pos, mass


# ---- Law R (reference) direct sum -------------------------------------------

def test_ref_two_body_closed_form(oracle):
    """F_0 = G m0 m1 / (r^2 + S) toward particle 1 (Physics.hpp:25-35, BruteForceCPU.cpp:36-40)."""
    pos, mass = _two_body()
    F = oracle.ref_forces(pos, mass)
    r = float(pos[1, 0] - pos[0, 0])
    mag = REF_G * mass[0] * mass[1] / (r * r + REF_S)
    assert F[0] == pytest.approx([mag, 0.0, 0.0], rel=1e-12, abs=1e-30 * mag)
    assert F[1] == pytest.approx([-mag, 0.0, 0.0], rel=1e-12, abs=1e-30 * mag)


def test_ref_softening_enters_additively(oracle):
    """Law R is not Plummer: at r^2 = S the force is half of G m m / r^2."""
    S = 4.0
    pos = np.array([[0, 0, 0], [2, 0, 0]], np.float32)
    mass = np.array([1.0, 1.0])
    F = oracle.ref_forces(pos, mass, G=1.0, S=S)
    assert F[0, 0] == pytest.approx(1.0 / 8.0, rel=1e-14)  # 1 / (4 + 4)


def test_ref_single_and_coincident(oracle):
    pos = np.array([[1.0, 2.0, 3.0]], np.float32)
    assert np.all(oracle.ref_forces(pos, np.array([1e25])) == 0)
    pos = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0]], np.float32)
    F = oracle.ref_forces(pos, np.array([1e25, 1e25]))
    assert np.all(F == 0)  # Normalize of a zero vector is zero


def test_ref_newton_third_law(oracle, pu):
    pos, vel, mass = pu.ic_uniform(300, 11, half=500.0, mass_lo=1e20, mass_hi=1e30)
    F = oracle.ref_forces(pos, mass)
    total = np.abs(F).sum(0)
    assert np.all(np.abs(F.sum(0)) <= 1e-6 * total)


def test_ref_octahedron_symmetry(oracle):
    """Equal masses on a regular octahedron around a centre particle."""
    v = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float32) * 8.0
    pos = np.vstack([[[0, 0, 0]], v]).astype(np.float32)
    mass = np.full(7, 3.0e24)
    F = oracle.ref_forces(pos, mass)
    assert np.all(np.abs(F[0]) < 1e-12 * np.abs(F[1]).max())
    mags = np.linalg.norm(F[1:], axis=1)
    assert np.allclose(mags, mags[0], rtol=1e-12)
    # every vertex is pulled toward the centre
    cos = (F[1:] * -v).sum(1) / (mags * np.linalg.norm(v, axis=1))
    assert np.allclose(cos, 1.0, atol=1e-12)


def test_ref_rows_subset_matches_full(oracle, pu):
    pos, vel, mass = pu.ic_uniform(200, 5)
    full = oracle.ref_forces(pos, mass)
    rows = np.array([0, 17, 199, 5])
    part = oracle.ref_forces(pos, mass, rows=rows, workers=3)
    assert np.array_equal(part, full[rows])


def test_ref_worker_count_independent(oracle, pu):
    """Each target's sum runs over j in index order regardless of the partition."""
    pos, vel, mass = pu.ic_uniform(257, 9)
    a = oracle.ref_forces(pos, mass, workers=1)
    b = oracle.ref_forces(pos, mass, workers=7)
    assert np.array_equal(a, b)


def test_ref_vs_numpy_restatement(oracle, pu):
    """An independent numpy statement of the same FP32-geometry / FP64-magnitude pair law."""
    pos, vel, mass = pu.ic_uniform(96, 21, half=300.0)
    F = oracle.ref_forces(pos, mass)
    p = pos.astype(np.float32)
    ref = np.zeros((96, 3))
    for i in range(96):
        d = (p[i] - p).astype(np.float32)  # b.Position - a.Position, FP32
        ln = np.sqrt((d * d).sum(1, dtype=np.float32)).astype(np.float32)
        n = np.where(ln[:, None] > 0, d / np.where(ln > 0, ln, 1)[:, None], 0).astype(np.float32)
        dd = ((p - p[i]) ** 2).sum(1, dtype=np.float32)  # DistanceSquared, FP32
        f = -(REF_G * mass * mass[i]) / (dd.astype(np.float64) + REF_S)
        f[i] = 0.0
        ref[i] = (f[:, None] * n.astype(np.float64)).sum(0)
    assert np.max(rel_err_rows(F, ref)) < 1e-12


def test_ref_close_to_fp64_law_r(oracle, pu):
    """The FP32 Normalize / r^2 of the reference differ from FP64 by FP32 rounding only."""
    pos, vel, mass = pu.ic_uniform(400, 3)
    a_ref = oracle.ref_accel(pos, mass)
    a64 = oracle.fp64_accel(0, pos, mass, REF_G, REF_S)
    assert np.max(rel_err_rows(a_ref, a64)) < 2e-6


# ---- integrator --------------------------------------------------------------

def test_ref_step_closed_form(oracle):
    """One kick-then-drift step (BruteForceCPU.cpp:61-73): v1 = v0 + F/m dt, x1 = x0 + (float)(v1 dt / scale)."""
    pos, mass = _two_body()
    vel = np.array([[1e10, -2e10, 0.0], [0.0, 3e9, -1e9]])
    dt = np.float32(0.02 / 60)
    F = oracle.ref_forces(pos, mass)
    v1 = vel + F / mass[:, None] * np.float64(dt)
    x1 = pos + (v1 * np.float64(dt) / REF_SCALE).astype(np.float32)
    p, v = oracle.ref_step(pos, vel, mass, dt, 1)
    assert np.array_equal(v, v1)
    assert np.array_equal(p, x1.astype(np.float32))


def test_circular_orbit_energy(oracle):
    """Kick-drift on a two-body circular orbit (G = 1, S = 0): bounded energy error over 5 periods."""
    m = np.array([1.0, 1e-6])
    r = 1.0
    vc = np.sqrt(1.0 * (m[0] + m[1]) / r)
    pos = np.array([[0, 0, 0], [r, 0, 0]], np.float32)
    vel = np.array([[0, -vc * m[1] / (m[0] + m[1]), 0], [0, vc * m[0] / (m[0] + m[1]), 0]])
    period = 2 * np.pi * r / vc
    steps = 2000
    dt = period / 400
    e0 = oracle.fp64_energy(0, pos, vel.astype(np.float32), m, 1.0, 0.0)["total"]
    p, v = oracle.ref_step(pos, vel, m, dt, steps, G=1.0, S=0.0, scale=1.0)
    e1 = oracle.fp64_energy(0, p, v.astype(np.float32), m, 1.0, 0.0)["total"]
    rad = np.linalg.norm(p[1] - p[0])
    assert abs(rad - r) < 0.05
    assert abs(e1 - e0) < 0.05 * abs(e0)


# ---- FP64 oracle ----------------------------------------------------------------

@pytest.mark.parametrize("law", [0, 1])
def test_fp64_two_body(oracle, law):
    pos = np.array([[0, 0, 0], [0, 0, 1.5]], np.float32)
    mass = np.array([2.0, 3.0])
    soft = 0.25
    a = oracle.fp64_accel(law, pos, mass, 1.0, soft)
    r = 1.5
    if law == 0:
        mag = 3.0 / (r * r + soft)
    else:
        mag = 3.0 * r / (r * r + soft) ** 1.5
    assert a[0] == pytest.approx([0, 0, mag], rel=1e-15, abs=1e-300)
    assert a[1] == pytest.approx([0, 0, -2.0 / 3.0 * mag], rel=1e-15, abs=1e-300)  # a_1 = -G m_0 g(r)


@pytest.mark.parametrize("law", [0, 1])
def test_fp64_vs_numpy(oracle, pu, law):
    pos, vel, mass = pu.ic_plummer(128, 4)
    soft = 1e-3
    a = oracle.fp64_accel(law, pos, mass, 1.0, soft)
    p = pos.astype(np.float64)
    for i in (0, 63, 127):
        d = p - p[i]
        r2 = (d * d).sum(1)
        if law == 0:
            with np.errstate(divide="ignore", invalid="ignore"):
                w = np.where(r2 > 0, mass / (np.sqrt(r2) * (r2 + soft)), 0.0)
        else:
            w = mass / (r2 + soft) ** 1.5
        w[i] = 0
        ref = (d * w[:, None]).sum(0)
        assert np.linalg.norm(a[i] - ref) <= 1e-13 * np.linalg.norm(ref)


def test_energy_potential_is_force_integral(oracle):
    """-dU/dr of the law-R pair potential equals the pair force (central difference)."""
    S = 0.3
    m = np.array([2.0, 5.0])
    def U(r):
        pos = np.array([[0, 0, 0], [r, 0, 0]], np.float32)
        return oracle.fp64_energy(0, pos, np.zeros((2, 3), np.float32), m, 1.0, S)["potential"]
    r, h = 1.25, 1.0 / 64
    dU = (U(r + h) - U(r - h)) / (2 * h)
    force = 2.0 * 5.0 / (r * r + S)
    assert dU == pytest.approx(force, rel=5e-4)


def test_energy_momenta_closed_form(oracle):
    """Two bodies worked by hand: P = sum m v, L = sum m x cross v, K = sum m v^2 /
    (2 scale).  Every component differs, so a swapped index, a sign or a
    factor (v x r instead of r x v, a missing mass) changes the answer."""
    pos = np.array([[1, 0, 0], [0, 2, 1]], np.float32)
    vel = np.array([[0, 3, 0], [4, 0, -2]], np.float32)
    m = np.array([2.0, 0.5])
    e = oracle.fp64_energy(0, pos, vel, m, 1.0, 0.1, scale=2.0)
    assert np.array_equal(e["momentum"], [2.0, 6.0, -1.0])          # 2 (0,3,0) + 0.5 (4,0,-2)
    assert np.array_equal(e["angular_momentum"], [-2.0, 2.0, 2.0])  # 2 (0,0,3) + 0.5 (-4,4,-8)
    assert e["kinetic"] == pytest.approx(14.0 / 2.0, rel=1e-15)     # (2 * 9 + 0.5 * 20) / 2 / scale
    assert e["mass"] == 2.5
    # the law-R pair potential at r = sqrt(6): -(G m1 m2 / sqrt S)(pi/2 - atan(r / sqrt S))
    r, S = np.sqrt(6.0), 0.1
    assert e["potential"] == pytest.approx(-(1.0 / np.sqrt(S)) * (np.pi / 2 - np.arctan(r / np.sqrt(S))), rel=1e-12)


def test_plummer_virial_energy(oracle, pu):
    """Plummer sphere (G = M = a = 1): E = -3 pi / 64 = -0.14726 in the large-N limit."""
    pos, vel, mass = pu.ic_plummer(4096, 0x5EED0001)
    e = oracle.fp64_energy(1, pos, vel, mass, 1.0, 1e-6)
    assert e["total"] == pytest.approx(-3 * np.pi / 64, rel=0.05)
    assert 2 * e["kinetic"] / abs(e["potential"]) == pytest.approx(1.0, rel=0.06)
    assert np.all(np.abs(e["momentum"]) < 1e-6)


# ---- Morton keys ------------------------------------------------------------------

def _interleave(ix, iy, iz):
    k = 0
    for b in range(21):
        k |= ((ix >> b) & 1) << (3 * b)
        k |= ((iy >> b) & 1) << (3 * b + 1)
        k |= ((iz >> b) & 1) << (3 * b + 2)
    return k


def test_morton_known_answers(oracle):
    """Child order x + 2y + 4z of Octree::Split (Octree.cpp:25-41): x is the lowest bit."""
    pos = np.array(
        [[1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [2, 0, 0], [2 ** 21 - 1] * 3, [0, 0, 0]], np.float32
    )
    keys, outside = oracle.morton_keys(pos, [0, 0, 0], 2.0 ** 21)
    assert outside == 0
    assert list(keys) == [1, 2, 4, 7, 8, 2 ** 63 - 1, 0]


def test_morton_exact_vs_rational(oracle):
    """Keys equal an exact rational floor((x - lo) / cell) on random points, for the
    reference's non-power-of-two root [-4000, 4000)^3 (BarnesHut.cpp:14-19)."""
    rng = np.random.default_rng(7)
    pos = rng.uniform(-4000, 4000, size=(500, 3)).astype(np.float32)
    # points on and within an ulp of cell boundaries
    pos[:20] = np.float32(0.0)
    pos[20:40] = np.nextafter(np.float32(0), np.float32(-1))
    pos[40:60] = np.float32(-4000.0)
    pos[60:80] = np.float32(125.0 / 64)  # a level-12 boundary
    pos[80:100] = np.nextafter(np.float32(125.0 / 64), np.float32(0))
    keys, outside = oracle.morton_keys(pos, [-4000, -4000, -4000], 8000.0)
    assert outside == 0
    cell = Fraction(8000, 2 ** 21)
    for i in range(len(pos)):
        idx = [int((Fraction(float(x)) + 4000) // cell) for x in pos[i]]
        assert int(keys[i]) == _interleave(*idx), i


def test_morton_clamps_outside(oracle):
    pos = np.array([[-5000, 0, 0], [4000, 0, 0], [0, 0, 3999.999]], np.float32)
    keys, outside = oracle.morton_keys(pos, [-4000, -4000, -4000], 8000.0)
    assert outside == 2


def test_morton_cells_are_reference_octree_cells(oracle, pu):
    """A particle's leaf depth in the reference's pointer octree (Octree::Add/Split)
    is one more than the longest 3-bit-group prefix its key shares with any other key."""
    pos, vel, mass = pu.ic_uniform(600, 13, half=500.0)
    keys, _ = oracle.morton_keys(pos, [-4000] * 3, 8000.0)
    depth = oracle.bh_leaf_depths(pos, mass, [-4000] * 3, 8000.0)
    k = [int(x) for x in keys]
    for i in range(len(k)):
        best = 0
        for j in range(len(k)):
            if i == j:
                continue
            x = k[i] ^ k[j]
            common = 21 if x == 0 else (63 - x.bit_length()) // 3
            best = max(best, common)
        assert depth[i] == best + 1, i


def test_sort_order_is_stable(oracle):
    keys = np.array([5, 3, 5, 1, 3, 5, 0], np.uint64)
    assert list(oracle.sort_order(keys)) == [6, 3, 1, 4, 0, 2, 5]


def test_dynamic_root_covers_and_aligns(oracle, pu):
    for seed, half in ((1, 3.0), (2, 1000.0), (3, 1e-3)):
        pos, vel, mass = pu.ic_uniform(1000, seed, half=half)
        pos = pos + np.float32(half * 0.37)
        lo, size = oracle.dynamic_root(pos)
        assert size > 0 and np.log2(size) == int(np.log2(size))
        cell = size / 2 ** 21
        assert np.all(np.asarray(lo, np.float64) / cell == np.floor(np.asarray(lo, np.float64) / cell))
        assert np.all(pos >= lo) and np.all(pos < lo.astype(np.float64) + size)
        keys, outside = oracle.morton_keys(pos, lo, size)
        assert outside == 0


# ---- Barnes-Hut restatement ---------------------------------------------------------

def test_bh_theta_zero_is_direct_sum(oracle, pu):
    """theta = 0 opens every internal node (d / r < 0 never holds), so
    CalculateForce reduces to the exact pair sum of BruteForceCPU."""
    pos, vel, mass = pu.ic_uniform(300, 17, half=500.0)
    F_bh, st = oracle.bh_forces(pos, mass, 0.0, [-4000] * 3, 8000.0)
    F = oracle.ref_forces(pos, mass)
    assert st["dropped"] == 0 and st["leaves"] == 300
    assert st["interactions"] == 300 * 299
    assert np.max(rel_err_rows(F_bh, F)) < 1e-12


def test_bh_large_theta_is_root_monopole(oracle, pu):
    """A huge theta approximates at the root: F_p = Gravity(p, COM, M) along (COM - p),
    the root monopole including p itself (no self-exclusion, Octree.cpp:122-132)."""
    pos, vel, mass = pu.ic_uniform(50, 23, half=200.0)
    F_bh, st = oracle.bh_forces(pos, mass, 1e30, [-4000] * 3, 8000.0)
    assert st["interactions"] == 50  # one root monopole per target
    M = mass.sum()
    com = (pos.astype(np.float64) * mass[:, None]).sum(0) / M
    for i in range(50):
        d = com - pos[i].astype(np.float64)
        r = np.linalg.norm(d)
        ref = REF_G * mass[i] * M / (r * r + REF_S) * d / r
        assert np.linalg.norm(F_bh[i] - ref) <= 1e-4 * np.linalg.norm(ref)


def test_bh_converges_to_direct(oracle, pu):
    pos, vel, mass = pu.ic_uniform(800, 29, half=500.0, mass_lo=1e25, mass_hi=1e25)
    F = oracle.ref_forces(pos, mass)
    errs = []
    for theta in (1.0, 0.5, 0.25):
        F_bh, _ = oracle.bh_forces(pos, mass, theta, [-4000] * 3, 8000.0)
        errs.append(np.sqrt(np.mean(rel_err_rows(F_bh, F) ** 2)))
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 1e-2


def test_bh_drops_particles_outside_root(oracle):
    pos = np.array([[0, 0, 0], [1, 1, 1], [5000, 0, 0]], np.float32)
    mass = np.array([1e25, 1e25, 1e25])
    _, st = oracle.bh_forces(pos, mass, 0.5, [-4000] * 3, 8000.0)
    assert st["dropped"] == 1


@pytest.mark.parametrize("k", [-2, -1, 0, 1, 2, 3])
def test_bh_opening_test_at_the_boundary(oracle, k):
    """Octree.cpp:122-124 approximates iff side / r < theta.  A two-particle
    cell of side 125 seen from exactly 250 (quotient 0.5, theta 0.5) is opened;
    one FP32 ulp farther it is approximated by its monopole.  The pair's exact
    forces and the monopole differ by about 5%, so the decision is visible."""
    pos, mass, t = opening_boundary_case(k)
    F, st = oracle.bh_forces(pos, mass, 0.5, [-4000] * 3, 8000.0)
    assert st["dropped"] == 0
    # particles 0 and 1 each meet the other and the target's leaf; the target
    # meets the pair's two leaves, or their cell's monopole once
    assert st["interactions"] == (5 if k >= 1 else 6)
    exact = law_r_force(pos[t], pos[0], mass[t], mass[0]) + law_r_force(pos[t], pos[1], mass[t], mass[1])
    mono = law_r_force(pos[t], [62.5, 62.5, 62.5], mass[t], mass[0] + mass[1])
    assert np.linalg.norm(exact - mono) > 1e-2 * np.linalg.norm(exact)
    want = mono if k >= 1 else exact
    assert np.linalg.norm(F[t] - want) < 1e-5 * np.linalg.norm(want)


@pytest.mark.parametrize("c", [(-9.28220844, 3.83515167, -2.42638111), (1.0, 2.0, 3.0)])
def test_bh_coincident_pair_is_dropped(oracle, c):
    """Octree::Split (Octree.cpp:16-51) cannot separate two particles at one
    position.  At the first coordinate the cube containing them stops
    shrinking once it is one FP32 ulp wide (lo + size/2 rounds up), so the
    reference's Add recursion would never end; at the second a child cube
    comes out empty and the reference drops both.  The oracle drops both in
    either case: the pair feels the third particle as the direct sum gives
    it, and the third particle feels nothing (the tree holds no mass of the
    pair)."""
    pos = np.array([c, c, (7.5, -6.25, 8.0)], np.float32)
    m = np.array([2.0, 3.0, 5.0])
    for theta in (0.0, 0.5, 2.0):
        F, st = oracle.bh_forces(pos, m, theta, [-20.0, -20.0, -20.0], 40.0, G=1.0, S=0.01)
        assert st["dropped"] == 2 and st["leaves"] == 1
        assert np.all(F[2] == 0)
        for i in (0, 1):
            pair = oracle.ref_forces(pos[[i, 2]], m[[i, 2]], G=1.0, S=0.01)
            np.testing.assert_allclose(F[i], pair[0], rtol=1e-12)
