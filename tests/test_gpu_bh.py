"""Parity of the device Barnes-Hut with the CPU restatement of the reference's
Octree / BarnesHut (src/Sim/Octree.cpp:16-145, src/Sim/BarnesHut.cpp:44-110).

* Morton keys and the sorted order: bit-exact against the CPU.
* Accelerations: the reference's opening test (side / |p - COM| < theta) and
  monopoles on the same cells; per-particle relative error 1e-5 (the cell
  centres of mass are FP64 here and FP32 in the reference).
* Positions after 10 steps: relative error 1e-4.
"""
import numpy as np
import pytest

from helpers import REF_G, REF_S, opening_boundary_case, rel_err_rows

pytestmark = pytest.mark.gpu

REF_LO = [-4000.0] * 3
REF_SIZE = 8000.0


def _bh(pu, cfg, pos, vel, mass):
    sim = pu.NBodySim(cfg)
    sim.init_soa(pos, vel, mass)
    return sim


def _ref_input(pu, n, seed, half=500.0):
    return pu.ic_uniform(n, seed, half=half, mass_lo=1e20, mass_hi=1e30)


def test_keys_and_order_bit_exact_fixed_root(gpu, oracle):
    pu = gpu
    pos, vel, mass = _ref_input(pu, 20000, 0x5EED0004)
    pos[100:140] = pos[7]  # identical keys: the order among them must be stable
    pos[500] = np.float32(0.0)
    pos[501] = np.nextafter(np.float32(0), np.float32(-1))
    with _bh(pu, pu.reference_config(pu.METHOD_BARNES_HUT), pos, vel, mass) as sim:
        sim.accelerations()
        keys, order = sim.bh_tree_order()
        st = sim.bh_stats()
    ref_keys, outside = oracle.morton_keys(pos, REF_LO, REF_SIZE)
    ref_order = oracle.sort_order(ref_keys)
    assert outside == 0 and st["outside_root"] == 0
    assert np.array_equal(order, ref_order)
    assert np.array_equal(keys, ref_keys[ref_order])


def test_keys_and_order_bit_exact_dynamic_root(gpu, oracle):
    pu = gpu
    pos, vel, mass = pu.ic_plummer(50000, 0x5EED0005)
    pos += np.float32(3.25)
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=1e-4, theta=0.5)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        sim.accelerations()
        keys, order = sim.bh_tree_order()
    lo, size = oracle.dynamic_root(pos)
    ref_keys, outside = oracle.morton_keys(pos, lo, size)
    ref_order = oracle.sort_order(ref_keys)
    assert outside == 0
    assert np.array_equal(order, ref_order)
    assert np.array_equal(keys, ref_keys[ref_order])


@pytest.mark.parametrize("theta", [2.0, 0.5, 1.0])
def test_reference_octree_forces(gpu, oracle, theta):
    """Reference constants, fixed root [-4000, 4000)^3 (BarnesHut.cpp:14-19)."""
    pu = gpu
    n = 4096
    pos, vel, mass = _ref_input(pu, n, 0x5EED0006)
    cfg = pu.reference_config(pu.METHOD_BARNES_HUT, theta=theta)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a = sim.accelerations()
        st = sim.bh_stats()
    F, info = oracle.bh_forces(pos, mass, theta, REF_LO, REF_SIZE)
    ref = F / mass[:, None]
    err = rel_err_rows(a, ref)
    dcount = st["interactions"] - info["interactions"]
    print(f"theta {theta}: max rel {err.max():.2e}, p99.9 {np.quantile(err, 0.999):.2e}, "
          f"interactions/particle {st['interactions'] / n:.0f} (device - oracle {dcount}), "
          f"stack overflow {st['stack_overflow']}")
    assert st["stack_overflow"] == 0
    assert info["dropped"] == 0
    assert err.max() < 1e-5
    # The same walk: one interaction per accepted cell and per other particle
    # reached.  A dropped or repeated source of small mass would pass the force
    # bound; the count does not.  Decisions may flip only where |p - COM| is
    # within FP32 rounding of the threshold (FP64 vs FP32 cell centres).
    assert abs(dcount) <= 1e-5 * info["interactions"]


def test_dynamic_root_forces(gpu, oracle):
    """G = 1 Plummer sphere with the per-step power-of-two root; the CPU octree
    is built over the same root."""
    pu = gpu
    n, soft, theta = 6000, 1e-4, 0.5
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0007)
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=soft, theta=theta)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a = sim.accelerations()
        st = sim.bh_stats()
    lo, size = oracle.dynamic_root(pos)
    F, info = oracle.bh_forces(pos, mass, theta, lo, size, G=1.0, S=soft)
    err = rel_err_rows(a, F / mass[:, None])
    dcount = st["interactions"] - info["interactions"]
    print(f"dynamic root: max rel {err.max():.2e}, interactions device - oracle {dcount}")
    assert err.max() < 1e-5
    assert abs(dcount) <= 1e-5 * info["interactions"]


def test_theta_zero_is_direct_sum(gpu):
    pu = gpu
    n, soft = 3000, 1e-4
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0008)
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=soft, theta=0.0)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a_bh = sim.accelerations()
        st = sim.bh_stats()
    with _bh(pu, pu.default_config(law=pu.LAW_R, softening=soft), pos, vel, mass) as sim:
        a_d = sim.accelerations()
    assert st["interactions"] == n * (n - 1)
    assert rel_err_rows(a_bh, a_d).max() < 1e-5  # two FP32 sums of the same pairs in different orders


def test_theta_zero_full_octree_keeps_every_pair(gpu):
    """theta = 0 opens every cell for every lane.  Two particles in each of the
    64^3 level-7 cells of [0, 4000)^3 under the reference's fixed root make
    every cell of six levels have eight child cells, the deepest stack the walk
    can meet: batches of eight cells with no room kept for a depth-first
    descent would stack 8, 64, 120, 176, 218 entries and then overflow 224.
    Every pair is evaluated exactly once and the stack never overflows."""
    pu = gpu
    h = 62.5  # side of a level-7 cell of the 8000-wide root
    g = np.arange(64, dtype=np.float64) * h
    corner = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = np.concatenate([corner + 0.25 * h, corner + 0.75 * h]).astype(np.float32)
    n = len(pos)
    vel = np.zeros_like(pos)
    mass = np.full(n, 1e25)
    cfg = pu.reference_config(pu.METHOD_BARNES_HUT, theta=0.0)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a_bh = sim.accelerations()
        st = sim.bh_stats()
    with _bh(pu, pu.reference_config(pu.METHOD_DIRECT), pos, vel, mass) as sim:
        a_d = sim.accelerations()
    print(f"theta 0 lattice, N={n}: stack overflow {st['stack_overflow']}, interactions {st['interactions']}")
    assert st["stack_overflow"] == 0
    assert st["interactions"] == n * (n - 1)
    # the lattice's symmetry cancels most of each sum: compare against the RMS
    scale = np.sqrt((a_d.astype(np.float64) ** 2).sum(1).mean())
    assert np.abs(a_bh.astype(np.float64) - a_d).max() < 1e-5 * scale


def test_plummer_law_against_fp64(gpu, oracle):
    """Law P (G m r / (r^2 + eps^2)^(3/2)) through the octree: theta 0 is the
    FP64 direct sum to FP32 rounding, theta 0.5 converges towards it."""
    pu = gpu
    n, soft = 4000, 1e-4
    pos, vel, mass = pu.ic_plummer(n, 0x5EED000B)
    ref = oracle.fp64_accel(1, pos, mass, 1.0, soft)
    errs = {}
    for theta in (0.0, 0.5):
        cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_P, softening=soft, theta=theta)
        with _bh(pu, cfg, pos, vel, mass) as sim:
            errs[theta] = rel_err_rows(sim.accelerations(), ref)
    print(f"law P: theta 0 max {errs[0.0].max():.2e}, theta 0.5 rms {np.sqrt(np.mean(errs[0.5] ** 2)):.2e}")
    assert errs[0.0].max() < 1e-5
    assert np.sqrt(np.mean(errs[0.5] ** 2)) < 2e-2


def test_reference_ten_steps(gpu, oracle):
    pu = gpu
    n = 2048
    pos, _, mass = _ref_input(pu, n, 0x5EED0009)
    vel = (pos / np.linalg.norm(pos, axis=1, keepdims=True) * 1e16).astype(np.float32)
    dt = np.float32(0.02 / 60)
    with _bh(pu, pu.reference_config(pu.METHOD_BARNES_HUT), pos, vel, mass) as sim:
        sim.step(dt, 10)
        p, v, _ = sim.read_soa()
    p_ref, v_ref = oracle.bh_step(pos, vel.astype(np.float64), mass, dt, 10, 2.0, REF_LO, REF_SIZE)
    perr = rel_err_rows(p, p_ref)
    print(f"BH 10 steps: pos max rel {perr.max():.2e}")
    assert perr.max() < 1e-4


def test_identical_positions_form_a_bucket(gpu):
    """Coincident particles share one 63-bit key; their cell is evaluated pair by pair."""
    pu = gpu
    pos, vel, mass = pu.ic_plummer(1000, 0x5EED000A)
    pos[200:260] = pos[3]
    soft = 1e-4
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=soft, theta=0.0)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a_bh = sim.accelerations()
    with _bh(pu, pu.default_config(law=pu.LAW_R, softening=soft), pos, vel, mass) as sim:
        a_d = sim.accelerations()
    assert np.all(np.isfinite(a_bh))
    assert rel_err_rows(a_bh, a_d).max() < 1e-5  # two FP32 sums of the same pairs in different orders


def test_bucket_acts_as_one_particle_of_its_total_mass(gpu):
    """At theta 0.5 a bucket of 61 coincident particles has the centre of mass,
    mass, cells and opening decisions of one particle of their total mass at
    that point: every other particle feels the same force, and the coincident
    ones (same position, same decisions, zero contribution from each other)
    get the single particle's acceleration (equal to rounding: lanes in two
    wavefronts flush their lists at different points)."""
    pu = gpu
    n, soft = 1000, 1e-4
    pos, vel, mass = pu.ic_plummer(n, 0x5EED000B)
    pos[200:260] = pos[3]
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=soft, theta=0.5)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a_bucket = sim.accelerations()
    keep = np.r_[0:200, 260:n]
    mass_one = mass[keep].copy()
    mass_one[3] = mass[3] + mass[200:260].sum()
    with _bh(pu, cfg, pos[keep], vel[keep], mass_one) as sim:
        a_one = sim.accelerations()
    coincident = a_bucket[np.r_[3, 200:260]]
    assert np.all(np.isfinite(a_bucket))
    assert rel_err_rows(coincident, np.broadcast_to(coincident[0], coincident.shape)).max() < 1e-6
    err = rel_err_rows(a_bucket[keep], a_one)
    print(f"bucket vs one particle: max rel {err.max():.2e}")
    assert err.max() < 1e-5


def test_all_particles_at_one_point(gpu):
    """Every key equal: the root itself is a bucket (or, with a root side near
    2^-99, a monopole at the particles' own position), every pair has d = 0,
    and every acceleration is exactly zero."""
    pu = gpu
    n = 50
    pos = np.tile(np.array([[0.25, -0.5, 1.0]], np.float32), (n, 1))
    vel = np.zeros((n, 3), np.float32)
    mass = np.full(n, 1.0 / n)
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=1e-4, theta=0.5)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a = sim.accelerations()
        st = sim.bh_stats()
    assert np.all(a == 0.0)
    assert st["stack_overflow"] == 0


@pytest.mark.parametrize("n", [1, 2, 3, 17])
def test_tiny_sizes(gpu, oracle, n):
    pu = gpu
    pos, vel, mass = _ref_input(pu, n, 300 + n, half=100.0)
    with _bh(pu, pu.reference_config(pu.METHOD_BARNES_HUT), pos, vel, mass) as sim:
        a = sim.accelerations()
        sim.step(np.float32(1e-3), 2)
        p, _, _ = sim.read_soa()
    F, _ = oracle.bh_forces(pos, mass, 2.0, REF_LO, REF_SIZE)
    if n == 1:
        assert np.all(a == 0)
    else:
        assert rel_err_rows(a, F / mass[:, None]).max() < 1e-5
    p_ref, _ = oracle.bh_step(pos, vel.astype(np.float64), mass, np.float32(1e-3), 2, 2.0, REF_LO, REF_SIZE)
    assert rel_err_rows(p, p_ref).max() < 1e-4


def test_outside_fixed_root_is_counted(gpu):
    pu = gpu
    pos, vel, mass = _ref_input(pu, 100, 5)
    pos[0] = [5000.0, 0.0, 0.0]
    with _bh(pu, pu.reference_config(pu.METHOD_BARNES_HUT), pos, vel, mass) as sim:
        sim.accelerations()
        assert sim.bh_stats()["outside_root"] == 1


def test_two_disc_accuracy_vs_direct(gpu, oracle):
    """Two-disc galaxies, theta = 0.5: parity with the CPU octree on the same
    root, and the monopole approximation's own error against the device direct
    sum (about 1% RMS for the reference's opening test at theta = 0.5)."""
    pu = gpu
    n, soft = 65536, 0.0025
    p1, v1, m1 = pu.ic_disc_bulge(n // 2, 11, bulge_fraction=0.0)
    p2, v2, m2 = pu.ic_disc_bulge(n // 2, 12, bulge_fraction=0.0)
    pos = np.vstack([p1 - np.float32([20, 0, 0]), p2 + np.float32([20, 0, 0])]).astype(np.float32)
    vel = np.vstack([v1, v2]).astype(np.float32)
    mass = np.concatenate([m1, m2])
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=soft, theta=0.5)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a_bh = sim.accelerations()
        st = sim.bh_stats()
    with _bh(pu, pu.default_config(law=pu.LAW_R, softening=soft), pos, vel, mass) as sim:
        a_d = sim.accelerations()
    lo, size = oracle.dynamic_root(pos)
    # same cells and opening test with FP64 centres of mass: the traversal itself
    F64, info64 = oracle.bh_forces(pos, mass, 0.5, lo, size, G=1.0, S=soft, com_fp64=True)
    perr = rel_err_rows(a_bh, F64 / mass[:, None])
    # the reference's FP32 centre-of-mass accumulation differs by its own rounding
    F32, info32 = oracle.bh_forces(pos, mass, 0.5, lo, size, G=1.0, S=soft)
    rerr = rel_err_rows(a_bh, F32 / mass[:, None])
    err = rel_err_rows(a_bh, a_d)
    rms = np.sqrt(np.mean(err ** 2))
    print(f"two discs N={n}: vs CPU octree (FP64 COM) max {perr.max():.2e}; vs reference FP32 COM "
          f"rms {np.sqrt(np.mean(rerr ** 2)):.2e} max {rerr.max():.2e}; vs direct rms {rms:.2e} max {err.max():.2e}; "
          f"interactions/particle {st['interactions'] / n:.0f} (device - oracle {st['interactions'] - info64['interactions']} "
          f"with FP64 COM, {st['interactions'] - info32['interactions']} with FP32 COM)")
    assert perr.max() < 1e-5
    assert np.sqrt(np.mean(rerr ** 2)) < 1e-4
    assert rms < 3e-2
    assert st["interactions"] < n * n / 4
    # the same walk as the octree with the same centres of mass
    assert abs(st["interactions"] - info64["interactions"]) <= 1e-5 * info64["interactions"]


@pytest.mark.parametrize("k", [-2, -1, 0, 1, 2, 3])
def test_opening_decision_at_the_boundary(gpu, oracle, k):
    """The device's per-level opening threshold makes the reference's decision
    exactly where side / |p - COM| is 0.5 or one rounding step below it."""
    pu = gpu
    pos, mass, t = opening_boundary_case(k)
    vel = np.zeros_like(pos)
    cfg = pu.reference_config(pu.METHOD_BARNES_HUT, theta=0.5)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a = sim.accelerations()
        st = sim.bh_stats()
    F, info = oracle.bh_forces(pos, mass, 0.5, REF_LO, REF_SIZE)
    err = rel_err_rows(a, F / mass[:, None])
    assert err.max() < 1e-5, err
    assert st["interactions"] == info["interactions"] == (5 if k >= 1 else 6)


def test_cell_export_describes_the_octree(gpu):
    """nbody_bh_cells: every cell is a cube of its level on the root's grid that
    holds exactly its particles, with their FP64 mass and centre of mass, and
    (below level 21) splits them among two or more octants."""
    pu = gpu
    n = 20000
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0009)
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=1e-4, theta=0.5)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        sim.accelerations()
        keys, order = sim.bh_tree_order()
        cells = sim.bh_cells()
    assert cells.size > n // 8
    root = cells[0]
    assert root["level"] == 0 and root["first"] == 0 and root["count"] == n
    p = pos[order].astype(np.float64)
    m = mass[order]
    for c in cells[:: max(1, cells.size // 2000)]:
        f, k = int(c["first"]), int(c["count"])
        q = p[f:f + k]
        lo = c["lo"].astype(np.float64)
        assert np.all(q >= lo) and np.all(q < lo + float(c["side"]))
        assert float(c["side"]) == float(root["side"]) * 2.0 ** -int(c["level"])
        M = m[f:f + k].sum()
        assert c["mass"] == pytest.approx(M, rel=1e-12)
        com = (q * m[f:f + k, None]).sum(0) / M
        assert np.all(np.abs(c["com"] - com) <= 1e-6 * float(root["side"]))
        if c["level"] < 21:
            octant = (((q - lo) / (float(c["side"]) / 2)).astype(np.int64) * np.array([1, 2, 4])).sum(1)
            assert np.unique(octant).size >= 2
    # depth-first preorder: by first particle, enclosing cells first
    first = cells["first"].astype(np.int64)
    count = cells["count"].astype(np.int64)
    assert np.all(np.diff(first) >= 0)
    same = np.diff(first) == 0
    assert np.all(np.diff(count)[same] < 0)


@pytest.mark.parametrize("n", [8192, 8193, 131072, 131073, 262143, 262144, 524287, 524288,
                               1048575, 1048576, 1048577])
def test_sort_and_mass_paths_at_their_thresholds(gpu, oracle, n):
    """The sort's tiles hold 512 keys up to N = 131,072, 1,024 up to
    1,048,576 and 2,048 above; its scatter scans its own digit counts up to 16
    tiles (N <= 8,192) and uses a separate scan of the counts above.  The mass
    chunks hold 1,024 elements up to N = 262,143, 2,048 up to 524,287 and
    4,096 above, and the prefix pass folds the chunk scan up to 256 chunks
    (N + 1 <= 1,048,576).  On both sides of each threshold the keys and the
    stable order are bit-exact and the cells' FP64 masses and centres of mass
    are those of their particles."""
    pu = gpu
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0300 + (n & 0xff))
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=1e-4, theta=0.5)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a = sim.accelerations()
        keys, order = sim.bh_tree_order()
        cells = sim.bh_cells()
        st = sim.bh_stats()
    lo, size = oracle.dynamic_root(pos)
    ref_keys, outside = oracle.morton_keys(pos, lo, size)
    ref_order = oracle.sort_order(ref_keys)
    assert outside == 0 and st["stack_overflow"] == 0
    assert np.array_equal(order, ref_order)
    assert np.array_equal(keys, ref_keys[ref_order])
    assert np.all(np.isfinite(a))
    m = mass[order]
    csum = np.concatenate([[0.0], np.cumsum(m)])
    root = cells[0]
    assert root["count"] == n and root["mass"] == pytest.approx(mass.sum(), rel=1e-12)
    p = pos[order].astype(np.float64)
    for c in cells[:: max(1, cells.size // 500)]:
        f, k = int(c["first"]), int(c["count"])
        M = m[f:f + k].sum()
        assert c["mass"] == pytest.approx(M, rel=1e-12, abs=1e-15 * csum[-1])
        com = (p[f:f + k] * m[f:f + k, None]).sum(0) / M
        assert np.all(np.abs(c["com"] - com) <= 1e-6 * float(root["side"]))


def test_sort_duplicate_heavy_above_tile_threshold(gpu, oracle):
    """600,000 particles on 1,000 distinct points (about 600 identical keys
    each) sort in 293 tiles, through the separate scan of the tile counts:
    the keys and the stable order are bit-exact, and the buckets' coincident
    particles give finite accelerations."""
    pu = gpu
    rng = np.random.default_rng(0x5EED0400)
    pts = rng.uniform(-1.0, 1.0, size=(1000, 3)).astype(np.float32)
    pos = np.ascontiguousarray(pts[rng.integers(0, 1000, 600000)])
    vel = np.zeros_like(pos)
    mass = rng.uniform(1.0, 2.0, len(pos))
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=1e-4, theta=0.5)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        a = sim.accelerations()
        keys, order = sim.bh_tree_order()
        st = sim.bh_stats()
    lo, size = oracle.dynamic_root(pos)
    ref_keys, outside = oracle.morton_keys(pos, lo, size)
    ref_order = oracle.sort_order(ref_keys)
    assert outside == 0 and st["stack_overflow"] == 0
    assert np.array_equal(order, ref_order)
    assert np.array_equal(keys, ref_keys[ref_order])
    assert np.all(np.isfinite(a))


def test_full_size_encounter_keys_and_tree(gpu, oracle):
    """N = 4,194,304 (two 2,097,152-particle discs on a parabolic encounter,
    theta 0.5): keys and stable order bit-exact against the CPU on the full N,
    no stack overflow, the root cell holds every particle and the tree's FP64
    total mass equals the particles'."""
    pu = gpu
    n = 1 << 22
    pos, vel, mass = pu.ic_disc_encounter(n, 0x5EED0003, 3.0, 0.3, 40.0, 5.0, 30.0)
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=0.0025, theta=0.5)
    with _bh(pu, cfg, pos, vel, mass) as sim:
        sim.step(5e-4, 1)
        st = sim.bh_stats()
        keys, order = sim.bh_tree_order()  # the tree of the last step's state
        p1, _, _ = sim.read_soa()
        cells = sim.bh_cells()
    lo, size = oracle.dynamic_root(pos)
    ref_keys, outside = oracle.morton_keys(pos, lo, size)
    ref_order = oracle.sort_order(ref_keys)
    assert outside == 0
    assert np.array_equal(order, ref_order)
    assert np.array_equal(keys, ref_keys[ref_order])
    assert st["stack_overflow"] == 0
    ipp = st["interactions"] / n
    print(f"full size: {ipp:.0f} interactions/particle, {cells.size} cells")
    assert 300 < ipp < 5000
    assert cells[0]["count"] == n and cells[0]["mass"] == pytest.approx(mass.sum(), rel=1e-12)
    assert np.all(np.isfinite(p1))
