"""Opening angle changed on a live simulation (the reference's BHThetaChanged,
src/Sim/BarnesHut.cpp:29-31), and captured step graphs replaced while earlier
launches may still be queued.

* nbody_set_theta: after each change (2.0 -> 0.5 -> 1.0) the accelerations
  equal the oracle octree's at the new theta (max relative 1e-5) with the
  same interaction count, and three steps match the restated
  BarnesHut::Update (positions 1e-4); a handle replaying captured graphs
  gives the same bits as one launching plainly.
* Graph replacement: a sequence of nbody_step / nbody_batch_step calls whose
  (steps, dt) change between calls without any synchronisation, so a new
  graph is captured while the previous one may be queued, equals plain
  launches bit for bit.
"""
import numpy as np
import pytest

from helpers import bh_step_dynamic, rel_err_rows

pytestmark = pytest.mark.gpu

REF_LO = [-4000.0] * 3
REF_SIZE = 8000.0


def _check_theta(oracle, sim, pos_now, mass, theta, **kw):
    a = sim.accelerations()
    st = sim.bh_stats()
    if "root" in kw:
        lo, size = kw.pop("root")
    else:
        lo, size = REF_LO, REF_SIZE
    F, info = oracle.bh_forces(pos_now, mass, theta, lo, size, com_fp64=True, **kw)
    err = rel_err_rows(a, F / mass[:, None])
    dcount = st["interactions"] - info["interactions"]
    print(f"theta {theta}: max rel {err.max():.2e}, interactions {st['interactions']} (device - oracle {dcount})")
    assert err.max() < 1e-5
    assert abs(dcount) <= 1e-5 * info["interactions"]
    return st["interactions"]


def test_set_theta_on_a_live_simulation_fixed_root(gpu, oracle):
    pu = gpu
    n = 4096
    pos, _, mass = pu.ic_uniform(n, 0x5EED0101, half=500.0, mass_lo=1e20, mass_hi=1e30)
    vel = (pos / np.linalg.norm(pos, axis=1, keepdims=True) * 1e16).astype(np.float32)
    dt = np.float32(0.02 / 60)
    cfg = pu.reference_config(pu.METHOD_BARNES_HUT)  # theta 2.0, fixed root [-4000, 4000)^3
    counts = {}
    with pu.NBodySim(cfg) as sim, pu.NBodySim(cfg) as g:
        sim.init_soa(pos, vel, mass)
        g.init_soa(pos, vel, mass)
        g.set_graph(True)
        for theta in (2.0, 0.5, 1.0):
            if theta != 2.0:
                sim.set_theta(theta)
                g.set_theta(theta)
            p0, v0, _ = sim.read_soa()
            counts[theta] = _check_theta(oracle, sim, p0, mass, theta)
            sim.step(dt, 3)
            g.step(dt, 3)
            p1, v1, _ = sim.read_soa()
            pg, vg, _ = g.read_soa()
            assert np.array_equal(p1, pg) and np.array_equal(v1, vg), theta
            p_ref, _ = oracle.bh_step(p0, v0.astype(np.float64), mass, dt, 3, theta, REF_LO, REF_SIZE)
            perr = rel_err_rows(p1, p_ref)
            print(f"theta {theta}: 3 steps, position max rel {perr.max():.2e}")
            assert perr.max() < 1e-4
    # a smaller opening angle opens more cells
    assert counts[0.5] > counts[1.0] > counts[2.0]


def test_set_theta_on_a_live_simulation_dynamic_root(gpu, oracle):
    pu = gpu
    n, soft, dt = 6000, 1e-4, 1e-3
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0102)
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=soft, theta=2.0)
    with pu.NBodySim(cfg) as sim, pu.NBodySim(cfg) as g:
        sim.init_soa(pos, vel, mass)
        g.init_soa(pos, vel, mass)
        g.set_graph(True)
        for theta in (2.0, 0.5, 1.0):
            if theta != 2.0:
                g.step(dt, 2)  # queued with the old theta, not waited for
                sim.step(dt, 2)
                sim.set_theta(theta)
                g.set_theta(theta)
            p0, v0, _ = sim.read_soa()
            _check_theta(oracle, sim, p0, mass, theta, G=1.0, S=soft, root=oracle.dynamic_root(p0))
            sim.step(dt, 3)
            g.step(dt, 3)
            p1, v1, _ = sim.read_soa()
            pg, vg, _ = g.read_soa()
            assert np.array_equal(p1, pg) and np.array_equal(v1, vg), theta
            p_ref, _ = bh_step_dynamic(oracle, p0, v0, mass, dt, 3, theta, G=1.0, S=soft)
            assert rel_err_rows(p1, p_ref).max() < 1e-4
    with pytest.raises(pu.NBodyError) as e:
        pu.NBodySim(cfg).set_theta(float("nan"))
    assert e.value.status == 1


@pytest.mark.parametrize("method", ["direct", "bh"])
def test_graph_replaced_between_unsynchronised_steps(gpu, method):
    pu = gpu
    m = pu.METHOD_BARNES_HUT if method == "bh" else pu.METHOD_DIRECT
    cfg = pu.default_config(method=m, law=pu.LAW_R, softening=1e-4, theta=0.6)
    calls = [(1e-3, 1), (1e-3, 1), (5e-4, 2), (1e-3, 3), (5e-4, 2), (2e-3, 1), (1e-3, 1)]
    with pu.NBodySim(cfg) as eager, pu.NBodySim(cfg) as graph:
        eager.init_plummer(30000, 21)
        graph.init_plummer(30000, 21)
        graph.set_graph(True)
        for dt, k in calls:  # no synchronisation between the calls
            eager.step(dt, k)
            graph.step(dt, k)
        pe, ve, ae = eager.read_soa()
        pg, vg, ag = graph.read_soa()
    assert np.array_equal(pe, pg) and np.array_equal(ve, vg) and np.array_equal(ae, ag)


def test_batch_graph_replaced_between_unsynchronised_steps(gpu):
    pu = gpu
    cfg = pu.default_config(law=pu.LAW_R, softening=1e-4)
    gal = pu.batch_galaxies(10, 31, 1000, 6000)
    calls = [(1e-3, 1), (1e-3, 1), (5e-4, 2), (1e-3, 3), (2e-3, 1)]
    with pu.Batch(cfg, 100000) as g, pu.Batch(cfg, 100000) as e:
        gi = [g.add(*x) for x in gal[:6]]
        ei = [e.add(*x) for x in gal[:6]]
        e.set_graph(False)
        for dt, k in calls:
            g.step(dt, k)
            e.step(dt, k)
        # systems join between unsynchronised steps: the item list grows and
        # is reallocated while the previous graph may still be queued
        gi += [g.add(*x) for x in gal[6:]]
        ei += [e.add(*x) for x in gal[6:]]
        for dt, k in calls[::-1]:
            g.step(dt, k)
            e.step(dt, k)
        for a, c in zip(gi, ei):
            for x, y in zip(g.read(a), e.read(c)):
                assert np.array_equal(x, y)
