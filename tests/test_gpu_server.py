"""Device initial conditions and the continuous-batching request server."""
import time

import numpy as np
import pytest

from helpers import rel_err_rows

pytestmark = pytest.mark.gpu


def test_device_plummer_matches_host_generator(gpu):
    pu = gpu
    n = 20000
    pos, vel, mass = pu.ic_plummer(n, 77)
    with pu.NBodySim(pu.default_config(law=pu.LAW_R, softening=1e-4)) as sim:
        sim.init_plummer(n, 77)
        p, v, _ = sim.read_soa()
        d = sim.diagnostics()
    # same counter-based stream; FP64 libm and summation order may move the last bits
    assert np.max(np.abs(p - pos)) < 1e-5 * np.max(np.abs(pos))
    assert np.max(np.abs(v - vel)) < 1e-5 * np.max(np.abs(vel))
    assert d["mass"] == pytest.approx(1.0, rel=1e-12)
    with pu.Batch(pu.default_config(law=pu.LAW_R, softening=1e-4), 40000) as b:
        i = b.add_plummer(n, 77)
        pb, vb, _ = b.read(i)
    assert np.array_equal(pb, p) and np.array_equal(vb, v)


def _serve(pu, requests, budget=1 << 34, keep=True, threshold=3000, inject=0, no_graph=0):
    cfg = pu.server_config(memory_budget=budget, direct_capacity=1 << 20, bh_threshold=threshold, keep_results=1,
                           dt=1e-3, inject_alloc_failures=inject, bh_no_graph=no_graph)
    out = {}
    with pu.Server(cfg) as srv:
        ids = [srv.submit_plummer(n, steps, seed) for (n, steps, seed) in requests]
        results = []
        peak = 0
        deadline = time.monotonic() + 120.0
        while time.monotonic() < deadline:
            srv.tick()
            results += srv.poll()
            st = srv.stats()
            peak = max(peak, st["memory_used"])
            if st["completed"] == len(requests):
                break
        assert len(results) == len(requests)
        for rid, (n, steps, seed) in zip(ids, requests):
            out[rid] = srv.fetch(rid, n)
        return ids, out, results, peak, srv.stats()


def test_server_result_same_alone_and_in_the_mix(gpu):
    pu = gpu
    rng = np.random.default_rng(5)
    reqs = [(int(n), int(s), 1000 + k) for k, (n, s) in
            enumerate(zip(rng.integers(500, 6000, 14), rng.integers(2, 9, 14)))]
    ids, mixed, results, _, st = _serve(pu, reqs)
    assert st["completed"] == len(reqs) and st["pending"] == 0
    methods = {r["id"]: r["method"] for r in results}
    assert {pu.METHOD_DIRECT, pu.METHOD_BARNES_HUT} <= set(methods.values())
    for k in (0, 3, 7, 11):
        aid, alone, _, _, _ = _serve(pu, [reqs[k]])
        for x, y in zip(mixed[ids[k]], alone[aid[0]]):
            assert np.array_equal(x, y), k


def test_server_admission_respects_budget(gpu):
    pu = gpu
    reqs = [(4000, 3, k) for k in range(12)] + [(20000, 2, 99)]
    # room for about four direct requests at a time (64 B per particle)
    budget = 4 * 4096 * 64 + 20480 * 432
    ids, out, results, peak, st = _serve(pu, reqs, budget=budget, threshold=10000)
    assert st["completed"] == len(reqs)
    assert peak <= budget
    order = sorted(results, key=lambda r: r["admitted"])
    # continuous batching: some requests were admitted after others completed
    first_done = min(r["completed"] for r in results)
    assert any(r["admitted"] > first_done for r in order)


def test_server_matches_single_simulation(gpu):
    pu = gpu
    reqs = [(2500, 4, 7), (5000, 3, 8)]
    ids, out, results, _, _ = _serve(pu, reqs, threshold=4000)
    # direct request == batch of one (bit for bit); BH request == its own handle
    with pu.Batch(pu.default_config(law=pu.LAW_R, softening=0.02 ** 2), 8192) as b:
        i = b.add_plummer(2500, 7)
        b.step(1e-3, 4)
        p, v, _ = b.read(i)
    assert np.array_equal(out[ids[0]][0], p)
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=0.02 ** 2, theta=0.6)
    with pu.NBodySim(cfg) as sim:
        sim.init_plummer(5000, 8)
        sim.step(1e-3, 3)
        p2, v2, _ = sim.read_soa()
    assert np.array_equal(out[ids[1]][0], p2)


def test_query_reports_completion(gpu):
    pu = gpu
    cfg = pu.default_config(method=pu.METHOD_BARNES_HUT, law=pu.LAW_R, softening=1e-4, theta=0.6)
    with pu.NBodySim(cfg) as sim:
        sim.init_plummer(50000, 3)
        sim.step(1e-3, 4)
        deadline = time.monotonic() + 60.0
        while not sim.query() and time.monotonic() < deadline:
            time.sleep(0.001)
        assert sim.query()
        sim.synchronize()
        assert sim.query()
    with pu.Batch(pu.default_config(law=pu.LAW_R, softening=1e-4), 8192) as b:
        b.add_plummer(3000, 4)
        b.step(1e-3, 2)
        b.synchronize()
        assert b.query()


@pytest.mark.parametrize("method", ["direct", "bh"])
def test_graph_replays_match_plain_launches(gpu, method):
    """nbody_set_graph: every replay (odd step counts flip the buffer parity,
    so both parities' graphs are replayed) equals plain launches bit for bit,
    with eager work and synchronisations between the replays."""
    pu = gpu
    m = pu.METHOD_BARNES_HUT if method == "bh" else pu.METHOD_DIRECT
    cfg = pu.default_config(method=m, law=pu.LAW_R, softening=1e-4, theta=0.6)
    with pu.NBodySim(cfg) as eager, pu.NBodySim(cfg) as graph:
        eager.init_plummer(20000, 11)
        graph.init_plummer(20000, 11)
        graph.set_graph(True)
        for rep in range(5):
            eager.step(1e-3, 3)
            graph.step(1e-3, 3)
            graph.synchronize()
            a_eager = eager.accelerations()  # eager work between the replays
            a_graph = graph.accelerations()
            pe, ve, _ = eager.read_soa()
            pg, vg, _ = graph.read_soa()
            assert np.array_equal(pe, pg) and np.array_equal(ve, vg), rep
            assert np.array_equal(a_eager, a_graph), rep
        if method == "bh":
            assert graph.bh_stats()["interactions"] > 0


def test_server_defers_requests_whose_allocation_fails(gpu):
    """Device allocations that fail (forced from the host: the server's next
    five request allocations report NBODY_ERR_OUT_OF_MEMORY) put the request
    back in the queue instead of failing the tick; every request still
    completes, with the same bits as without the failures."""
    pu = gpu
    reqs = [(2000, 3, 41), (6000, 2, 42), (1500, 4, 43), (8000, 2, 44), (2500, 3, 45), (5000, 2, 46)]
    ids0, out0, _, _, st0 = _serve(pu, reqs, threshold=4000)
    ids1, out1, res1, _, st1 = _serve(pu, reqs, threshold=4000, inject=5)
    assert st0["alloc_deferrals"] == 0
    assert st1["alloc_deferrals"] == 5
    assert st1["completed"] == len(reqs) and st1["pending"] == 0
    assert {r["method"] for r in res1} == {pu.METHOD_DIRECT, pu.METHOD_BARNES_HUT}
    for a, b in zip(ids0, ids1):
        for x, y in zip(out0[a], out1[b]):
            assert np.array_equal(x, y)


def test_server_plain_launches_equal_graph_replay(gpu):
    """bh_no_graph = 1 steps the Barnes-Hut requests with plain launches: every
    request's final state equals the graph-replayed server's bit for bit."""
    pu = gpu
    rng = np.random.default_rng(9)
    reqs = [(int(n), int(s), 2000 + k) for k, (n, s) in
            enumerate(zip(rng.integers(500, 9000, 10), rng.integers(2, 7, 10)))]
    ids_g, res_g, _, _, _ = _serve(pu, reqs)
    ids_p, res_p, results, _, _ = _serve(pu, reqs, no_graph=1)
    assert pu.METHOD_BARNES_HUT in {r["method"] for r in results}
    for a, b in zip(ids_g, ids_p):
        for x, y in zip(res_g[a], res_p[b]):
            assert np.array_equal(x, y)
