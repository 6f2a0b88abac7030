"""The batch executor: many galaxies in one arena, one persistent kernel per
step, hipGraph-captured.  A system's result must be the same alone and in the
mix (bit for bit: the same per-target code and order), and match the
single-system path and the oracle within the parity tolerances."""
import numpy as np
import pytest

from helpers import rel_err_rows

pytestmark = pytest.mark.gpu

SOFT = 1e-4  # eps = 0.01
DT = 1e-3


def _cfg(pu):
    return pu.default_config(law=pu.LAW_R, softening=SOFT)


def _galaxies(pu, count=24, seed=7, lo=300, hi=5000):
    return pu.batch_galaxies(count, seed, lo, hi)


def test_alone_and_in_the_mix_bitwise(gpu):
    pu = gpu
    gal = _galaxies(pu)
    with pu.Batch(_cfg(pu), 200000) as b:
        ids = [b.add(p, v, m) for (p, v, m) in gal]
        b.step(DT, 3)
        mixed = [b.read(i) for i in ids]
    for k in (0, 5, 17, 23):
        p, v, m = gal[k]
        with pu.Batch(_cfg(pu), 20000) as b:
            i = b.add(p, v, m)
            b.step(DT, 3)
            alone = b.read(i)
        for x, y in zip(mixed[k], alone):
            assert np.array_equal(x, y), k


def test_batch_matches_single_system_and_oracle(gpu, oracle):
    pu = gpu
    gal = _galaxies(pu, count=10, seed=3)
    with pu.Batch(_cfg(pu), 100000) as b:
        ids = [b.add(p, v, m) for (p, v, m) in gal]
        b.step(DT, 10)
        res = [b.read(i) for i in ids]
    for k, (p, v, m) in enumerate(gal):
        with pu.NBodySim(_cfg(pu)) as sim:
            sim.init_soa(p, v, m)
            sim.step(DT, 10)
            ps, vs, _ = sim.read_soa()
        assert rel_err_rows(res[k][0], ps).max() < 1e-5
        if k < 3:
            p_ref, _ = oracle.ref_step(p, v.astype(np.float64), m, DT, 10, G=1.0, S=SOFT, scale=1.0)
            assert rel_err_rows(res[k][0], p_ref).max() < 1e-4


def test_graph_replays_match_eager(gpu):
    """Replay the captured graph several times, with eager work and syncs in
    between, and check every replay against plain launches of the same steps."""
    pu = gpu
    gal = _galaxies(pu, count=12, seed=11)
    with pu.Batch(_cfg(pu), 100000) as g, pu.Batch(_cfg(pu), 100000) as e:
        gi = [g.add(p, v, m) for (p, v, m) in gal]
        ei = [e.add(p, v, m) for (p, v, m) in gal]
        e.set_graph(False)
        for rep in range(4):
            g.step(DT, 2)
            e.step(DT, 2)
            g.synchronize()
            e.synchronize()
            for a, c in zip(gi, ei):
                ga, ea = g.read(a), e.read(c)
                for x, y in zip(ga, ea):
                    assert np.array_equal(x, y), rep


def test_join_and_leave_between_steps(gpu):
    pu = gpu
    gal = _galaxies(pu, count=8, seed=5)
    with pu.Batch(_cfg(pu), 60000) as b:
        ids = [b.add(*gal[k]) for k in range(4)]
        b.step(DT, 2)
        b.remove(ids[1])
        late = b.add(*gal[5])
        b.step(DT, 2)
        got_late = b.read(late)
        got0 = b.read(ids[0])
    with pu.Batch(_cfg(pu), 20000) as b:
        i = b.add(*gal[5])
        b.step(DT, 2)
        want_late = b.read(i)
    with pu.Batch(_cfg(pu), 20000) as b:
        i = b.add(*gal[0])
        b.step(DT, 4)
        want0 = b.read(i)
    for x, y in zip(got_late, want_late):
        assert np.array_equal(x, y)
    for x, y in zip(got0, want0):
        assert np.array_equal(x, y)


def test_capacity_and_errors(gpu):
    pu = gpu
    p, v, m = pu.ic_plummer(1000, 1)
    with pu.Batch(_cfg(pu), 2048) as b:
        a = b.add(p, v, m)
        b.add(p, v, m)
        with pytest.raises(pu.NBodyError) as e:
            b.add(p, v, m)
        assert e.value.status == 6  # NBODY_ERR_CAPACITY
        b.remove(a)
        b.add(p, v, m)
        inf = b.info()
        assert inf.systems == 2 and inf.particles == 2000
        with pytest.raises(pu.NBodyError):
            b.remove(99)


def test_sequential_baseline_agrees(gpu):
    pu = gpu
    gal = _galaxies(pu, count=6, seed=9)
    with pu.Batch(_cfg(pu), 50000) as b, pu.Batch(_cfg(pu), 50000) as s:
        bi = [b.add(*g) for g in gal]
        si = [s.add(*g) for g in gal]
        b.step(DT, 3)
        s.step_sequential(DT, 3)  # plain launches: the same per-target sums
        for x, y in zip(bi, si):
            for u, w in zip(b.read(x), s.read(y)):
                assert np.array_equal(u, w)
        s.step_sequential(DT, 2, split_small=True)
        b.step(DT, 2)
        for x, y in zip(bi, si):
            assert rel_err_rows(b.read(x)[0], s.read(y)[0]).max() < 1e-5


def test_tuner_keeps_state(gpu):
    pu = gpu
    gal = _galaxies(pu, count=10, seed=13)
    with pu.Batch(_cfg(pu), 80000) as b, pu.Batch(_cfg(pu), 80000) as c:
        bi = [b.add(*g) for g in gal]
        ci = [c.add(*g) for g in gal]
        t = b.tune(DT)
        assert t["i_per_lane"] in (1, 2) and t["block"] in (128, 256) and t["ms_per_step"] > 0
        c.set_tuning(t["i_per_lane"], t["block"], t["wg_per_cu"])
        b.step(DT, 2)
        c.step(DT, 2)
        for x, y in zip(bi, ci):
            for u, w in zip(b.read(x), c.read(y)):
                assert np.array_equal(u, w)


def test_plummer_law_batch(gpu, oracle):
    """Law P (a = G m r / (r^2 + eps^2)^(3/2)) through the batch: a galaxy is the
    same alone and in the mix bit for bit and matches the single-system path
    after five steps, whose first accelerations match the FP64 oracle."""
    pu = gpu
    cfg = pu.default_config(law=pu.LAW_P, softening=SOFT)
    gal = _galaxies(pu, count=8, seed=11)
    with pu.Batch(cfg, 60000) as b:
        ids = [b.add(p, v, m) for (p, v, m) in gal]
        b.step(DT, 5)
        mixed = [b.read(i) for i in ids]
    for k in (0, 3, 7):
        p, v, m = gal[k]
        with pu.Batch(cfg, 10000) as b:
            i = b.add(p, v, m)
            b.step(DT, 5)
            alone = b.read(i)
        for x, y in zip(mixed[k], alone):
            assert np.array_equal(x, y), k
        with pu.NBodySim(cfg) as sim:
            sim.init_soa(p, v, m)
            a = sim.accelerations()
            sim.step(DT, 5)
            ps, _, _ = sim.read_soa()
        assert rel_err_rows(mixed[k][0], ps).max() < 1e-5
        a64 = oracle.fp64_accel(pu.LAW_P, p, m, 1.0, SOFT)
        assert rel_err_rows(a, a64).max() < 1e-5
