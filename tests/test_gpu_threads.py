"""Distinct simulation handles used from several host threads at once: each
handle runs on its own stream, so the results are those of the same runs one
after another, bit for bit (include/nbody.h: one handle must not be used from
two threads at once; distinct handles are independent)."""
import threading

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _run(pu, method, seed, n, steps):
    cfg = pu.default_config(method=method, law=pu.LAW_R, softening=1e-4, theta=0.5)
    pos, vel, mass = pu.ic_plummer(n, seed)
    with pu.NBodySim(cfg) as sim:
        sim.init_soa(pos, vel, mass)
        sim.step(1e-3, steps)
        return sim.read_soa()


def test_handles_in_parallel_threads_match_sequential(gpu):
    pu = gpu
    jobs = [(pu.METHOD_DIRECT, 11, 5000, 5), (pu.METHOD_BARNES_HUT, 12, 20000, 4),
            (pu.METHOD_DIRECT, 13, 3000, 7), (pu.METHOD_BARNES_HUT, 14, 8000, 6)] * 2
    seq = [_run(pu, *j) for j in jobs]
    out = [None] * len(jobs)
    errors = []
    start = threading.Barrier(len(jobs))

    def work(k):
        try:
            start.wait()
            for _ in range(3):  # create, run, read, destroy, again, while the others do the same
                out[k] = _run(pu, *jobs[k])
        except Exception as e:  # reported below, from the main thread
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=work, args=(k,)) for k in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    for k, (a, b) in enumerate(zip(seq, out)):
        for x, y in zip(a, b):
            assert np.array_equal(x, y), k
