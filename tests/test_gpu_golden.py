"""The device against the committed golden fixtures (tests/golden/*.npz, the
oracle's outputs for seeded inputs; tests/test_golden.py pins them on the CPU):
accelerations to 1e-5 relative per particle, positions after 10 steps to 1e-4,
Barnes-Hut keys and order bit-exact."""
import os

import numpy as np
import pytest

from helpers import rel_err_rows

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _run(pu, cfg, g):
    vel = g["vel"].astype(np.float32)
    with pu.NBodySim(cfg) as sim:
        sim.init_soa(g["pos"], vel, g["mass"])
        acc = sim.accelerations()
        extra = {}
        if cfg.method == pu.METHOD_BARNES_HUT:
            extra["keys"], extra["order"] = sim.bh_tree_order()
        sim.step(float(g["dt"]), int(g["nsteps"]))
        p, v, _ = sim.read_soa()
    return acc, p, v, extra


@pytest.mark.parametrize("name", ["ref_law_r_n8.npz", "ref_law_r_n64.npz", "ref_law_r_n257.npz",
                                  "g1_law_r_n64.npz", "g1_law_r_n257.npz"])
def test_direct_matches_golden(gpu, name):
    pu = gpu
    g = load(name)
    cfg = pu.default_config(method=pu.METHOD_DIRECT, law=pu.LAW_R, G=float(g["G"]), softening=float(g["S"]),
                            scale=float(g["scale"]))
    acc, p, v, _ = _run(pu, cfg, g)
    aerr = rel_err_rows(acc, g["acc"])
    perr = rel_err_rows(p, g["pos10"])
    verr = rel_err_rows(v, g["vel10"])
    print(f"{name}: acc {aerr.max():.2e}, pos10 {perr.max():.2e}, vel10 {verr.max():.2e}")
    assert aerr.max() < 1e-5
    assert perr.max() < 1e-4
    assert verr.max() < 1e-4


def test_barnes_hut_matches_golden(gpu):
    pu = gpu
    g = load("bh_theta05_n257.npz")
    cfg = pu.reference_config(pu.METHOD_BARNES_HUT, theta=float(g["theta"]))
    acc, p, v, extra = _run(pu, cfg, g)
    assert np.array_equal(extra["order"], g["order"])
    assert np.array_equal(extra["keys"], g["keys"][g["order"]])
    aerr = rel_err_rows(acc, g["forces_com64"] / g["mass"][:, None])
    perr = rel_err_rows(p, g["pos10"])
    print(f"bh golden: acc {aerr.max():.2e}, pos10 {perr.max():.2e}")
    assert aerr.max() < 1e-5
    assert perr.max() < 1e-4
