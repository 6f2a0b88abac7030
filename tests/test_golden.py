"""The committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py): the oracle reproduces them bit for bit, and an
independent numpy restatement agrees with them, so a change of the oracle or a
corrupted fixture is caught on the CPU."""
import glob
import os

import numpy as np
import pytest

from helpers import rel_err_rows

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def law_r_numpy(pos, mass, G, S):
    """a_i = sum_j G m_j / (r^2 + S) * d / |d|, d = p_j - p_i (Physics.hpp:25-35,
    BruteForceCPU.cpp:36-40), in FP64 with the i == j term skipped."""
    p = pos.astype(np.float64)
    d = p[None, :, :] - p[:, None, :]
    r2 = (d * d).sum(-1)
    np.fill_diagonal(r2, 1.0)
    w = G * mass[None, :] / (r2 + S) / np.sqrt(r2)
    np.fill_diagonal(w, 0.0)
    return (w[:, :, None] * d).sum(1)


def test_fixtures_present():
    names = sorted(os.path.basename(f) for f in glob.glob(os.path.join(GOLDEN, "*.npz")))
    assert names == ["bh_theta05_n257.npz", "g1_law_r_n257.npz", "g1_law_r_n64.npz", "ref_law_r_n257.npz",
                     "ref_law_r_n64.npz", "ref_law_r_n8.npz"]


@pytest.mark.parametrize("name", ["ref_law_r_n8.npz", "ref_law_r_n64.npz", "ref_law_r_n257.npz",
                                  "g1_law_r_n64.npz", "g1_law_r_n257.npz"])
def test_direct_fixture(oracle, name):
    g = load(name)
    G, S, scale, dt = float(g["G"]), float(g["S"]), float(g["scale"]), float(g["dt"])
    acc = oracle.ref_accel(g["pos"], g["mass"], G=G, S=S, workers=3)
    assert np.array_equal(acc, g["acc"])
    p10, v10 = oracle.ref_step(g["pos"], g["vel"], g["mass"], dt, int(g["nsteps"]), G=G, S=S, scale=scale, workers=2)
    assert np.array_equal(p10, g["pos10"]) and np.array_equal(v10, g["vel10"])
    assert rel_err_rows(g["acc"], law_r_numpy(g["pos"], g["mass"], G, S)).max() < 1e-6
    if "acc_fp64" in g:
        assert rel_err_rows(g["acc"], g["acc_fp64"]).max() < 1e-6


def test_bh_fixture(oracle):
    g = load("bh_theta05_n257.npz")
    lo, size, theta = g["root_lo"], float(g["root_size"]), float(g["theta"])
    keys, outside = oracle.morton_keys(g["pos"], lo, size)
    assert outside == 0 and np.array_equal(keys, g["keys"])
    assert np.array_equal(g["order"], np.argsort(g["keys"], kind="stable").astype(np.uint32))
    F, _ = oracle.bh_forces(g["pos"], g["mass"], theta, lo, size, workers=3)
    F64, _ = oracle.bh_forces(g["pos"], g["mass"], theta, lo, size, workers=3, com_fp64=True)
    assert np.array_equal(F, g["forces"]) and np.array_equal(F64, g["forces_com64"])
    p10, v10 = oracle.bh_step(g["pos"], g["vel"], g["mass"], float(g["dt"]), int(g["nsteps"]), theta, lo, size,
                              workers=2)
    assert np.array_equal(p10, g["pos10"]) and np.array_equal(v10, g["vel10"])
    # theta 0.5 on 257 scattered particles: monopole error of a few percent at most
    exact = law_r_numpy(g["pos"], g["mass"], float(g["G"]), float(g["S"])) * g["mass"][:, None]
    err = rel_err_rows(g["forces"], exact)
    assert np.sqrt(np.mean(err ** 2)) < 2e-2
