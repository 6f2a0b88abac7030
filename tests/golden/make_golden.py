"""Write the golden fixtures of this directory from the CPU oracle.

    make -C oracle && python tests/golden/make_golden.py

Every fixture is a small .npz (numpy arrays only, no pickles): seeded inputs
drawn here with numpy's PCG64, and the oracle's outputs for them:

* ref_law_r_n{8,64,257}.npz - the reference's constants (src/Sim/Physics.hpp:9-16):
  accelerations of BruteForceCPU::Exec (src/Sim/BruteForceCPU.cpp:25-43) and
  positions / velocities after 10 steps of BruteForceCPU::Update (:45-73).
* g1_law_r_n{64,257}.npz - G = 1, S = eps^2 = 1e-4, scale 1, dt 1e-3 (the
  softening and step of the N = 4096 Plummer configuration at small N):
  accelerations and 10 steps as above, plus the FP64 direct-sum oracle.
* bh_theta05_n257.npz - Octree / BarnesHut (src/Sim/Octree.cpp:16-145) on the
  reference's fixed root [-4000, 4000)^3 with theta 0.5: Morton keys, the
  stable sort order, forces with the reference's FP32 centres of mass and with
  FP64 ones, and positions after 10 Barnes-Hut steps.

The oracle is deterministic (every row sums its sources in index order), so
regenerating reproduces the files bit for bit.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from helpers import REF_G, REF_S, REF_SCALE, Oracle  # noqa: E402

REF_LO = np.array([-4000.0] * 3, np.float32)
REF_SIZE = 8000.0


def ref_inputs(n, seed):
    """Uniform cube of half-width 500 with log-uniform masses 1e20..1e30 and
    small velocities, as the reference's precompute seeder produces."""
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-500.0, 500.0, (n, 3)).astype(np.float32)
    vel = rng.normal(0.0, 1e3, (n, 3))
    mass = 10.0 ** rng.uniform(20.0, 30.0, n)
    return pos, vel, mass


def plummer_inputs(n, seed):
    """Equal-mass Plummer sphere (a = 1, M = 1, G = 1): radius from the inverted
    mass profile, isotropic directions, speeds by rejection on q^2 (1 - q^2)^3.5."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(1e-6, 1.0, n)
    r = np.minimum(1.0 / np.sqrt(u ** (-2.0 / 3.0) - 1.0), 20.0)

    def directions(k):
        c = rng.uniform(-1.0, 1.0, k)
        p = rng.uniform(0.0, 2.0 * np.pi, k)
        s = np.sqrt(1.0 - c * c)
        return np.stack([s * np.cos(p), s * np.sin(p), c], 1)

    q = np.empty(n)
    for i in range(n):
        while True:
            x, y = rng.uniform(0.0, 1.0), rng.uniform(0.0, 0.1)
            if y < x * x * (1.0 - x * x) ** 3.5:
                q[i] = x
                break
    ve = np.sqrt(2.0) * (1.0 + r * r) ** -0.25
    pos = (r[:, None] * directions(n)).astype(np.float32)
    vel = (q * ve)[:, None] * directions(n)
    mass = np.full(n, 1.0 / n)
    return pos, vel, mass


def main():
    o = Oracle()
    w = 1  # one worker: the oracle's result does not depend on it, but keep it fixed
    for n, seed in ((8, 0x601D0008), (64, 0x601D0040), (257, 0x601D0101)):
        pos, vel, mass = ref_inputs(n, seed)
        dt = 1.0 / 60.0
        acc = o.ref_accel(pos, mass, workers=w)
        p10, v10 = o.ref_step(pos, vel, mass, dt, 10, workers=w)
        np.savez_compressed(
            os.path.join(HERE, f"ref_law_r_n{n}.npz"), pos=pos, vel=vel, mass=mass, G=REF_G, S=REF_S,
            scale=REF_SCALE, dt=dt, nsteps=10, acc=acc, pos10=p10, vel10=v10,
        )
    for n, seed in ((64, 0x601D1040), (257, 0x601D1101)):
        pos, vel, mass = plummer_inputs(n, seed)
        G, S, dt = 1.0, 1e-4, 1e-3
        acc = o.ref_accel(pos, mass, G=G, S=S, workers=w)
        acc64 = o.fp64_accel(0, pos, mass, G, S, workers=w)
        p10, v10 = o.ref_step(pos, vel, mass, dt, 10, G=G, S=S, scale=1.0, workers=w)
        np.savez_compressed(
            os.path.join(HERE, f"g1_law_r_n{n}.npz"), pos=pos, vel=vel, mass=mass, G=G, S=S, scale=1.0, dt=dt,
            nsteps=10, acc=acc, acc_fp64=acc64, pos10=p10, vel10=v10,
        )
    n = 257
    pos, vel, mass = ref_inputs(n, 0x601D2101)
    keys, outside = o.morton_keys(pos, REF_LO, REF_SIZE)
    assert outside == 0
    order = o.sort_order(keys)
    F, _ = o.bh_forces(pos, mass, 0.5, REF_LO, REF_SIZE, workers=w)
    F64, _ = o.bh_forces(pos, mass, 0.5, REF_LO, REF_SIZE, workers=w, com_fp64=True)
    dt = 1.0 / 60.0
    p10, v10 = o.bh_step(pos, vel, mass, dt, 10, 0.5, REF_LO, REF_SIZE, workers=w)
    np.savez_compressed(
        os.path.join(HERE, f"bh_theta05_n{n}.npz"), pos=pos, vel=vel, mass=mass, G=REF_G, S=REF_S, scale=REF_SCALE,
        theta=0.5, root_lo=REF_LO, root_size=REF_SIZE, keys=keys, order=order, forces=F, forces_com64=F64, dt=dt,
        nsteps=10, pos10=p10, vel10=v10,
    )
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)), "bytes")


if __name__ == "__main__":
    main()
