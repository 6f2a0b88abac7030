"""nbody_field_noalloc: the direct-sum field at arbitrary points without any
device allocation (its staging lives in a fixed device array).  Called
through the C ABI, it must equal nbody_field bit for bit, on two calls in a
row, the second with fewer points, and refuse more points than its array
holds."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_POINTS = 65536  # NBODY_FIELD_NOALLOC_MAX_POINTS


def _noalloc(pu, sim, pts):
    lib = pu.load_library()
    pts = np.ascontiguousarray(pts, dtype=np.float32)
    out = np.full_like(pts, np.nan)
    fp = ctypes.POINTER(ctypes.c_float)
    st = lib.nbody_field_noalloc(sim._h, pts.ctypes.data_as(fp), pts.shape[0], out.ctypes.data_as(fp))
    return st, out


@pytest.mark.parametrize("law", [0, 1])
def test_field_noalloc_equals_field(gpu, law):
    pu = gpu
    n = 20000
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0301)
    rng = np.random.default_rng(7)
    with pu.NBodySim(pu.default_config(law=law, softening=1e-4)) as sim:
        sim.init_soa(pos, vel, mass)
        for m in (MAX_POINTS, 3001):  # a full array, then a smaller call after it
            pts = np.concatenate([pos[rng.integers(0, n, m // 2)],
                                  rng.normal(0.0, 2.0, (m - m // 2, 3))]).astype(np.float32)
            want = sim.field(pts)
            st, got = _noalloc(pu, sim, pts)
            assert st == 0, pu.load_library().nbody_last_error()
            assert np.array_equal(got, want), m
        st, _ = _noalloc(pu, sim, np.zeros((MAX_POINTS + 1, 3), np.float32))
        assert st == 6  # NBODY_ERR_CAPACITY
        st, got = _noalloc(pu, sim, np.zeros((0, 3), np.float32))
        assert st == 0
