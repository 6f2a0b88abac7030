"""Range of law R's FP32 pair arithmetic on the device.

The direct sum runs in per-system power-of-two units (a length from the
system's extent, a mass scale from its largest G m), so positions scaled by
2^k with the softening scaled by 2^2k give accelerations scaled by exactly
2^-2k, bit for bit, for any k that keeps the inputs normal FP32 numbers: the
old single-rsq form overflowed FP32 at |d| ~ 2.6e6 and lost softening-free
close pairs to underflow.  Checked at 2^22 (|d| up to ~2e8) and 2^-20, with
S = 0, and against the FP64 oracle.
"""
import numpy as np
import pytest

from helpers import rel_err_rows

pytestmark = pytest.mark.gpu


def _acc(pu, pos, vel, mass, soft, law=0):
    with pu.NBodySim(pu.default_config(law=law, softening=soft)) as sim:
        sim.init_soa(pos, vel, mass)
        return sim.accelerations()


@pytest.mark.parametrize("k", [22, -20, 40])
def test_scaling_by_a_power_of_two_is_exact(gpu, oracle, k):
    pu = gpu
    n, soft = 4096, 1e-4
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0201)
    a1 = _acc(pu, pos, vel, mass, soft)
    f = np.float32(2.0 ** k)
    pos_k = (pos * f).astype(np.float32)
    ak = _acc(pu, pos_k, vel, mass, soft * 4.0 ** k)
    assert np.array_equal(ak * np.float32(2.0 ** (2 * k)), a1)
    a64 = oracle.fp64_accel(0, pos_k, mass, 1.0, soft * 4.0 ** k)
    err = rel_err_rows(ak, a64)
    print(f"positions x 2^{k}: max rel {err.max():.2e} against FP64")
    assert err.max() < 1e-5


def test_unsoftened_law_r_with_close_pairs(gpu, oracle):
    """S = 0: a = G m_j rhat / r^2, with pairs 1e-5 apart."""
    pu = gpu
    n = 4096
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0202)
    pos[1::64] = pos[0::64] + np.float32(1e-5)  # 64 close pairs
    a = _acc(pu, pos, vel, mass, 0.0)
    a64 = oracle.fp64_accel(0, pos, mass, 1.0, 0.0)
    err = rel_err_rows(a, a64)
    print(f"S = 0: max rel {err.max():.2e}, rms {np.sqrt(np.mean(err ** 2)):.2e}")
    assert np.all(np.isfinite(a))
    assert err.max() < 1e-5


def test_far_system_and_wide_mass_range(gpu, oracle):
    """A compact cluster 3e6 units from the origin next to a far companion, and
    masses spanning 1e-6 to 1 (the light sources' pulls are 1e6 times weaker)."""
    pu = gpu
    n = 2048
    pos, vel, mass = pu.ic_plummer(n, 0x5EED0203)
    pos = (pos + np.float32(3.0e6)).astype(np.float32)
    pos[-1] = pos[-1] + np.float32(2.0e7)  # a companion 2e7 away
    mass = mass * np.where(np.arange(n) % 2 == 0, 1.0, 1e-6)
    a = _acc(pu, pos, vel, mass, 1e-4)
    a64 = oracle.fp64_accel(0, pos, mass, 1.0, 1e-4)
    err = rel_err_rows(a, a64)
    print(f"far system, masses over 6 decades: max rel {err.max():.2e}")
    # positions near 3e6 resolve only ~0.25 units: the FP32 differences (the
    # oracle's inputs too) are exact, the sums are FP32
    assert err.max() < 1e-5


def test_negative_or_non_finite_masses_are_rejected(gpu):
    pu = gpu
    pos = np.zeros((4, 3), np.float32)
    pos[:, 0] = np.arange(4)
    for bad in (-1.0, float("nan"), float("inf")):
        mass = np.ones(4)
        mass[2] = bad
        with pu.NBodySim(pu.default_config()) as sim:
            with pytest.raises(pu.NBodyError) as e:
                sim.init_soa(pos, None, mass)
            assert e.value.status == 1
        with pu.Batch(pu.default_config(law=pu.LAW_R), 1024) as b:
            with pytest.raises(pu.NBodyError) as e:
                b.add(pos, None, mass)
            assert e.value.status == 1
            assert b.info().free_records == 1024  # nothing was allocated
