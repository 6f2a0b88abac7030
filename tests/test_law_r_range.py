"""Range of law R's FP32 pair arithmetic, on the CPU.

The direct sum evaluates law R (src/Sim/Physics.hpp:25-35) as
a' += d' rsq(r2 t^2 + tiny), t = fma(r2, h, k), in scaled units that the
library picks per system (nbody_direct_scale: a length lambda and a mass
scale c, powers of two).  This test takes the library's own choice of units
and checks, for systems that strain FP32 (a cluster 3e6 units from the origin
with a companion 2e7 away, masses over six and ten decades, no softening with
pairs 1e-5 apart, the reference's star system, coordinates scaled by 2^40),
that (1) every rsq argument of a pair that matters stays a normal FP32
number, and (2) a float32 model of the pair arithmetic summed over all
sources gives the FP64 direct sum's accelerations to 1e-5.  Before the
rescaling the argument r2 (r2 + S)^2 overflowed FP32 beyond |d| ~ 2.6e6 and
underflowed for close unsoftened pairs.

The model rounds each operation to float32 (an FMA as its float64 value
rounded once); it is a model of the kernel's arithmetic for range checks, the
device itself is checked in tests/test_gpu_range.py.
"""
import numpy as np
import pytest

FLT_MIN = np.float32(np.finfo(np.float32).tiny)
F32 = np.float32


def _fma(a, b, c):
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(F32)


def _model(pos, gm, S, sc):
    """Accelerations from the float32 model of the kernel's law-R pair loop."""
    il, c, soft, out = F32(sc["inv_len"]), F32(sc["mass_k"]), F32(sc["soft"]), sc["out"]
    p = (pos.astype(F32) * il).astype(F32)
    gm32 = gm.astype(F32)
    live = gm32 > 0
    h = np.where(live, (c / np.where(live, gm32, F32(1))).astype(F32), F32(np.inf)).astype(F32)
    k = np.where(live, (soft * h).astype(F32), F32(np.inf)).astype(F32)
    d = (p[None, :, :] - p[:, None, :]).astype(F32)  # [target, source, 3]
    with np.errstate(over="ignore", invalid="ignore", divide="ignore", under="ignore"):
        r2 = _fma(d[..., 2], d[..., 2], _fma(d[..., 1], d[..., 1], _fma(d[..., 0], d[..., 0], np.full(d.shape[:2], FLT_MIN, F32))))
        t = _fma(r2, np.broadcast_to(h, r2.shape), np.broadcast_to(k, r2.shape))
        u = (r2 * t).astype(F32)
        q = _fma(u, t, np.full(u.shape, FLT_MIN, F32))
        s = (F32(1) / np.sqrt(q)).astype(F32)
        a = (d.astype(np.float64) * s[..., None].astype(np.float64)).sum(1) * out
    return a, q


def _fp64(pos, gm, S):
    x = pos.astype(np.float64)
    d = x[None, :, :] - x[:, None, :]
    r2 = (d ** 2).sum(-1)
    with np.errstate(divide="ignore", invalid="ignore"):
        f = np.where(r2 > 0, gm[None, :] / (np.sqrt(r2) * (r2 + S)), 0.0)
    return (d * f[..., None]).sum(1)


def _extent(pos):
    return float((pos.max(0).astype(np.float64) - pos.min(0)).max())


def _check(pu, pos, gm, S, tol=1e-5):
    sc = pu.direct_scale(0, S, _extent(pos), float(gm.max()))
    a, q = _model(pos, gm, S, sc)
    ref = _fp64(pos, gm, S)
    err = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
    # every pair that matters: distinct points, a source within 1e-9 of the heaviest
    n = len(gm)
    distinct = np.any(pos[None, :, :] != pos[:, None, :], axis=-1)
    heavy = np.broadcast_to(gm[None, :] >= 1e-9 * gm.max(), (n, n))
    qm = q[distinct & heavy]
    return err, qm, sc


def _plummer(pu, n, seed):
    pos, _, mass = pu.ic_plummer(n, seed)
    return pos, mass


def test_far_cluster_with_companion_and_six_decades_of_mass(pu):
    pos, mass = _plummer(pu, 512, 0x5EED0203)
    pos = (pos + F32(3.0e6)).astype(F32)
    pos[-1] = pos[-1] + F32(2.0e7)
    gm = mass * np.where(np.arange(len(mass)) % 2 == 0, 1.0, 1e-6)
    err, qm, sc = _check(pu, pos, gm, 1e-4)
    assert np.all(qm >= FLT_MIN) and np.all(np.isfinite(qm))
    assert err.max() < 1e-5, err.max()


def test_unsoftened_close_pairs(pu):
    pos, mass = _plummer(pu, 512, 0x5EED0202)
    pos[1::64] = pos[0::64] + F32(1e-5)
    err, qm, _ = _check(pu, pos, mass, 0.0)
    assert np.all(qm >= FLT_MIN) and np.all(np.isfinite(qm))
    assert err.max() < 1e-5


@pytest.mark.parametrize("k", [40, -40])
def test_coordinates_scaled_by_powers_of_two(pu, k):
    pos, mass = _plummer(pu, 256, 0x5EED0201)
    f = F32(2.0 ** k)
    base = pu.direct_scale(0, 1e-4, _extent(pos), float(mass.max()))
    big = pu.direct_scale(0, 1e-4 * 4.0 ** k, _extent(pos * f), float(mass.max()))
    assert big["inv_len"] == base["inv_len"] * 2.0 ** -k       # the same scaled positions
    assert big["soft"] == base["soft"]
    assert big["out"] == base["out"] * 2.0 ** (-2 * k)
    err, qm, _ = _check(pu, (pos * f).astype(F32), mass, 1e-4 * 4.0 ** k)
    assert np.all(qm >= FLT_MIN) and np.all(np.isfinite(qm))
    assert err.max() < 1e-5


def test_reference_random_and_star_system_constants(pu):
    G = 6.674e-11
    # RandomSeeder-like: uniform in +-500, masses U(1e20, 1e30), S = 10
    pos, _, mass = pu.ic_uniform(512, 0x5EED0001, half=500.0, mass_lo=1e20, mass_hi=1e30)
    err, qm, _ = _check(pu, pos, G * mass, 10.0)
    assert np.all(qm >= FLT_MIN) and np.all(np.isfinite(qm))
    assert err.max() < 1e-5
    # StarSystemSeeder-like: a 1e30 star at the origin, bodies on the z axis at
    # 4-7 AU / 20 AU with masses U(1e10, 1e26) (StarSystemSeeder.cpp:17-55)
    rng = np.random.default_rng(5)
    n = 400
    pos = np.zeros((n, 3), F32)
    pos[1:, 2] = rng.uniform(4.0 / 20.0, 7.0 / 20.0, n - 1).astype(F32)
    mass = np.concatenate([[1e30], rng.uniform(1e10, 1e26, n - 1)])
    err, qm, _ = _check(pu, pos, G * mass, 10.0)
    assert np.all(qm >= FLT_MIN) and np.all(np.isfinite(qm))
    assert err.max() < 1e-5


def test_the_unscaled_form_was_out_of_range(pu):
    """The check has teeth: with lambda = c = 1 (the unscaled single-rsq form)
    the far system's arguments leave FP32's range."""
    pos, mass = _plummer(pu, 256, 0x5EED0203)
    pos = (pos + F32(3.0e6)).astype(F32)
    pos[-1] = pos[-1] + F32(2.0e7)
    unscaled = {"inv_len": 1.0, "mass_k": float(mass.max()), "soft": 1e-4, "out": float(mass.max())}
    _, q = _model(pos, mass, 1e-4, unscaled)
    distinct = np.any(pos[None, :, :] != pos[:, None, :], axis=-1)
    assert not np.all(np.isfinite(q[distinct]))
