"""The C-ABI library (no GPU needed): it loads, exports every declared entry
point, validates arguments before touching a device, and its host-side
initial-condition generators behave."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from helpers import ROOT

HEADER = os.path.join(ROOT, "include", "nbody.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nbody_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ("nbody_create", "nbody_init", "nbody_step", "nbody_read", "nbody_destroy"):
        assert required in names


def test_library_exports_every_declared_symbol(pu):
    lib = pu.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", pu.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\bT (nbody_\w+)", out))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, f"declared but not exported: {missing}"
    for n in declared_functions():
        getattr(lib, n)


def test_particle_record_layout(pu):
    dt = pu.PARTICLE_DTYPE
    assert dt.itemsize == 104
    assert dt.fields["position"][1] == 0
    assert dt.fields["colour"][1] == 12
    assert dt.fields["original_colour"][1] == 28
    assert dt.fields["velocity"][1] == 48
    assert dt.fields["forces"][1] == 72
    assert dt.fields["mass"][1] == 96


def test_names_and_versions(pu):
    lib = pu.load_library()
    assert lib.nbody_abi_version() == 2
    assert lib.nbody_method_name(pu.METHOD_DIRECT) == b"Brute Force GPU"
    assert lib.nbody_method_name(pu.METHOD_BARNES_HUT) == b"Barnes-Hut"
    assert lib.nbody_method_name(99) == b"Unknown"
    assert lib.nbody_status_string(0) == b"ok"
    assert pu.nbody_sim_get_name(pu.ENBodySim.BruteForceCPU) == "Brute Force CPU"
    assert pu.nbody_sim_get_name(pu.ENBodySim.BruteForceGPU) == "Brute Force GPU"
    assert pu.nbody_sim_get_name(pu.ENBodySim.BarnesHut) == "Barnes-Hut"
    assert pu.nbody_sim_get_name(7) == "Unknown"


def test_reference_config_constants(pu):
    c = pu.reference_config(pu.METHOD_BARNES_HUT)
    assert c.G == 6.674e-11 and c.softening == 10.0 and c.scale == 2.3e13
    assert c.theta == 2.0 and c.law == pu.LAW_R
    assert c.root_mode == pu.ROOT_FIXED and list(c.root_lo) == [-4000.0] * 3 and c.root_size == 8000.0
    d = pu.default_config()
    assert d.G == 1.0 and d.law == pu.LAW_P and d.method == pu.METHOD_DIRECT


@pytest.mark.parametrize(
    "overrides",
    [
        {"method": 7},
        {"law": 3},
        {"law": 1, "softening": 0.0},
        {"softening": -1.0},
        {"scale": 0.0},
        {"G": float("nan")},
        {"theta": -0.5},
        {"root_mode": 1, "root_size": 0.0},
        {"tune_i_per_lane": 3},
        {"tune_block": 96},
    ],
)
def test_invalid_config_rejected_before_device(pu, overrides):
    with pytest.raises(pu.NBodyError) as e:
        pu.NBodySim(pu.default_config(**overrides))
    assert e.value.status == 1  # NBODY_ERR_INVALID_ARG


def test_create_without_device_fails_cleanly(pu):
    if pu.device_count() > 0:
        pytest.skip("a device is visible")
    with pytest.raises(pu.NBodyError) as e:
        pu.NBodySim(pu.default_config())
    assert e.value.status == 3  # NBODY_ERR_DEVICE
    assert "device" in str(e.value).lower()


def test_null_arguments(pu):
    lib = pu.load_library()
    assert lib.nbody_create(None, None) == 1
    assert lib.nbody_step(None, ctypes.c_float(0.1), 1) == 1
    assert lib.nbody_read(None, None, 0, 1) == 1
    assert lib.nbody_destroy(None) == 0
    assert len(lib.nbody_last_error()) >= 0


def test_ic_counter_based_prefix_property(pu):
    """Particle i's draws depend on (seed, i) only."""
    a = pu.ic_uniform(1000, 42)
    b = pu.ic_uniform(300, 42)
    for x, y in zip(a, b):
        assert np.array_equal(x[:300], y)
    c = pu.ic_uniform(300, 43)
    assert not np.array_equal(a[0][:300], c[0])


def test_ic_plummer_properties(pu, oracle):
    pos, vel, mass = pu.ic_plummer(4096, 0x5EED0001)
    assert np.allclose(mass, 1 / 4096)
    com = (pos * mass[:, None]).sum(0)
    assert np.all(np.abs(com) < 1e-6) and np.all(np.abs((vel * mass[:, None]).sum(0)) < 1e-6)
    r = np.linalg.norm(pos, axis=1)
    assert r.max() <= 20.0 * 1.001
    # half-mass radius of a Plummer sphere: a / sqrt(2^(2/3) - 1) = 1.3048
    assert np.median(r) == pytest.approx(1.3048, rel=0.05)
    e = oracle.fp64_energy(1, pos, vel, mass, 1.0, 1e-8)
    assert 2 * e["kinetic"] / abs(e["potential"]) == pytest.approx(1.0, rel=0.06)


def test_ic_disc_bulge_properties(pu):
    n = 20000
    pos, vel, mass = pu.ic_disc_bulge(n, 7, Rd=3.0, z0=0.3, bulge_fraction=0.2, ab=0.6)
    nb = round(0.2 * n)
    nd = n - nb
    assert mass.sum() == pytest.approx(1.0, rel=1e-12)
    assert mass[:nd].sum() == pytest.approx(0.8, rel=1e-12)
    R = np.hypot(pos[:nd, 0], pos[:nd, 1])
    assert np.mean(R) == pytest.approx(2 * 3.0, rel=0.05)  # Gamma(2, Rd) mean
    assert np.std(pos[:nd, 2]) == pytest.approx(0.3 * np.pi / np.sqrt(12), rel=0.1)  # sech^2 std
    rb = np.linalg.norm(pos[nd:], axis=1)
    assert np.median(rb) == pytest.approx(0.6 * (1 + np.sqrt(2)), rel=0.08)  # Hernquist half-mass
    # disc rotates: angular momentum along +z dominates
    Lz = (pos[:nd, 0] * vel[:nd, 1] - pos[:nd, 1] * vel[:nd, 0]).sum()
    assert Lz > 0


def test_ic_invalid_arguments(pu):
    with pytest.raises(pu.NBodyError):
        pu.ic_plummer(10, 1, a=-1.0)
    with pytest.raises(pu.NBodyError):
        pu.ic_disc_bulge(10, 1, bulge_fraction=1.5)


def test_last_error_is_per_thread(pu):
    """nbody_last_error() belongs to the calling thread: a failure in one
    thread leaves another thread's message untouched (no GPU: the config is
    rejected before any device call)."""
    import threading

    lib = pu.load_library()
    bar = threading.Barrier(2)
    seen = {}

    def failing():
        c = pu.default_config()
        c.law = 7
        h = pu._P()
        seen["status"] = lib.nbody_create(ctypes.byref(c), ctypes.byref(h))
        seen["a"] = lib.nbody_last_error()
        bar.wait()
        bar.wait()
        seen["a_after"] = lib.nbody_last_error()

    def other():
        c = pu.default_config()
        c.theta = -1.0
        h = pu._P()
        bar.wait()
        seen["b_before"] = lib.nbody_last_error()
        seen["b_status"] = lib.nbody_create(ctypes.byref(c), ctypes.byref(h))
        seen["b"] = lib.nbody_last_error()
        bar.wait()

    threads = [threading.Thread(target=failing), threading.Thread(target=other)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(30)
    assert seen["status"] == 1 and b"law" in seen["a"]  # NBODY_ERR_INVALID_ARG
    assert not seen["b_before"]  # the other thread's failure is not visible here
    assert seen["b_status"] == 1 and b"theta" in seen["b"]
    assert seen["a_after"] == seen["a"]  # and this thread's failure is not visible there
