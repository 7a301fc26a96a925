"""The blue-noise oracle (oracle/blue_noise.py) against the reference's
generator as it evaluates (crates/ps-clouds/src/noise.rs:646-704) and the
host generator behind the ABI; the GPU entry points are exported. CPU only."""
import ctypes as C
import ctypes.util

import numpy as np
import pytest

import oracle.blue_noise as BN
from test_oracle_pins import check_blue_noise
from pedalsky_amd import abi, scene


def test_splat_weights_are_the_c_library_expf():
    # Rust's f32::exp is the C library's expf: the oracle's correctly rounded
    # exponentials of the f32 products must be those values exactly
    libm = C.CDLL(ctypes.util.find_library("m"))
    libm.expf.restype = C.c_float
    libm.expf.argtypes = [C.c_float]
    inv = np.float32(1.0) / (np.float32(2.0) * np.float32(1.9) * np.float32(1.9))
    w = BN.splat_weights()
    for r2 in range(73):
        x = np.float32(-np.float32(r2)) * inv
        assert np.float32(libm.expf(float(x))) == w[r2], r2
    assert w[0] == 1.0 and np.all(np.diff(w) < 0)


@pytest.fixture(scope="module")
def oracle_map():
    return BN.generate()


def test_oracle_satisfies_the_pins(oracle_map):
    check_blue_noise(oracle_map)


def test_host_generator_is_the_oracle(oracle_map):
    assert np.array_equal(scene.blue_noise(), oracle_map)


@pytest.mark.parametrize("slip", ["last index among equals", "splat radius 5"])
def test_restatement_slips_change_the_map(oracle_map, slip, monkeypatch):
    # two plausible misreadings of the greedy placement; each gives another map
    if slip == "splat radius 5":
        monkeypatch.setattr(BN, "RADIUS", 5)
        assert not np.array_equal(BN.generate(), oracle_map)
    else:
        real_argmin = np.argmin
        monkeypatch.setattr(BN.np, "argmin", lambda a: len(a) - 1 - int(real_argmin(a[::-1])))
        bad = BN.generate()
        monkeypatch.setattr(BN.np, "argmin", real_argmin)
        assert not np.array_equal(bad, oracle_map)
        with pytest.raises(AssertionError):
            check_blue_noise(bad)


def test_entry_points_exported():
    lib = abi.load()
    for name in ("ps_synthesise_blue_noise", "ps_read_blue_noise"):
        assert hasattr(lib, name)
