"""numpy's default generator as the device task draws from it (pcg64.h,
compiled for the host by tests/hostcheck), against numpy itself: the
SeedSequence seeding of PCG64 for one- and many-word seeds, and 64-bit outputs
at arbitrary offsets taken by jump-ahead against numpy's own advance() and
random_raw().  The 32-bit buffering and the uniform transform are checked
against Generator.integers / Generator.uniform on the same outputs."""

import ctypes

import numpy as np
import pytest

from helpers import HOSTCHECK

pytestmark = pytest.mark.skipif(not HOSTCHECK.exists(), reason="tests/hostcheck not built (run build())")

SEEDS = [0, 1, 7, 12345, 2**32 - 1, 2**32, 2**63 + 5, 2**64 - 1, 2**100 + 3, 2**200 + 99,
         0x9E3779B97F4A7C15F39CC0605CEDC834]


def words(seed: int) -> np.ndarray:
    """numpy's SeedSequence entropy words: little-endian 32-bit, at least one"""
    w = []
    while True:
        w.append(seed & 0xFFFFFFFF)
        seed >>= 32
        if not seed:
            return np.array(w, dtype=np.uint32)


@pytest.fixture(scope="module")
def lib():
    L = ctypes.CDLL(str(HOSTCHECK))
    L.hc_np_seed.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L.hc_np_draws.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    return L


@pytest.mark.parametrize("seed", SEEDS)
def test_seeding_matches_numpy(lib, seed):
    w = words(seed)
    out = np.zeros(4, dtype=np.uint64)
    lib.hc_np_seed(w.ctypes.data, len(w), out.ctypes.data)
    st = np.random.PCG64(seed).state["state"]
    assert (int(out[0]) << 64 | int(out[1])) == st["state"]
    assert (int(out[2]) << 64 | int(out[3])) == st["inc"]


@pytest.mark.parametrize("seed", [0, 12345, 2**100 + 3])
def test_jump_ahead_outputs_match_numpy(lib, seed):
    w = words(seed)
    seq = np.arange(600, dtype=np.uint64)
    far = np.array([1000, 4095, 65536, 123457, 2**20 + 1, 2**31 - 1, 2**33 + 12345, 2**39 + 7], dtype=np.uint64)
    ks = np.concatenate([seq, far])
    out = np.zeros(len(ks), dtype=np.uint64)
    lib.hc_np_draws(w.ctypes.data, len(w), ks.ctypes.data, len(ks), out.ctypes.data)
    assert np.array_equal(out[:600], np.random.PCG64(seed).random_raw(600))
    for k, v in zip(far, out[600:]):
        g = np.random.PCG64(seed)
        g.advance(int(k))
        assert int(g.random_raw()) == int(v), f"offset {k}"


def test_uint32_halves_and_uniform_transform(lib):
    """integers(0, 2**32, uint32) takes the low half of an output, then keeps
    the high half for the next 32-bit draw across 64-bit draws; uniform is
    lo + (hi - lo) * (x >> 11) * 2**-53 (the device's restatement)."""
    seed = 2024
    w = words(seed)
    ks = np.arange(8, dtype=np.uint64)
    raw = np.zeros(8, dtype=np.uint64)
    lib.hc_np_draws(w.ctypes.data, len(w), ks.ctypes.data, 8, raw.ctypes.data)
    g = np.random.default_rng(seed)
    a = g.integers(0, 2**32, size=3, dtype=np.uint32)
    assert list(a) == [raw[0] & 0xFFFFFFFF, raw[0] >> 32, raw[1] & 0xFFFFFFFF]
    u = g.uniform(-np.pi, np.pi, 2)
    unit = (raw[2:4] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    assert np.array_equal(u, -np.pi + (np.pi - -np.pi) * unit)
    b = g.integers(0, 2**32, size=2, dtype=np.uint32)
    assert list(b) == [raw[1] >> 32, raw[4] & 0xFFFFFFFF]
    st = g.bit_generator.state
    assert st["has_uint32"] == 1 and st["uinteger"] == int(raw[4] >> 32)
