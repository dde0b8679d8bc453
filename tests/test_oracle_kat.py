"""Known-answer tests pinning the oracle's random-number layer.

The reference's streams are TEA-seeded (shaders/random.glsl:13-27) 24-bit
Numerical-Recipes LCGs (random.glsl:31-43); per-sample frame seeds come from
PCG32 (O'Neill 2014).  Each generator is pinned by published outputs, not by
re-typing the oracle's formula.
"""
import numpy as np
import pytest

import rtxamd
from rtxamd import abi
from oracle.binding import lib

M32 = 0xFFFFFFFF


def tea_generic(v0, v1, key, cycles):
    """Wheeler & Needham's TEA encryption (1994) with `cycles` cycles."""
    s = 0
    for _ in range(cycles):
        s = (s + 0x9E3779B9) & M32
        v0 = (v0 + ((((v1 << 4) + key[0]) & M32) ^ ((v1 + s) & M32) ^ (((v1 >> 5) + key[1]) & M32))) & M32
        v1 = (v1 + ((((v0 << 4) + key[2]) & M32) ^ ((v0 + s) & M32) ^ (((v0 >> 5) + key[3]) & M32))) & M32
    return v0, v1


def test_tea_generic_matches_published_vector():
    # TEA test vector: all-zero key and block -> 41ea3a0a 94baa940 (32 cycles)
    assert tea_generic(0, 0, (0, 0, 0, 0), 32) == (0x41EA3A0A, 0x94BAA940)


def test_oracle_tea_is_16_cycle_tea_with_the_references_key():
    L = lib(abi)
    key = (0xA341316C, 0xC8013EA4, 0xAD90777D, 0x7E95761E)  # shaders/random.glsl:22-23
    rng = np.random.default_rng(1)
    for a, b in rng.integers(0, 2**32, size=(200, 2), dtype=np.uint64):
        assert L.orc_tea(int(a), int(b)) == tea_generic(int(a), int(b), key, 16)[0]
    assert L.orc_tea(0, 0) == tea_generic(0, 0, key, 16)[0]


def test_lcg_numerical_recipes_sequence():
    # Numerical Recipes "ranqd1": from 0 the generator yields 1013904223, 1196435762
    L = lib(abi)
    s1 = L.orc_lcg_step(0)
    s2 = L.orc_lcg_step(s1)
    assert (s1, s2) == (1013904223, 1196435762)
    # period structure: full period mod 2^32 means the low bit alternates
    bits = []
    s = 12345
    for _ in range(16):
        s = L.orc_lcg_step(s)
        bits.append(s & 1)
    assert bits == [bits[0], 1 - bits[0]] * 8


def test_pcg32_demo_outputs():
    # pcg32-demo: pcg32_srandom_r(&rng, 42u, 54u) prints 0xa15c02b7 first
    L = lib(abi)
    assert L.orc_pcg32_first(42, 54) == 0xA15C02B7
    assert rtxamd.pcg32_first(42, 54) == 0xA15C02B7


@pytest.mark.parametrize("state", [0, 1, 0x00FFFFFF, 0xDEADBEEF, M32])
def test_sampler_draw_count_and_24bit_floats(state):
    """orc_sample_cosine consumes exactly two draws, u then phi, each the
    24-bit LCG output over 2^24 (random.glsl:42, 86-94)."""
    L = lib(abi)
    import ctypes as C
    st = C.c_uint32(state)
    out = (C.c_double * 3)()
    L.orc_sample_cosine((C.c_double * 3)(0, 0, 1), C.byref(st), out)
    s1 = L.orc_lcg_step(state)
    s2 = L.orc_lcg_step(s1)
    assert st.value == s2
    u = (s1 & 0xFFFFFF) / 2**24
    phi = 2 * np.pi * (s2 & 0xFFFFFF) / 2**24
    # local direction (sqrt(u) cos phi, sqrt(u) sin phi, sqrt(1-u)) around +z
    assert out[2] == pytest.approx(np.sqrt(1 - u), abs=1e-12)
    assert np.hypot(out[0], out[1]) == pytest.approx(np.sqrt(u), abs=1e-12)
