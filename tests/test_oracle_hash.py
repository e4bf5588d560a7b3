"""Pin the oracle's hashers and probing schemes (CPU only).

Checks come from outside the oracle: SMHasher's published verification codes,
values recorded from the reference's own headers, and closed forms.
"""
import json
import os
import random

import pytest

GOLDEN = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")))
M64 = (1 << 64) - 1


def smhasher_verification(fn, hashbytes):
    key = bytearray(256)
    hashes = bytearray()
    for i in range(256):
        key[i] = i
        hashes += fn(bytes(key[:i]), 256 - i)
    return int.from_bytes(fn(bytes(hashes), 0)[:4], "little")


def test_murmur64a_smhasher_code(oracle_lib):
    o = oracle_lib
    code = smhasher_verification(lambda d, s: o.murmur64a(d, s).to_bytes(8, "little"), 8)
    assert code == int(GOLDEN["smhasher_verification"]["MurmurHash64A"], 16)


def test_murmur3_x64_128_smhasher_code_pins_fmix64(oracle_lib):
    # MurmurHash3_x64_128 finishes with two fmix64 calls; a wrong shift or
    # constant in or_fmix64 changes this code.
    o = oracle_lib
    code = smhasher_verification(lambda d, s: o.murmur3_x64_128(d, s), 16)
    assert code == int(GOLDEN["smhasher_verification"]["MurmurHash3_x64_128"], 16)


def test_recorded_reference_hash(oracle_lib):
    rec = GOLDEN["reference_recorded"]
    assert oracle_lib.murmur64a_u64(1, 42) == int(rec["murmur64a_key1_seed42"], 16)
    assert oracle_lib.hash64(oracle_lib.HASH_MURMUR64A, 1, 42) == int(rec["murmur64a_key1_seed42"], 16)


def test_recorded_double_hash_probes(oracle_lib):
    o = oracle_lib
    rec = [int(x, 16) for x in GOLDEN["reference_recorded"]["double_hash_probes_key1_seed42"]]
    got = o.probes(o.PROBE_DOUBLE, o.HASH_MURMUR64A, 20, 42, 1, 1, cap=2)
    assert got == rec


def fmix64_inverse(h):
    # Independently derived inverse: x ^= x >> 33 is an involution for
    # shifts >= 32; the multipliers are inverted mod 2^64.
    c1 = pow(0xFF51AFD7ED558CCD, -1, 1 << 64)
    c2 = pow(0xC4CEB9FE1A85EC53, -1, 1 << 64)
    h ^= h >> 33
    h = (h * c2) & M64
    h ^= h >> 33
    h = (h * c1) & M64
    h ^= h >> 33
    return h


def test_fmix64_inverts(oracle_lib):
    rng = random.Random(7)
    assert oracle_lib.fmix64(0) == 0
    for _ in range(2000):
        x = rng.getrandbits(64)
        assert fmix64_inverse(oracle_lib.fmix64(x)) == x


def test_hash_seeding(oracle_lib):
    o = oracle_lib
    # FMIX64 hasher = fmix64(key ^ seed)
    for k, s in [(1, 42), (0xDEADBEEF, 1764), (M64 - 5, 0)]:
        assert o.hash64(o.HASH_FMIX64, k, s) == o.fmix64(k ^ s)


@pytest.mark.parametrize("hasher", [0, 1])
def test_probe_closed_forms(oracle_lib, hasher):
    o = oracle_lib
    rng = random.Random(11 + hasher)
    seed = 42
    for _ in range(200):
        key = rng.getrandbits(64)
        tag = rng.getrandbits(16)
        h1 = o.hash64(hasher, key, seed)
        h2 = o.hash64(hasher, key, (seed * seed) & M64)
        # linear: h + i (linear_probing.cuh:72-78); stops at max_probes or ~0
        exp = []
        for i in range(7):
            v = (h1 + i) & M64
            if v == M64:
                break
            exp.append(v)
        assert o.probes(o.PROBE_LINEAR, hasher, 7, seed, key, tag) == exp
        # double: h1 + i*h2 with h2 seeded seed^2 (double_hashing.cuh:66-89)
        exp = [(h1 + i * h2) & M64 for i in range(5)]
        assert o.probes(o.PROBE_DOUBLE, hasher, 5, seed, key, tag) == exp
        # power of two: {h1, h2} (power_of_two.cuh:74-91)
        assert o.probes(o.PROBE_P2, hasher, 2, seed, key, tag) == [h1, h2]
        # xor power of two: {h1, h1 ^ (tag * 0x5bd1e995)} (xor_power_of_two.cuh:69-84)
        assert o.probes(o.PROBE_XOR_P2, hasher, 2, seed, key, tag) == [h1, (h1 ^ (tag * 0x5BD1E995)) & M64]
        # split64: low and high 32-bit halves of one hash
        assert o.probes(o.PROBE_SPLIT64, hasher, 2, seed, key, tag) == [h1 & 0xFFFFFFFF, h1 >> 32]


def test_key_stream_is_distinct_and_matches_splitmix(oracle_lib):
    import numpy as np
    keys = oracle_lib.gen_keys(123, 0, 1 << 16)
    assert len(np.unique(keys)) == len(keys)
    assert not np.any(keys == 0) and not np.any(keys == np.uint64(M64))
    # splitmix64 (Vigna): state += golden gamma, then the mix; first output
    # of a stream seeded s is mix(s + gamma).
    def mix(z):
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
        return z ^ (z >> 31)
    for i in range(50):
        assert int(keys[i]) == mix((123 + (i + 1) * 0x9E3779B97F4A7C15) & M64)
