"""Pin the oracle's 8/8 and 10/6 packed slots against hand-derived cases.

Slot words follow representations/grouped_storage_sub_bits.cuh (join :77-95,
retrieve :35-74, predicates :160-195) over the SmallKey remap of
representations/dynamic_container.cuh:35-54:

  8/8   (tests/src/delete_tests.cu:165-179)   uint16 word = val8 << 8 | fp8,
        fp8 = low byte of the key, 0 -> 1, 0xFF -> 0xFE; tombstone word 0x00FF.
  10/6  (tests/src/test_mhm_tcf.cu:105-116)   uint16 word = val6 << 10 |
        (fp16 & 0x3FF), fp16 remapped at 16 bits BEFORE the masking; the
        XOR-power-of-two probe tag is the unmasked fp16; tombstone word 0x03FF.

Each expected word below is written out by hand from those rules, and each
expected slot from the insert schemes (shortcut threshold, lowest empty
slot, CAS from the exact empty / tombstone word), not by calling the oracle.
"""
import random

import pytest

M64 = (1 << 64) - 1
SEED = 42


def fp16(key):
    f = key & 0xFFFF
    return 1 if f == 0 else (0xFFFE if f == 0xFFFF else f)


def fp8(key):
    f = key & 0xFF
    return 1 if f == 0 else (0xFE if f == 0xFF else f)


def w106(key, val):
    return ((val & 0x3F) << 10) | (fp16(key) & 0x3FF)


def w88(key, val):
    return ((val & 0xFF) << 8) | fp8(key)


# ---- 10/6: the reference's test_mhm_tcf composition, shrunk -------------
NB0, B0 = 4, 8      # primary: 32 slots -> 4 buckets of 8, threshold 6
NB1, B1 = 2, 8      # backing: 16 slots


def make106(o, prober=None, nb0=NB0):
    prober = o.PROBE_DOUBLE if prober is None else prober
    return o.Table(o.REP_FP10_V6, [
        o.level(nb0 * B0, o.SCHEME_SHORTCUT, B0, tile=1, prober=prober, hasher=o.HASH_MURMUR64A,
                max_probes=2),
        o.level(NB1 * B1, o.SCHEME_LINEAR_BUCKETS, B1, tile=1, prober=o.PROBE_DOUBLE,
                hasher=o.HASH_MURMUR64A, max_probes=20),
    ], seed=SEED)


def first_bucket(o, key, nb=NB0, prober=None):
    prober = o.PROBE_DOUBLE if prober is None else prober
    return o.probes(prober, o.HASH_MURMUR64A, 2, SEED, key, fp16(key))[0] % nb


def pick(o, rng, b0, low16=None, nb=NB0, avoid=(0, 0x3FF), used=None):
    """A random key whose first probe bucket is b0 (and low 16 bits low16)."""
    while True:
        k = rng.getrandbits(64)
        if low16 is not None:
            k = (k & ~0xFFFF) | low16
        elif (fp16(k) & 0x3FF) in avoid or (used is not None and (fp16(k) & 0x3FF) in used):
            continue
        if first_bucket(o, k, nb) == b0:
            if used is not None:
                used.add(fp16(k) & 0x3FF)
            return k


def bucket(t, lvl, b, B):
    return list(t.level_array(lvl)[b * B:(b + 1) * B])


def test_pack_roundtrip_exhaustive_10_6(oracle_lib):
    o = oracle_lib.lib()
    for w in range(1 << 16):
        k, v = o.or_unpack_fp(10, w), o.or_unpack_val(10, 6, w)
        assert k < 1024 and v < 64 and o.or_pack(10, 6, k, v) == w
    # join clips both fields (grouped_storage_sub_bits.cuh:85-87)
    assert o.or_pack(10, 6, 0x1234, 0x41) == (0x01 << 10) | 0x234
    assert o.or_pack(10, 6, 0xFFFF, 0xFF) == 0xFFFF


def test_10_6_insert_query_remove(oracle_lib):
    o = oracle_lib
    t = make106(o)
    rng = random.Random(1)
    used = set()
    ks = [pick(o, rng, 2, used=used) for _ in range(3)]
    for i, k in enumerate(ks):
        assert t.insert(k, 0x40 + i) == 1           # value keeps its low 6 bits
    assert bucket(t, 0, 2, B0) == [w106(ks[0], 0), w106(ks[1], 1), w106(ks[2], 2)] + [0] * 5
    assert t.query(ks[1]) == (True, 1)
    assert t.remove(ks[1]) == 1
    assert bucket(t, 0, 2, B0)[:3] == [w106(ks[0], 0), 0x03FF, w106(ks[2], 2)]
    assert t.query(ks[1]) == (False, 0)
    assert t.fill() == 2


def test_10_6_masked_fingerprint_collision(oracle_lib):
    # fp16 0x1155 and 0x0D55 differ, their low 10 bits (0x155) do not: the
    # stored word cannot tell them apart.
    o = oracle_lib
    t = make106(o)
    rng = random.Random(2)
    a = pick(o, rng, 1, low16=0x1155)
    b = pick(o, rng, 1, low16=0x0D55)
    t.insert(a, 9)
    assert t.query(b) == (True, 9)


def test_10_6_zero_key_bits_store_as_empty(oracle_lib):
    o = oracle_lib
    t = make106(o)
    rng = random.Random(3)
    z = pick(o, rng, 0, low16=0xFC00)              # fp16 0xFC00 -> key bits 0
    # On an empty table the query matches the first empty slot it sees.
    assert t.query(z) == (True, 0)
    # val 0: the word is 0, the CAS 0 -> 0 "succeeds", nothing changes.
    assert t.insert(z, 0) == 1
    assert all(w == 0 for w in t.level_array(0)) and t.fill() == 0
    # val 5: word 5 << 10 lands in slot 0 and still counts as empty.
    assert t.insert(z, 5) == 1
    assert bucket(t, 0, 0, B0)[0] == 5 << 10 and t.fill() == 0
    assert t.query(z) == (True, 5)
    # A later key sees slot 0 as empty, its CAS from 0 fails, slot 1 takes it.
    k = pick(o, rng, 0)
    assert t.insert(k, 7) == 1
    assert bucket(t, 0, 0, B0)[:3] == [5 << 10, w106(k, 7), 0]


def test_10_6_all_ones_key_bits_store_as_tombstone(oracle_lib):
    o = oracle_lib
    t = make106(o)
    rng = random.Random(4)
    y = pick(o, rng, 3, low16=0x07FF)              # fp16 0x07FF -> key bits 0x3FF
    assert t.insert(y, 0) == 1
    assert bucket(t, 0, 3, B0)[0] == 0x03FF and t.fill() == 0
    assert t.query(y) == (True, 0)
    # plain insert skips the tombstone-looking slot (not empty) ...
    k1 = pick(o, rng, 3)
    t.insert(k1, 1)
    assert bucket(t, 0, 3, B0)[:2] == [0x03FF, w106(k1, 1)]
    # ... insert_with_delete's CAS from the tombstone word 0x03FF succeeds on it
    k2 = pick(o, rng, 3)
    t.insert(k2, 2, with_delete=True)
    assert bucket(t, 0, 3, B0)[:3] == [w106(k2, 2), w106(k1, 1), 0]
    assert t.query(y) == (False, 0)
    # with a nonzero value the word is 0x07FF: the tombstone CAS fails on it
    y2 = pick(o, rng, 1, low16=0x07FF)
    t.insert(y2, 1)
    k3 = pick(o, rng, 1)
    t.insert(k3, 3, with_delete=True)
    assert bucket(t, 0, 1, B0)[:3] == [0x07FF, w106(k3, 3), 0]


def test_10_6_xor_tag_is_unmasked_fp16(oracle_lib):
    """xor_power_of_two.cuh: second choice = h1 ^ (tag * 0x5bd1e995) with tag
    the SmallKey (16-bit) fingerprint, not its 10 stored bits."""
    o = oracle_lib
    nb = 13  # not a power of two: the two tags' low bits agree, so % 2^k could not tell them apart
    t = make106(o, prober=o.PROBE_XOR_P2, nb0=nb)
    rng = random.Random(5)

    def choices(k, tag):
        h1 = o.hash64(o.HASH_MURMUR64A, k, SEED)
        return h1 % nb, (h1 ^ ((tag * 0x5bd1e995) & M64)) % nb

    while True:
        k = rng.getrandbits(64)
        f = fp16(k)
        if (f & 0x3FF) in (0, 0x3FF) or f == (f & 0x3FF):
            continue
        b0, b16 = choices(k, f)
        _, b10 = choices(k, f & 0x3FF)
        if len({b0, b16, b10}) == 3:
            break
    # fill b0 to the shortcut threshold (6 of 8) through keys whose first choice is b0
    used = {fp16(k) & 0x3FF}
    for i in range(6):
        assert t.insert(pick(o, rng, b0, nb=nb, used=used), i) == 1
    assert t.insert(k, 1) == 1
    assert bucket(t, 0, b16, B0)[0] == w106(k, 1)
    assert bucket(t, 0, b10, B0)[0] == 0


# ---- 8/8: the reference's del_TCF_small composition, shrunk --------------
NB8, B8 = 4, 16


def make88(o):
    return o.Table(o.REP_FP8_V8, [
        o.level(NB8 * B8, o.SCHEME_SHORTCUT, B8, tile=4, prober=o.PROBE_XOR_P2, hasher=o.HASH_MURMUR64A,
                max_probes=2),
        o.level(32, o.SCHEME_LINEAR_BUCKETS, 8, tile=4, prober=o.PROBE_LINEAR, hasher=o.HASH_MURMUR64A,
                max_probes=400),
    ], seed=SEED)


def choices8(o, k, nb=NB8):
    h1 = o.hash64(o.HASH_MURMUR64A, k, SEED)
    return h1 % nb, (h1 ^ ((fp8(k) * 0x5bd1e995) & M64)) % nb


def test_pack_roundtrip_exhaustive_8_8_words(oracle_lib):
    o = oracle_lib.lib()
    for w in range(1 << 16):
        assert o.or_pack(8, 8, o.or_unpack_fp(8, w), o.or_unpack_val(8, 8, w)) == w


def test_8_8_fingerprint_value_and_tombstone(oracle_lib):
    o = oracle_lib
    t = make88(o)
    rng = random.Random(6)
    # low byte 0x00 -> fp 1, 0xFF -> 0xFE; values keep 8 bits
    ks = []
    for low in (0x00, 0xFF, 0x37):
        while True:
            k = (rng.getrandbits(64) & ~0xFF) | low
            if choices8(o, k)[0] == 2:
                break
        ks.append(k)
    for i, k in enumerate(ks):
        assert t.insert(k, 0x100 + 0x11 * i) == 1
    assert bucket(t, 0, 2, B8)[:4] == [0x0001, (0x11 << 8) | 0xFE, (0x22 << 8) | 0x37, 0]
    assert [t.query(k) for k in ks] == [(True, 0x00), (True, 0x11), (True, 0x22)]
    assert t.remove(ks[1]) == 1
    assert bucket(t, 0, 2, B8)[:3] == [0x0001, 0x00FF, (0x22 << 8) | 0x37]
    assert t.fill() == 2


def test_8_8_xor_tag_is_fp8(oracle_lib):
    """For the 8/8 slot the SmallKey is uint8: the XOR tag is the 8-bit fp."""
    o = oracle_lib
    nb = 63  # not a power of two (see the 10/6 case)
    t = o.Table(o.REP_FP8_V8, [o.level(nb * B8, o.SCHEME_SHORTCUT, B8, tile=4, prober=o.PROBE_XOR_P2,
                                       hasher=o.HASH_MURMUR64A, max_probes=2)], seed=SEED)
    rng = random.Random(7)
    while True:
        k = rng.getrandbits(64)
        b0, b8 = choices8(o, k, nb)
        h1 = o.hash64(o.HASH_MURMUR64A, k, SEED)
        b16 = (h1 ^ ((fp16(k) * 0x5bd1e995) & M64)) % nb
        if len({b0, b8, b16}) == 3:
            break
    n = 0
    while n < 12:                                  # threshold 12 of 16
        f = rng.getrandbits(64)
        if choices8(o, f, nb)[0] == b0:
            assert t.insert(f, 0) == 1
            n += 1
    assert t.insert(k, 0x5A) == 1
    assert bucket(t, 0, b8, B8)[0] == w88(k, 0x5A)


@pytest.mark.parametrize("rep", ["FP8_V8", "FP10_V6"])
def test_narrow_formats_no_false_negatives_for_stored_keys(oracle_lib, rep):
    """Every key whose word is neither empty- nor tombstone-looking is found
    after a 90 % fill, and fill counts exactly those keys."""
    o = oracle_lib
    r = getattr(o, "REP_" + rep)
    nslots = 1 << 14
    t = o.Table(r, [
        o.level(nslots, o.SCHEME_SHORTCUT, 16, tile=4, prober=o.PROBE_XOR_P2, hasher=o.HASH_MURMUR64A),
        o.level(nslots // 16, o.SCHEME_LINEAR_BUCKETS, 8, tile=4, prober=o.PROBE_DOUBLE,
                hasher=o.HASH_MURMUR64A, max_probes=20),
    ])
    keys = o.gen_keys(11, 0, int(0.9 * nslots))
    st = t.bulk_insert(keys)
    kb = 8 if rep == "FP8_V8" else 10
    f = (keys & 0xFFFF).astype("int64") if kb == 10 else (keys & 0xFF).astype("int64")
    top = 0xFFFF if kb == 10 else 0xFF
    f[f == 0] = 1
    f[f == top] = top - 1
    key_bits = f & ((1 << kb) - 1)
    normal = (key_bits != 0) & (key_bits != (1 << kb) - 1)
    found, _ = t.bulk_query(keys)
    assert found[(st > 0) & normal].all()
    assert t.fill() == int(((st > 0) & normal).sum())
