"""Pin the oracle's 12-bit sub-word bucket and the cache-line-grouped bucket
layout against hand-derived words.

12-bit bucket (representations/12_bit_bucket.cuh:33-313 over the sub-word
helpers of representation_helpers.cuh:451-1005): slots are packed four to
three uint16 words, and each slot keeps its own 12 bits of the key's low 16
bits, with its own empty-avoiding remap:

  slot 4g+0  word 3g   bits 0-11   m = item & 0x0FFF,           0 -> 0x0001
  slot 4g+1  word 3g   bits 12-15  m = item & 0xF0FF,   (item & 0xF000) == 0 -> m |= 0x1000
             word 3g+1 bits 0-7
  slot 4g+2  word 3g+1 bits 8-15   m = item & 0xFF0F,   (item & 0xFF00) == 0 -> m |= 0x0100
             word 3g+2 bits 0-3
  slot 4g+3  word 3g+2 bits 4-15   m = item & 0xFFF0,           0 -> 0x0010

item = the key's low 16 bits (no 16-bit remap); a query matches slot i when
its composite equals m(item, i % 4); there are no values (queries return 0)
and no tombstones (the reference's remove does not compile for this bucket).

Grouped buckets (insert_schemes/grouped_power_buckets.cuh:27-388,
blocked_bucket_insert<K>): K buckets share one 128-byte line, bucket b lives
at byte 128*(b / K) + bucket_bytes*(b % K), the rest of the line is padding;
insert / query / remove are the shortcut scheme's.
"""
import random

SEED = 42


def proj(item, o):
    item &= 0xFFFF
    if o == 0:
        m = item & 0x0FFF
        return m if m else 1
    if o == 1:
        m = item & 0xF0FF
        return m if item & 0xF000 else m | 0x1000
    if o == 2:
        m = item & 0xFF0F
        return m if item & 0xFF00 else m | 0x0100
    m = item & 0xFFF0
    return m if m else 0x0010


def group_words(m0, m1, m2, m3):
    """The three uint16 words of one slot group holding composites m0..m3."""
    return [(m0 & 0x0FFF) | (m1 & 0xF000),
            (m1 & 0x00FF) | (m2 & 0xFF00),
            (m2 & 0x000F) | (m3 & 0xFFF0)]


def make12(o, nb=8, B=16, group=0):
    scheme = o.SCHEME_GROUPED if group else o.SCHEME_SHORTCUT
    return o.Table(o.REP_FP12, [o.level(nb * B, scheme, B, tile=4, prober=o.PROBE_DOUBLE,
                                        hasher=o.HASH_MURMUR64A, max_probes=2, group=group)], seed=SEED)


def first_bucket(o, key, nb):
    return o.probes(o.PROBE_DOUBLE, o.HASH_MURMUR64A, 2, SEED, key, 0)[0] % nb


def pick(o, rng, b0, nb, low16):
    while True:
        k = (rng.getrandbits(64) & ~0xFFFF) | low16
        if first_bucket(o, k, nb) == b0:
            return k


def test_projection_remaps_closed_forms():
    # the four slot offsets keep different key bits
    assert [proj(0x1234, o) for o in range(4)] == [0x234, 0x1034, 0x1204, 0x1230]
    # each remap fires exactly when the first stored part would be zero
    assert [proj(0x5000, 0), proj(0x0A5B, 1), proj(0x00C7, 2), proj(0x000F, 3)] == [0x001, 0x105B, 0x0107, 0x0010]
    assert group_words(0x001, 0x105B, 0x0107, 0x0010) == [0x1001, 0x015B, 0x0017]


def test_12bit_slots_words_and_queries(oracle_lib):
    o = oracle_lib
    nb = 8
    t = make12(o, nb)
    rng = random.Random(1)
    lows = [0x5000, 0x0A5B, 0x00C7, 0x000F, 0xBEEF]
    ks = [pick(o, rng, 3, nb, low) for low in lows]
    for k in ks:
        assert t.insert(k) == 1
    words = list(t.level_array(0))
    B_WORDS = 12                                   # (12*16 - 1) / 16 + 1
    bucket = words[3 * B_WORDS:4 * B_WORDS]
    m = [proj(low, i % 4) for i, low in enumerate(lows)]
    assert bucket[:6] == group_words(*m[:4]) + group_words(m[4], 0, 0, 0)
    assert all(w == 0 for w in bucket[6:]) and all(w == 0 for i, w in enumerate(words) if i // 12 != 3)
    assert t.fill() == 5
    assert all(t.query(k) == (True, 0) for k in ks)
    # slot 1 ignores key bits 8-11: a key differing only there is a false positive
    fp1 = pick(o, rng, 3, nb, 0x0A5B ^ 0x0300)
    assert t.query(fp1) == (True, 0)
    # a key whose slot-0 composite is 1 through the remap (0x7000) matches slot 0
    assert t.query(pick(o, rng, 3, nb, 0x7000)) == (True, 0)
    # a key matching no slot's composite misses
    assert t.query(pick(o, rng, 3, nb, 0x0F0F)) == (False, 0)


def test_12bit_shortcut_threshold(oracle_lib):
    """12 of 16 slots filled sends the next key to its second choice
    (power_of_n_shortcut_buckets.cuh:198-306, fill < 0.75 * B)."""
    o = oracle_lib
    nb = 8
    t = make12(o, nb)
    rng = random.Random(2)
    for i in range(12):
        t.insert(pick(o, rng, 5, nb, 0x1111 * (i % 15 + 1)))
    while True:
        k = rng.getrandbits(64)
        p = o.probes(o.PROBE_DOUBLE, o.HASH_MURMUR64A, 2, SEED, k, 0)
        if p[0] % nb == 5 and p[1] % nb != 5:
            break
    b1 = p[1] % nb
    assert t.insert(k) == 1
    words = list(t.level_array(0))
    assert words[b1 * 12] == proj(k, 0)                 # slot 0 of the second choice
    assert (words[5 * 12 + 9] & 0x0FFF) == 0            # slot 12 of the first stays empty


def test_grouped_layout_12bit(oracle_lib):
    """blocked_bucket_insert<5>: five 24-byte buckets per 128-byte line."""
    o = oracle_lib
    nb = 12
    t = make12(o, nb, group=5)
    assert len(t.level_bytes(0)) == 3 * 128          # ceil(12 / 5) lines
    rng = random.Random(3)
    k = pick(o, rng, 7, nb, 0x0123)
    assert t.insert(k) == 1
    raw = t.level_bytes(0)
    off = 128 * (7 // 5) + 24 * (7 % 5)
    assert int.from_bytes(raw[off:off + 2], "little") == 0x0123
    assert sum(raw[:off]) == 0 and sum(raw[off + 2:]) == 0
    assert t.query(k) == (True, 0)


def test_grouped_layout_16_16(oracle_lib):
    """blocked_bucket_insert<3> over 8-slot 16/16 buckets (32 B): three buckets
    per line, 32 B of padding."""
    o = oracle_lib
    nb = 7
    t = o.Table(o.REP_FP16_V16, [o.level(nb * 8, o.SCHEME_GROUPED, 8, tile=4, prober=o.PROBE_DOUBLE,
                                         hasher=o.HASH_MURMUR64A, max_probes=2, group=3)], seed=SEED)
    assert len(t.level_bytes(0)) == 3 * 128
    rng = random.Random(4)
    keys = []
    for b in range(nb):
        k = pick(o, rng, b, nb, 0x0100 + b)
        keys.append(k)
        assert t.insert(k, 0x40 + b) == 1
    raw = t.level_bytes(0)
    for b in range(nb):
        off = 128 * (b // 3) + 32 * (b % 3)
        assert int.from_bytes(raw[off:off + 4], "little") == ((0x40 + b) << 16) | (0x0100 + b)
    for line in range(3):
        assert sum(raw[128 * line + 96:128 * line + 128]) == 0
    assert [t.query(k) for k in keys] == [(True, 0x40 + b) for b in range(nb)]
    assert t.remove(keys[4]) == 1
    off = 128 * (4 // 3) + 32 * (4 % 3)
    assert int.from_bytes(raw[off:off + 4], "little") != 0xFFFF
    assert int.from_bytes(t.level_bytes(0)[off:off + 4], "little") == 0xFFFF
