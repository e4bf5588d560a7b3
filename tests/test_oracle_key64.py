"""Pin the oracle's key-only table: the free table of the CMS allocator,
static_table<uint64_t, uint16_t, dynamic_container<key_container, uint64_t>,
1, 4, bucket_insert, 20, doubleHasher, murmurHasher> (allocators/cms.cuh:203).

8-byte slots hold get_smallKey(key) with SmallKey = uint64_t (key 0 -> 1,
~0 -> ~0 - 1; representations/dynamic_container.cuh:35-50, key_only.cuh);
empty 0, tombstone ~0.  bucket_insert (insert_schemes/bucket_insert.cuh):
insert takes the lowest empty slot of the first probe bucket that has one;
query / remove walk stripes of P slots (P = 1 here: slot by slot) and stop
at the first empty slot.  Expected slots are derived by hand from the
double-hashing probes (pinned in test_oracle_hash.py)."""
import random

SEED = 42
M64 = (1 << 64) - 1
NB, B = 8, 4


def make(o, nb=NB):
    return o.Table(o.REP_KEY64, [o.level(nb * B, o.SCHEME_BUCKET, B, tile=1, prober=o.PROBE_DOUBLE,
                                         hasher=o.HASH_MURMUR64A, max_probes=20)], seed=SEED)


def probes(o, key, nb=NB):
    return [p % nb for p in o.probes(o.PROBE_DOUBLE, o.HASH_MURMUR64A, 20, SEED, key, key)]


def pick(o, rng, b0, low=None):
    while True:
        k = rng.getrandbits(64)
        if low is not None:
            k = low
        if probes(o, k)[0] == b0:
            return k
        if low is not None:
            return None


def bucket(t, b):
    return [int(x) for x in t.level_array(0)[b * B:(b + 1) * B]]


def test_insert_lowest_empty_then_next_probe(oracle_lib):
    o = oracle_lib
    t = make(o)
    rng = random.Random(1)
    ks = [pick(o, rng, 2) for _ in range(5)]
    for k in ks:
        assert t.insert(k) == 1
    assert bucket(t, 2) == ks[:4]
    # bucket 2 is full: the fifth key takes slot 0 of its second probe bucket
    b1 = probes(o, ks[4])[1]
    if b1 != 2:
        assert bucket(t, b1)[0] == ks[4]
    assert all(t.query(k) == (True, 0) for k in ks)      # key-only: value 0
    assert t.fill() == 5


def test_reserved_keys_are_remapped(oracle_lib):
    o = oracle_lib
    t = make(o)
    b0 = probes(o, 0)[0]
    assert t.insert(0) == 1
    assert bucket(t, b0)[0] == 1                          # 0 is stored as 1
    # key 1 is a different key to probe with; where its probes meet the
    # stored 1 it is found, as in the reference
    if probes(o, 1)[0] == b0:
        assert t.query(1) == (True, 0)
    bt = probes(o, M64)[0]
    assert t.insert(M64) == 1
    assert M64 - 1 in bucket(t, bt)                       # ~0 is stored as ~0 - 1
    assert t.query(M64) == (True, 0) and t.query(0) == (True, 0)


def test_remove_tombstone_and_query_stop_rule(oracle_lib):
    o = oracle_lib
    t = make(o)
    rng = random.Random(2)
    ks = [pick(o, rng, 5) for _ in range(3)]
    for k in ks:
        t.insert(k)
    assert t.remove(ks[0]) == 1
    assert bucket(t, 5) == [M64, ks[1], ks[2], 0]
    # the tombstone is not empty: the query walks past it
    assert t.query(ks[1]) == (True, 0) and t.query(ks[0]) == (False, 0)
    # a plain insert does not reuse the tombstone: lowest EMPTY slot
    k4 = pick(o, rng, 5)
    t.insert(k4)
    assert bucket(t, 5) == [M64, ks[1], ks[2], k4]
    # an absent key whose first bucket has an empty slot stops there
    t2 = make(o)
    k = pick(o, rng, 6)
    t2.insert(k)
    other = pick(o, rng, 6)
    assert t2.query(other) == (False, 0)
    assert t2.fill() == 1
