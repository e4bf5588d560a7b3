"""Pin the oracle's flat power-of-n schemes over key-value slots and the
4-byte dynamic_container<key_val_pair, uint16_t> slot (REP_KV16), by
hand-derived placements and the reference's own checks.

  power_of_n_insert_shortcut_scheme (insert_schemes/power_of_n_shortcut.cuh,
      the reference's data_structs::tcf, data_structs/tcf.cuh:29-30):
      like the packed-bucket shortcut scheme, but a bucket's fill counts
      every non-empty slot, tombstones included (check_fill_bucket
      :234-261), and an insert claims empty slots only (:134-191);
  power_of_n_insert_scheme (power_of_n.cuh:326-422): no 75 % shortcut, the
      key goes to the least-filled probe bucket, ties in probe order;
  key_val_pair::atomic_reset (key_val_pair.cuh:143-151) turns the key into
      the tombstone and leaves the value.

A KV16 slot is one little-endian word val << 16 | key16, key16 the key's
low 16 bits with 0 -> 1 and 0xFFFF -> 0xFFFE (dynamic_container.cuh:35-50).
"""
import random

import numpy as np

SEED = 42


def fp16(key):
    f = key & 0xFFFF
    return 1 if f == 0 else (0xFFFE if f == 0xFFFF else f)


def kv16(o, nb, B=16, scheme=None, prober=None, probes=2):
    scheme = o.SCHEME_POWER_OF_N_SHORTCUT_SLOTS if scheme is None else scheme
    prober = o.PROBE_DOUBLE if prober is None else prober
    return o.Table(o.REP_KV16, [o.level(nb * B, scheme, B, tile=4, prober=prober, hasher=o.HASH_MURMUR64A,
                                        max_probes=probes)], seed=SEED)


def buckets_of(o, key, nb, probes=2, prober=None):
    prober = o.PROBE_DOUBLE if prober is None else prober
    return [p % nb for p in o.probes(prober, o.HASH_MURMUR64A, probes, SEED, key, fp16(key))]


def pick(o, rng, nb, want, probes=2):
    """A random key whose probe buckets satisfy want(list)."""
    while True:
        k = rng.getrandbits(64)
        if k & 0xFFFF in (0, 0xFFFF):
            continue
        b = buckets_of(o, k, nb, probes)
        if want(b):
            return k, b


def test_kv16_slot_word_and_remove_keeps_value(oracle_lib):
    o = oracle_lib
    nb, B = 8, 16
    t = kv16(o, nb)
    rng = random.Random(1)
    k, b = pick(o, rng, nb, lambda b: b[0] != b[1])
    assert t.insert(k, 0xBEEF) == 1
    w = t.level_array(0).reshape(nb, B)
    assert int(w[b[0], 0]) == (0xBEEF << 16) | fp16(k)
    assert t.query(k) == (True, 0xBEEF)
    assert t.remove(k) == 1
    assert int(t.level_array(0).reshape(nb, B)[b[0], 0]) == (0xBEEF << 16) | 0xFFFF   # value kept
    assert t.query(k) == (False, 0) and t.fill() == 0
    # the small-key remap: low 16 bits 0 and 0xFFFF store as 1 and 0xFFFE
    for low, stored in ((0, 1), (0xFFFF, 0xFFFE)):
        kk = (rng.getrandbits(48) << 16) | low
        assert t.insert(kk, 5) == 1 and t.query(kk) == (True, 5)
        w = t.level_array(0)
        assert (w == ((5 << 16) | stored)).sum() == 1
        assert t.remove(kk) == 1 and t.query(kk) == (False, 0)


def test_shortcut_counts_tombstones_in_the_fill(oracle_lib):
    """12 keys fill bucket A to the shortcut threshold; removing 4 leaves 8
    live keys and 4 tombstones.  The flat scheme still sees fill 12 and sends
    the next key for A to its second choice; the packed-bucket scheme (which
    does not count tombstones) puts it in A's first empty slot."""
    o = oracle_lib
    nb, B = 16, 16
    rng = random.Random(2)
    A = 3
    first = [pick(o, rng, nb, lambda b: b[0] == A and b[1] != A)[0] for _ in range(12)]
    k, b = pick(o, rng, nb, lambda b: b[0] == A and b[1] != A)
    flat = kv16(o, nb)
    packed = o.Table(o.REP_FP16_V16, [o.level(nb * B, o.SCHEME_SHORTCUT, B, tile=4, prober=o.PROBE_DOUBLE,
                                              hasher=o.HASH_MURMUR64A, max_probes=2)], seed=SEED)
    for t in (flat, packed):
        for q in first:
            assert t.insert(q, 1) == 1
        for q in first[:4]:
            assert t.remove(q) == 1
        assert t.insert(k, 9) == 1
    wf = flat.level_array(0).reshape(nb, B)
    wp = packed.level_array(0).reshape(nb, B)
    assert int(wf[b[1], 0]) == (9 << 16) | fp16(k)          # flat: second choice, slot 0
    assert int(wp[A, 12]) == (9 << 16) | fp16(k)            # packed: A's first empty slot (12)
    assert all(int(w) & 0xFFFF != fp16(k) for w in wf[A])


def test_power_of_n_goes_to_the_least_filled_bucket(oracle_lib):
    """Without the shortcut a key takes the least-filled probe bucket even
    when its first choice is nearly empty; equal fills go in probe order."""
    o = oracle_lib
    nb, B = 16, 8
    rng = random.Random(3)
    t = kv16(o, nb, B, scheme=o.SCHEME_POWER_OF_N)
    A, C = 5, 9
    for _ in range(2):    # A gets 2 keys through keys whose only bucket is A
        q, _ = pick(o, rng, nb, lambda b: b[0] == A and b[1] == A)
        assert t.insert(q, 1) == 1
    q, _ = pick(o, rng, nb, lambda b: b[0] == C and b[1] == C)
    assert t.insert(q, 1) == 1   # C holds 1
    k, b = pick(o, rng, nb, lambda b: b[0] == A and b[1] == C)
    assert t.insert(k, 7) == 1
    w = t.level_array(0).reshape(nb, B)
    assert int(w[C, 1]) == (7 << 16) | fp16(k)               # fill 1 < 2: C, although A came first
    k2, _ = pick(o, rng, nb, lambda b: b[0] == A and b[1] == C)
    assert t.insert(k2, 8) == 1                             # A and C both hold 2: probe order, A
    assert int(t.level_array(0).reshape(nb, B)[A, 2]) == (8 << 16) | fp16(k2)


def test_reference_p2_table_checks(oracle_lib):
    """tests/src/test_components.cu:54,88-200: key_val_pair<u64,u64> slots,
    8-slot buckets, power_of_n_insert_scheme with 3 double-hashing probes,
    seed 1.  Inserted pairs are found with their values, and tables of 1..10
    slots hold that many keys (keys start at 1 here: key 0 is the empty
    key)."""
    o = oracle_lib

    def p2(nslots):
        return o.Table(o.REP_KV64, [o.level(nslots, o.SCHEME_POWER_OF_N, 8, tile=8, prober=o.PROBE_DOUBLE,
                                            hasher=o.HASH_MURMUR64A, max_probes=3)], seed=1)
    t = p2(1100)
    for k, v in ((1, 1), (2, 2), (3, 4)):
        assert t.insert(k, v) == 1
        assert t.query(k) == (True, v)
    for i in range(1, 11):
        t = p2(i)
        keys = np.arange(1, i + 1, dtype=np.uint64)
        assert (t.bulk_insert(keys, keys) == 1).all()
        f, v = t.bulk_query(keys)
        assert f.all() and (v == keys).all()


def test_flat_tcf_with_bucket_backing(oracle_lib):
    """The reference's data_structs::tcf: a flat shortcut KV16 level backed by
    tiny_static_table_4 (KV16, 4-slot buckets, bucket_insert, 20 double-hash
    probes).  Overfilled, it spills into the backing level; every stored key
    is found with its value; removes hit the right level."""
    o = oracle_lib
    t = o.Table(o.REP_KV16, [
        o.level(1 << 10, o.SCHEME_POWER_OF_N_SHORTCUT_SLOTS, 16, tile=4, prober=o.PROBE_DOUBLE,
                hasher=o.HASH_MURMUR64A, max_probes=2),
        o.level(1 << 7, o.SCHEME_BUCKET, 4, tile=4, prober=o.PROBE_DOUBLE, hasher=o.HASH_MURMUR64A,
                max_probes=20)], seed=SEED)
    keys = o.gen_keys(4, 0, 1300)
    vals = ((keys >> np.uint64(20)) & np.uint64(0xFFFF)).astype(np.uint16)
    st = t.bulk_insert(keys, vals)
    assert (st == 2).sum() > 0 and (st == 0).sum() > 0
    assert t.fill() == int((st > 0).sum())
    f, v = t.bulk_query(keys)
    assert f[st > 0].all()
    back = keys[st == 2]
    assert (t.bulk_remove(back) == 1).all()
    w1 = t.level_array(1)
    assert (((w1 & 0xFFFF) == 0xFFFF).sum()) == len(back)   # backing tombstones, one per removed key
