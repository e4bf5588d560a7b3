"""Pin the oracle's power_of_n_insert_shortcut_replace_scheme
(insert_schemes/power_of_n_shortcut_key_replace.cuh:39-523) by hand-derived
placements and a dictionary model.

The scheme is the flat shortcut scheme (power_of_n_shortcut.cuh) with two
changes to insert: a probe bucket that holds the key is entered whatever its
fill (check_fill_bucket returns Bucket_Size + 1, :233-270), and a claim takes
the lowest slot that is empty or holds the key (insert_into_bucket,
:134-190), so a held key gets the new value in place.  The reference header
does not compile; the oracle reads it with its three slips fixed (the
missing brace after :366, found_ballot starting true at :235, atomic_swap
swapping only from the empty key).  Because a bucket's fill counts
tombstones and never falls, a key that went to a later probe bucket finds
every earlier one at 3B/4 or fuller for good, so a sequential re-insert
always lands on the key's own slot.
"""
import random


SEED = 42


def fp16(key):
    f = key & 0xFFFF
    return 1 if f == 0 else (0xFFFE if f == 0xFFFF else f)


def table(o, rep, nb, B=16, scheme=None, probes=2):
    scheme = o.SCHEME_POWER_OF_N_SHORTCUT_REPLACE if scheme is None else scheme
    return o.Table(rep, [o.level(nb * B, scheme, B, tile=4, prober=o.PROBE_DOUBLE, hasher=o.HASH_MURMUR64A,
                                 max_probes=probes)], seed=SEED)


def buckets_of(o, rep, key, nb):
    tag = key if rep == o.REP_KV64 else fp16(key) if rep == o.REP_KV16 else key
    return [p % nb for p in o.probes(o.PROBE_DOUBLE, o.HASH_MURMUR64A, 2, SEED, key, tag)]


def pick(o, rep, rng, nb, want):
    while True:
        k = rng.getrandbits(64)
        if k & 0xFFFF in (0, 0xFFFF) or k in (0, (1 << 64) - 1):
            continue
        b = buckets_of(o, rep, k, nb)
        if want(b):
            return k, b


def test_reinsert_replaces_the_value_in_place(oracle_lib):
    o = oracle_lib
    nb, B = 8, 16
    rng = random.Random(1)
    t = table(o, o.REP_KV16, nb)
    k, b = pick(o, o.REP_KV16, rng, nb, lambda b: b[0] != b[1])
    assert t.insert(k, 5) == 1
    assert int(t.level_array(0).reshape(nb, B)[b[0], 0]) == (5 << 16) | fp16(k)
    assert t.insert(k, 9) == 1
    w = t.level_array(0).reshape(nb, B)
    assert int(w[b[0], 0]) == (9 << 16) | fp16(k)          # same slot, new value
    assert (w & 0xFFFF != 0).sum() == 1 and t.fill() == 1
    assert t.query(k) == (True, 9)
    # the flat shortcut scheme without replace stores a second copy instead
    s = table(o, o.REP_KV16, nb, scheme=o.SCHEME_POWER_OF_N_SHORTCUT_SLOTS)
    assert s.insert(k, 5) == 1 and s.insert(k, 9) == 1
    ws = s.level_array(0).reshape(nb, B)
    assert int(ws[b[0], 0]) == (5 << 16) | fp16(k) and int(ws[b[0], 1]) == (9 << 16) | fp16(k)
    assert s.fill() == 2 and s.query(k) == (True, 5)


def test_replace_enters_a_full_bucket(oracle_lib):
    """Bucket A full (16 keys whose two probes are both A); re-inserting its
    first key still lands on that key's slot (fill B + 1 enters A in the
    first pass) and the table stays full."""
    o = oracle_lib
    nb, B = 4, 16
    rng = random.Random(2)
    A = 1
    keys = [pick(o, o.REP_KV64, rng, nb, lambda b: b[0] == A and b[1] == A)[0] for _ in range(B)]
    t = table(o, o.REP_KV64, nb)
    for i, k in enumerate(keys):
        assert t.insert(k, 100 + i) == 1
    w = t.level_array(0).reshape(nb, B, 2)
    assert [int(x) for x in w[A, :, 0]] == keys                 # filled in insertion order
    assert t.insert(keys[7], 0xFEEDFACECAFEBEEF) == 1
    w = t.level_array(0).reshape(nb, B, 2)
    assert int(w[A, 7, 0]) == keys[7] and int(w[A, 7, 1]) == 0xFEEDFACECAFEBEEF
    assert t.fill() == B and t.query(keys[7]) == (True, 0xFEEDFACECAFEBEEF)
    # a new key of A has nowhere to go
    k2 = pick(o, o.REP_KV64, rng, nb, lambda b: b[0] == A and b[1] == A)[0]
    assert t.insert(k2, 1) == 0


def test_reinsert_finds_the_key_in_its_second_bucket(oracle_lib):
    """A at the 3B/4 mark when k arrives: k goes to its second bucket C.  A
    re-insert skips A (fill 12, no k) and replaces k's value in C."""
    o = oracle_lib
    nb, B = 16, 16
    rng = random.Random(3)
    A = 5
    k, b = pick(o, o.REP_KV16, rng, nb, lambda b: b[0] == A and b[1] != A)
    C = b[1]
    t = table(o, o.REP_KV16, nb)
    for _ in range(12):
        q = pick(o, o.REP_KV16, rng, nb, lambda bb: bb[0] == A and bb[1] not in (A, C))[0]
        assert t.insert(q, 1) == 1
    assert t.insert(k, 7) == 1
    w = t.level_array(0).reshape(nb, B)
    assert int(w[C, 0]) == (7 << 16) | fp16(k)
    assert t.insert(k, 8) == 1
    w = t.level_array(0).reshape(nb, B)
    assert int(w[C, 0]) == (8 << 16) | fp16(k) and t.fill() == 13
    assert all(int(x) & 0xFFFF != fp16(k) for x in w[A])


def test_removed_key_is_inserted_afresh(oracle_lib):
    """A tombstone does not hold the key: re-inserting a removed key takes
    the next empty slot (the tombstone keeps counting in the fill)."""
    o = oracle_lib
    nb, B = 8, 16
    rng = random.Random(4)
    k, b = pick(o, o.REP_KV16, rng, nb, lambda b: b[0] != b[1])
    t = table(o, o.REP_KV16, nb)
    assert t.insert(k, 3) == 1 and t.remove(k) == 1
    assert t.insert(k, 4) == 1
    w = t.level_array(0).reshape(nb, B)
    assert int(w[b[0], 0]) == (3 << 16) | 0xFFFF and int(w[b[0], 1]) == (4 << 16) | fp16(k)
    assert t.query(k) == (True, 4)


def test_key_only_slots_keep_one_copy(oracle_lib):
    o = oracle_lib
    nb, B = 8, 16
    rng = random.Random(5)
    k, b = pick(o, o.REP_KEY64, rng, nb, lambda b: b[0] != b[1])
    t = table(o, o.REP_KEY64, nb)
    s = table(o, o.REP_KEY64, nb, scheme=o.SCHEME_POWER_OF_N_SHORTCUT_SLOTS)
    for x in (t, s):
        assert x.insert(k) == 1 and x.insert(k) == 1
    assert t.fill() == 1 and s.fill() == 2


def test_insert_or_assign_dictionary_model(oracle_lib):
    """Random inserts (a third of them re-inserts of live keys), removes and
    queries on 2^12 slots of each key-value slot type against a dict with
    insert-or-assign semantics.  Removes leave tombstones that keep counting
    in the fills, so the run stays well below the slot count (at most 1 024
    live keys, about 1 900 slots used)."""
    o = oracle_lib
    for rep, vmask in ((o.REP_KV64, (1 << 64) - 1), (o.REP_KV16, 0xFFFF), (o.REP_KEY64, 0)):
        rng = random.Random(6 + rep)
        nb, B = 256, 16
        t = table(o, rep, nb)
        model, live = {}, []
        for step in range(4000):
            r = rng.random()
            if r < 0.45 and len(model) < 1024:
                if live and rng.random() < 0.35:
                    k = rng.choice(live)
                else:
                    k = rng.getrandbits(64)
                    if rep == o.REP_KV16:
                        # KV16 stores the low 16 bits: keep them distinct
                        while fp16(k) in {fp16(x) for x in model} or k & 0xFFFF in (0, 0xFFFF):
                            k = rng.getrandbits(64)
                v = rng.getrandbits(64) & vmask
                assert t.insert(k, v) == 1, (rep, step)
                if k not in model:
                    live.append(k)
                model[k] = v
            elif r < 0.6 and live:
                k = live.pop(rng.randrange(len(live)))
                assert t.remove(k) == 1
                del model[k]
            else:
                k = rng.choice(live) if live and rng.random() < 0.7 else rng.getrandbits(64)
                found, v = t.query(k)
                if k in model:
                    assert found and v == model[k], (rep, step)
                elif rep != o.REP_KV16:
                    assert not found
        assert t.fill() == len(model)
        for k, v in model.items():
            assert t.query(k) == (True, v)
