"""Pin the oracle's recursive levels (tables/bucketed_table.cuh:226-381 with
Is_Recursive: an op that level i refuses goes to level i + 1, whose table is
generated with the square of level i's seed, bucketed_table.cuh:170) on
tables of three and four levels, by hand-derived placements: a key reaches
level 2 only when levels 0 and 1 refuse it, and lands in slot 0 of the first
bucket its level-2 probes give, computed here from seed^4."""
import numpy as np

SEED = 42
M64 = (1 << 64) - 1


def fp16(key):
    f = key & 0xFFFF
    return 1 if f == 0 else (0xFFFE if f == 0xFFFF else f)


def three_level(o, nb=(2, 2, 4), B=4):
    return o.Table(o.REP_FP16_V16, [
        o.level(nb[0] * B, o.SCHEME_SHORTCUT, B, prober=o.PROBE_DOUBLE, hasher=o.HASH_MURMUR64A, max_probes=2),
        o.level(nb[1] * B, o.SCHEME_LINEAR_BUCKETS, B, prober=o.PROBE_LINEAR, hasher=o.HASH_MURMUR64A,
                max_probes=nb[1]),
        o.level(nb[2] * B, o.SCHEME_LINEAR_BUCKETS, B, prober=o.PROBE_LINEAR, hasher=o.HASH_MURMUR64A,
                max_probes=nb[2]),
    ], seed=SEED)


def test_third_level_takes_what_the_first_two_refuse(oracle_lib):
    o = oracle_lib
    B = 4
    t = three_level(o)
    keys = o.gen_keys(101, 0, 40)
    vals = np.arange(40, dtype=np.uint16) + 7
    st = t.bulk_insert(keys, vals)
    # levels 0 and 1 hold 2*4 + 2*4 slots: the first 16 keys fit there
    assert (st[:16] > 0).all() and (st[:16] < 3).all()
    assert (st[16:32] == 3).all() and (st[32:] == 0).all()     # level 2 holds 16, the rest fail
    first = int(np.argmax(st == 3))
    k = int(keys[first])
    seed2 = (SEED * SEED) * (SEED * SEED) & M64             # the third level's seed
    p0 = o.probes(o.PROBE_LINEAR, o.HASH_MURMUR64A, 4, seed2, k, fp16(k))[0]
    words = t.level_array(2).reshape(4, B)
    assert int(words[p0 % 4, 0]) == (int(vals[first]) << 16) | fp16(k)
    assert t.fill() == 32
    f, v = t.bulk_query(keys)
    assert (f[:32] == 1).all() and (v[:32] == vals[:32]).all() and not f[32:].any()


def test_fourth_level_and_removes_across_levels(oracle_lib):
    o = oracle_lib
    B = 4
    lv = [o.level(2 * B, o.SCHEME_SHORTCUT, B, prober=o.PROBE_DOUBLE, hasher=o.HASH_MURMUR64A, max_probes=2)]
    for nb in (2, 2, 2):
        lv.append(o.level(nb * B, o.SCHEME_LINEAR_BUCKETS, B, prober=o.PROBE_LINEAR, hasher=o.HASH_MURMUR64A,
                          max_probes=nb))
    t = o.Table(o.REP_FP16_V16, lv, seed=SEED)
    keys = o.gen_keys(102, 0, 40)
    st = t.bulk_insert(keys)
    assert sorted(set(st.tolist())) == [0, 1, 2, 3, 4] and (st > 0).sum() == 32
    assert [int((st == lvl + 1).sum()) for lvl in range(4)] == [8, 8, 8, 8]
    # removing a level-3 key turns its slot into a tombstone there only
    k3 = keys[st == 4][:1]
    before = [t.level_array(lvl).copy() for lvl in range(4)]
    assert (t.bulk_remove(k3) == 1).all()
    after = [t.level_array(lvl) for lvl in range(4)]
    assert all((before[lvl] == after[lvl]).all() for lvl in range(3))
    assert int((before[3] != after[3]).sum()) == 1 and int(after[3][before[3] != after[3]][0]) == 0xFFFF
    f, _ = t.bulk_query(keys)
    assert (f == ((st > 0) & (keys != k3[0]))).all()
