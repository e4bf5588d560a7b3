"""Pin the oracle's key-value schemes (bucket_insert / single_slot_insert).

Checks: the reference tests' known answers, Knuth's closed forms for linear
probing, and a dictionary model on small random workloads.
"""
import random

import numpy as np
import pytest


def kv(o, nslots, B=1, P=1, prober=None, hasher=None, max_probes=None, scheme=None):
    prober = o.PROBE_LINEAR if prober is None else prober
    hasher = o.HASH_FMIX64 if hasher is None else hasher
    scheme = o.SCHEME_BUCKET if scheme is None else scheme
    if max_probes is None:
        max_probes = (nslots + B - 1) // B if prober == o.PROBE_LINEAR else 20
    return o.Table(o.REP_KV64, [o.level(nslots, scheme, B, tile=P, prober=prober, hasher=hasher,
                                        max_probes=max_probes)], seed=42)


def test_atomic_swap_known_answer(oracle_lib):
    # tests/src/test_key_val_pair.cu:50-62: swap into an empty slot succeeds
    # (1), the second swap into the same slot fails (0).
    o = oracle_lib
    t = kv(o, 1, max_probes=1)
    assert t.insert(34, 0) == 1
    assert t.insert(34, 1) == 0
    assert t.query(34) == (True, 0)


def test_static_table_known_answer(oracle_lib):
    # tests/src/test_components.cu:345-366: bucket_insert, P=8, B=8,
    # doubleHasher with 20 probes, murmur; key = 39*id, val = 44*key; every
    # insert and query succeeds.  (id 0 -> key 0 is the reserved empty key
    # and is left out.)
    o = oracle_lib
    t = kv(o, 1_000_000, B=8, P=8, prober=o.PROBE_DOUBLE, hasher=o.HASH_MURMUR64A, max_probes=20)
    keys = np.arange(1, 100_001, dtype=np.uint64) * np.uint64(39)
    vals = keys * np.uint64(44)
    st = t.bulk_insert(keys, vals)
    assert (st == 1).all()
    found, got = t.bulk_query(keys)
    assert found.all() and (got == vals).all()


def test_linear_probing_closed_forms(oracle_lib):
    """Knuth: at load a, a successful search probes (1 + 1/(1-a))/2 slots and
    an unsuccessful one (1 + 1/(1-a)^2)/2 on average.  The unsuccessful cost
    of one table also follows exactly from its occupancy map: a query that
    starts at slot s walks to the next empty slot."""
    o = oracle_lib
    m = 1 << 20
    a = 0.85
    n = int(a * m)
    t = kv(o, m)
    keys = o.gen_keys(5, 0, n)
    assert (t.bulk_insert(keys, keys) == 1).all()
    t.stats_reset()
    found, vals = t.bulk_query(keys)
    assert found.all() and (vals == keys).all()
    hit = t.stats()["buckets_read_l0"] / n
    t.stats_reset()
    absent = o.gen_keys(5, 1 << 40, 1 << 18)
    found, _ = t.bulk_query(absent)
    assert not found.any()
    miss = t.stats()["buckets_read_l0"] / len(absent)
    succ, unsucc = 0.5 * (1 + 1 / (1 - a)), 0.5 * (1 + 1 / (1 - a) ** 2)
    assert abs(hit - succ) / succ < 0.03, hit
    assert abs(miss - unsucc) / unsucc < 0.04, miss
    occ = t.level_array(0).reshape(-1, 2)[:, 0] != 0
    empties = np.nonzero(~occ)[0]
    pos = np.searchsorted(empties, np.arange(m))
    nxt = np.where(pos < len(empties), empties[np.minimum(pos, len(empties) - 1)], empties[0] + m)
    exact = float(((nxt - np.arange(m)) + 1).mean())
    assert abs(miss - exact) / exact < 0.02, (miss, exact)


def test_linear_probe_positions(oracle_lib):
    # Hand-derived: with B=1 and linear probing, colliding keys occupy
    # consecutive slots h, h+1, ... in insert order (wrapping mod nslots).
    o = oracle_lib
    m = 16
    t = kv(o, m)
    rng = random.Random(1)
    target = 15
    ks = []
    while len(ks) < 3:
        k = rng.getrandbits(64)
        if k and o.hash64(o.HASH_FMIX64, k, 42) % m == target:
            ks.append(k)
    for i, k in enumerate(ks):
        assert t.insert(k, i + 1) == 1
    arr = t.level_array(0).reshape(-1, 2)
    assert [int(arr[s, 0]) for s in (15, 0, 1)] == ks
    assert [int(arr[s, 1]) for s in (15, 0, 1)] == [1, 2, 3]


@pytest.mark.parametrize("B,P,prober", [(1, 1, 0), (4, 2, 1), (8, 8, 1), (16, 4, 1), (2, 4, 3)])
def test_dict_model(oracle_lib, B, P, prober):
    o = oracle_lib
    rng = random.Random(B * 100 + P)
    m = 512
    t = kv(o, m, B=B, P=P, prober=prober, max_probes=(m // B if prober == 0 else (2 if prober == 3 else 20)))
    model = {}
    removed = set()
    for step in range(3000):
        r = rng.random()
        if r < 0.5 and len(model) < int(0.7 * m):
            k = rng.getrandbits(63) + 1
            v = rng.getrandbits(64)
            if k in model or k in removed:
                continue
            st = t.insert(k, v)
            if st:
                model[k] = v
        elif r < 0.75 and model:
            k = rng.choice(list(model))
            assert t.remove(k) == 1
            del model[k]
            removed.add(k)
        else:
            k = rng.choice(list(model)) if model and rng.random() < 0.7 else rng.getrandbits(63) + 1
            found, v = t.query(k)
            if k in model:
                assert found and v == model[k]
            else:
                assert not found
    for k, v in model.items():
        assert t.query(k) == (True, v)
    for k in removed:
        assert t.query(k)[0] is False
