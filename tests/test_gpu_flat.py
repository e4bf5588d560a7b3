"""GPU parity of the flat power-of-n schemes over key-value slots:

  * the reference's flat TCF (data_structs/tcf.cuh:29-30): 16/16 key_val_pair
    slots (REP_KV16), 16-slot buckets, power_of_n_insert_shortcut_scheme with
    two double-hashing probes, backed by tiny_static_table_4 (4-slot buckets,
    bucket_insert, 20 probes);
  * p2_table of tests/src/test_components.cu:54 (key_val_pair<u64,u64>,
    8-slot buckets, tile 8, power_of_n_insert_scheme, 3 double-hashing
    probes, seed 1);
  * the k-mer store's tier-two table (key_val_pair<u64,u64>, 8-slot buckets,
    tile 8, power_of_n_insert_scheme over the power-of-two prober).

Deterministic mode: statuses, level bytes, queries (found and values),
removes and re-inserts after removes are bit-exact with the oracle.
Concurrently: failures stay near the key-order count, every stored key is
found with its value and fill equals the stored count."""
import numpy as np
import pytest
import torch

import oracle
import poggers_amd as pg
from tests.gpu_util import Ops, pair

pytestmark = pytest.mark.gpu

M, D = pg.HASH_MURMUR64A, pg.PROBE_DOUBLE


def p2_table(nslots, tile=8):
    return pg.TableDesc(pg.REP_KV64, [pg.Level(nslots, pg.SCHEME_POWER_OF_N, 8, tile, D, M, 3)])


def tier_two(nslots):
    return pg.TableDesc(pg.REP_KV64, [pg.Level(nslots, pg.SCHEME_POWER_OF_N, 8, 8, pg.PROBE_P2, M, 2)])


def kv16_pow(nslots, tile=4, B=8, probes=4):
    return pg.TableDesc(pg.REP_KV16, [pg.Level(nslots, pg.SCHEME_POWER_OF_N, B, tile, D, M, probes)])


def key_only_shortcut(nslots, tile=2):
    return pg.TableDesc(pg.REP_KEY64, [pg.Level(nslots, pg.SCHEME_POWER_OF_N_SHORTCUT_SLOTS, 16, tile, D, M, 2)])


# (descriptor, seed, keys as a fraction of level 0's slots)
CASES = {
    "flat_tcf": (pg.flat_tcf_desc(1 << 12, 1 << 7), 42, 1.05),
    "flat_tcf_tile16": (pg.flat_tcf_desc(1 << 12, 1 << 7, tile=16), 42, 1.03),
    "flat_tcf_tile2": (pg.flat_tcf_desc(1 << 12, 1 << 7, tile=2), 42, 1.04),
    "flat_tcf_tile1": (pg.flat_tcf_desc(1 << 12, 1 << 7, tile=1), 42, 1.02),
    "p2_table": (p2_table(1 << 12), 1, 0.97),
    "p2_table_tile2": (p2_table(1 << 12, tile=2), 1, 0.97),
    "tier_two": (tier_two(1 << 12), 42, 0.95),
    "kv16_pow4": (kv16_pow(1 << 12), 42, 0.99),
    "kv16_pow4_tile2_b16": (kv16_pow(1 << 12, tile=2, B=16), 42, 0.99),
    "key64_shortcut": (key_only_shortcut(1 << 12), 42, 0.98),
}


def vals_for(desc, keys):
    if desc.rep == pg.REP_KV64:
        return keys * np.uint64(5) + np.uint64(1)
    return ((keys >> np.uint64(24)) & np.uint64(0xFFFF)).astype(np.uint16)


@pytest.mark.parametrize("name", list(CASES))
def test_deterministic_bit_exact(name):
    desc, seed, over = CASES[name]
    o, g = pair(desc, seed=seed)
    g.set_mode(pg.DETERMINISTIC)
    ops = Ops(g)
    n = int(over * desc.levels[0].nslots)
    keys = oracle.gen_keys(121, 0, n)
    vals = vals_for(desc, keys)
    st_o = o.bulk_insert(keys, vals)
    st_g, fails = ops.insert(keys, vals)
    assert (st_g == st_o).all() and fails == int((st_o == 0).sum())
    if len(desc.levels) > 1:
        assert (st_o == 2).any(), "some keys must reach the backing level"
    for lvl in range(len(desc.levels)):
        assert g.level_bytes(lvl) == o.level_bytes(lvl), lvl
    assert g.get_fill() == o.fill()
    q = np.concatenate([keys, oracle.gen_keys(121, 1 << 40, 4096)])
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all() and (v_g == v_o).all()
    # removes leave tombstones that keep counting toward the fill: the
    # re-inserts then place keys differently from a table without them
    rm = keys[::3]
    assert (ops.remove(rm) == o.bulk_remove(rm)).all()
    more = oracle.gen_keys(122, 0, n // 8)
    st_o = o.bulk_insert(more, vals_for(desc, more))
    st_g, _ = ops.insert(more, vals_for(desc, more))
    assert (st_g == st_o).all()
    for lvl in range(len(desc.levels)):
        assert g.level_bytes(lvl) == o.level_bytes(lvl), lvl
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all() and (v_g == v_o).all()
    assert g.get_fill() == o.fill()


@pytest.mark.parametrize("name", list(CASES))
def test_concurrent_stored_keys_found(name):
    desc, seed, over = CASES[name]
    g = pg.Table(desc, seed=seed)
    ops = Ops(g)
    n = int(min(over, 0.9) * desc.levels[0].nslots)
    keys = oracle.gen_keys(123, 0, n)
    vals = vals_for(desc, keys)
    st, fails = ops.insert(keys, vals)
    ok = st > 0
    assert fails == int((~ok).sum()) and g.get_fill() == int(ok.sum())
    # the same keys in key order.  Concurrent claims obey the fill rule when
    # they land, but with every key of a small table in flight they land in
    # another order than key order, which costs a few tenths of a percent
    # more failures on a single level (a CPU model of the concurrent insert
    # shows the same); stale-fill herding, which the bounds prevent, loses
    # several percent
    o = oracle.Table(desc.rep, [lv.as_dict() for lv in desc.levels], seed=seed)
    fails_o = int((o.bulk_insert(keys, vals) == 0).sum())
    assert fails <= 2 * fails_o + 0.01 * n, (fails, fails_o)
    f, v = ops.query(keys)
    assert f[ok].all()
    if desc.rep != pg.REP_KEY64:
        # a 16-bit key can match an earlier key's slot: compare values only
        # where the oracle, given the same placement, has no collision
        if desc.rep == pg.REP_KV64:
            assert (v[ok] == vals[ok]).all()
        else:
            assert (v[ok] == vals[ok]).mean() > 0.99
    removed = ops.remove(keys[ok])
    # each successful remove tombstones one live slot
    assert g.get_fill() == int(ok.sum()) - int(removed.sum())
    if desc.rep == pg.REP_KV16:
        # a 16-bit key can remove a colliding key's slot, whose own remove
        # then finds nothing when its buckets differ
        assert removed.mean() > 0.999
    else:
        assert removed.all()


def test_reference_p2_table_known_answers():
    """tests/src/test_components.cu:88-200 on the GPU, in both modes."""
    for mode in (pg.DETERMINISTIC, pg.CONCURRENT):
        g = pg.Table(p2_table(1100), seed=1)
        g.set_mode(mode)
        ops = Ops(g)
        for k, v in ((1, 1), (2, 2), (3, 4)):
            st, _ = ops.insert(np.array([k], np.uint64), np.array([v], np.uint64))
            assert st[0] == 1
            f, got = ops.query(np.array([k], np.uint64))
            assert f[0] == 1 and got[0] == v
        for i in range(1, 11):
            g = pg.Table(p2_table(i), seed=1)
            g.set_mode(mode)
            ops = Ops(g)
            keys = np.arange(1, i + 1, dtype=np.uint64)
            st, _ = ops.insert(keys, keys)
            assert (st == 1).all()
            f, v = ops.query(keys)
            assert f.all() and (v == keys).all()


def fill_order_violations(words, buckets, keys, vals):
    """Keys at slot index i >= ceil(3B/4) of one probe bucket whose other
    probe bucket ends with fill below i.  Such a slot is only claimed in the
    fill-ordered pass, when the key's bucket is no fuller than the other one;
    fills only grow, so the other bucket ends at least that full.  Returns
    (violations, keys checked)."""
    fills = ((words & 0xFFFF) != 0).sum(1)
    fp = keys & np.uint64(0xFFFF)
    fp = np.where(fp == 0, 1, np.where(fp == 0xFFFF, 0xFFFE, fp)).astype(np.uint32)
    want = (vals.astype(np.uint32) << 16) | fp
    bad = checked = 0
    for x in range(2):
        hit = words[buckets[:, x]] == want[:, None]
        idx = hit.argmax(1)
        sel = hit.any(1) & (idx >= 12) & (buckets[:, 0] != buckets[:, 1])
        checked += int(sel.sum())
        bad += int((sel & (fills[buckets[:, 1 - x]] < idx)).sum())
    return bad, checked


def test_flat_shortcut_claims_follow_the_fill_rule_concurrent():
    """Concurrent inserts into the flat TCF: the shortcut claims only below
    ceil(3B/4) and the fill-ordered pass claims a bucket only while it is no
    fuller than the key's other bucket.  Every key in an upper slot therefore
    has an other bucket at least that full, as in key order (where it holds
    by construction; a CPU model of the concurrent insert without the bound
    breaks it for about 10 % of the upper-slot keys).  The bucket-fill
    histogram itself is not compared: with every key in flight the claims
    land in a different order from key order, which spreads the fills
    (the same CPU model reproduces the spread)."""
    desc = pg.flat_tcf_desc(1 << 16, 1 << 10)
    n = int(0.9 * (1 << 16))
    keys = oracle.gen_keys(124, 0, n)
    vals = vals_for(desc, keys)
    o = oracle.Table(desc.rep, [lv.as_dict() for lv in desc.levels], seed=42)
    st_o = o.bulk_insert(keys, vals)
    g = pg.Table(desc)
    st_g, _ = Ops(g).insert(keys, vals)
    assert abs(int((st_g == 1).sum()) - int((st_o == 1).sum())) <= 0.01 * n
    nb = (1 << 16) // 16
    buckets = np.array([[p % nb for p in oracle.probes(oracle.PROBE_DOUBLE, oracle.HASH_MURMUR64A, 2, 42, int(k), 0)]
                        for k in keys])
    w_o = np.frombuffer(o.level_bytes(0), np.uint32).reshape(-1, 16)
    w_g = np.frombuffer(g.level_bytes(0), np.uint32).reshape(-1, 16)
    bad_o, n_o = fill_order_violations(w_o, buckets, keys, vals)
    bad_g, n_g = fill_order_violations(w_g, buckets, keys, vals)
    assert bad_o == 0 and n_o > 0.1 * n
    assert bad_g == 0 and n_g > 0.1 * n, (bad_g, n_g)


@pytest.mark.parametrize("name,nslots,load", [("p2_table", 1 << 20, 0.93), ("flat_tcf", 1 << 20, 0.97),
                                             ("flat_tcf_tile2", 1 << 20, 0.97)])
def test_concurrent_large_failures_near_key_order(name, nslots, load):
    """At 2^20 slots every key of a launch reads its fills before most
    claims land; the bounded claims keep failures near the key-order count."""
    tile = 2 if name.endswith("tile2") else 4
    desc = p2_table(nslots) if name == "p2_table" else pg.flat_tcf_desc(nslots, nslots // 64, tile=tile)
    seed = 1 if name == "p2_table" else 42
    n = int(load * nslots)
    keys = oracle.gen_keys(125, 0, n)
    vals = vals_for(desc, keys)
    o = oracle.Table(desc.rep, [lv.as_dict() for lv in desc.levels], seed=seed)
    fails_o = int((o.bulk_insert(keys, vals) == 0).sum())
    g = pg.Table(desc, seed=seed)
    st, fails = Ops(g).insert(keys, vals)
    assert fails == int((st == 0).sum()) and g.get_fill() == int((st > 0).sum())
    assert fails <= 2 * fails_o + 0.0005 * n, (fails, fails_o)


# The flat TCF filled past its primary level, so its bucket_insert backing
# level takes keys (bucket_insert.cuh:127-357).
FLAT = {
    "flat_tile4": pg.flat_tcf_desc(1 << 14, 1 << 9),
    "flat_tile2": pg.flat_tcf_desc(1 << 14, 1 << 9, tile=2),
    "flat_tile1": pg.flat_tcf_desc(1 << 14, 1 << 9, tile=1),
}


def kv16_vals(keys, salt):
    return (((keys >> np.uint64(24)) + np.uint64(salt)) & np.uint64(0xFFFF)).astype(np.uint16)


@pytest.mark.parametrize("name", list(FLAT))
def test_flat_tcf_answers_equal_reference_over_same_slots(name, tmp_path):
    """Concurrent inserts past the primary level's slots, then concurrent
    removes: found flags and values equal the oracle's query over the same
    slot bytes.  From that state, removes and inserts in key order are
    bit-exact with the oracle.  A saved table answers the same after loading,
    and a cleared one finds nothing."""
    d = FLAT[name]
    g = pg.Table(d, seed=42)
    ops = Ops(g)
    keys = oracle.gen_keys(331, 0, int(0.97 * (d.levels[0].nslots + d.levels[1].nslots)))
    st, _ = ops.insert(keys, kv16_vals(keys, 1))
    assert (st == 2).sum() > 0, "no key reached the backing level: the case tests nothing"
    ops.remove(keys[::3])
    q = np.concatenate([keys, oracle.gen_keys(331, 1 << 41, 1 << 14)])
    o = oracle.Table(d.rep, [lv.as_dict() for lv in d.levels], seed=42)
    for lvl in range(2):
        o.set_level_bytes(lvl, g.level_bytes(lvl))
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all()
    assert (v_g[f_g > 0] == v_o[f_o > 0].astype(v_g.dtype)).all()
    g.set_mode(pg.DETERMINISTIC)
    rm = np.concatenate([q[::7], q[-1000:]])
    assert (ops.remove(rm) == o.bulk_remove(rm)).all()
    again = oracle.gen_keys(332, 0, 3000)
    va = kv16_vals(again, 2)
    st_g, _ = ops.insert(again, va)
    assert (st_g == o.bulk_insert(again, va.astype(o.val_dtype))).all()
    for lvl in range(2):
        assert g.level_bytes(lvl) == o.level_bytes(lvl), lvl
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all() and (v_g == v_o.astype(v_g.dtype)).all()
    path = str(tmp_path / "flat.pgt")
    g.save(path)
    t2 = pg.Table.load(path)
    f2, v2 = Ops(t2).query(q)
    assert (f2 == f_g).all() and (v2 == v_g).all()
    t2.clear()
    torch.cuda.synchronize()
    f3, _ = Ops(t2).query(q)
    assert not f3.any()
