"""The occupancy index of linear-bucket backing levels (16/16 and 8/8 filter
slots): one bit per bucket, set before any slot of the bucket is claimed and
cleared only with the slots.  Queries and removes stop at a bucket whose bit is
clear without reading it; the reference's linear walk stops at the first
bucket that is not full (linear_insert_buckets.cuh:323-398), and an empty
bucket is not full and holds no match.  The index must never change an
answer:

* after concurrent inserts, removes and tombstone-reusing inserts (a layout
  that depends on timing), the concurrent query answers equal the oracle's
  sequential query over the same slot bytes;
* deterministic removes and inserts on that state stay bit-exact with the
  oracle (statuses and slot bytes);
* a saved table loads with its index rebuilt from the slots, pg_clear clears
  it, and a slot image written through pg_level_ptr is synced;
* with nothing in the backing level, absent-key queries examine no backing
  bucket (10/6 slots, which keep no index, examine one each).
10/6 compositions run the same checks without an index."""
import numpy as np
import pytest
import torch

import oracle
import poggers_amd as pg
from tests.gpu_util import Ops, pair

pytestmark = pytest.mark.gpu

F16, F8, F10 = pg.REP_FP16_V16, pg.REP_FP8_V8, pg.REP_FP10_V6


def desc(rep, tile, bb, prober=pg.PROBE_LINEAR, probes=4000, nslots=1 << 14, backing=1 << 9):
    return pg.tcf_desc(nslots, backing, bucket_size=16, tile=tile, prober=pg.PROBE_DOUBLE, rep=rep,
                       backing_bucket=bb, backing_prober=prober, backing_probes=probes)


CASES = {
    "fp16_b8_tile4": desc(F16, 4, 8),
    "fp16_b16_tile1": desc(F16, 1, 16),
    "fp16_b4_tile16_double": desc(F16, 16, 4, pg.PROBE_DOUBLE, 20),
    "fp8_b8_tile4": desc(F8, 4, 8),
    "fp8_b16_tile2": desc(F8, 2, 16),
    "fp10_b8_tile4": desc(F10, 4, 8),
}


def vals_for(d, keys, salt):
    v = ((keys >> np.uint64(20)) + np.uint64(salt)) & np.uint64(0xFFFF)
    return (v & np.uint64(0xFF)).astype(np.uint8) if d.rep in (F8, F10) else v.astype(np.uint16)


def churned(d, seed):
    """A table filled past its primary level with concurrent inserts, a third
    of its keys removed, then tombstone-reusing inserts: keys in the backing
    level, and backing buckets holding tombstones."""
    g = pg.Table(d, seed=42)
    ops = Ops(g)
    total = d.levels[0].nslots + d.levels[1].nslots
    keys = oracle.gen_keys(seed, 0, int(0.97 * total))
    ops.insert(keys, vals_for(d, keys, 1))
    rm = keys[::3]
    ops.remove(rm)
    more = oracle.gen_keys(seed, 1 << 40, len(keys) // 6)
    ops.insert(more, vals_for(d, more, 2), with_delete=True)
    q = np.concatenate([keys, more, oracle.gen_keys(seed, 1 << 41, 1 << 14)])
    return g, ops, q


def oracle_on(d, g):
    o = oracle.Table(d.rep, [lv.as_dict() for lv in d.levels], seed=42)
    for lvl in range(len(d.levels)):
        o.set_level_bytes(lvl, g.level_bytes(lvl))
    return o


@pytest.mark.parametrize("name", list(CASES))
def test_concurrent_answers_equal_reference_over_same_slots(name):
    d = CASES[name]
    g, ops, q = churned(d, 301)
    backing = ops.level(1)
    kmask = 0xFFFF if d.rep == F16 else (0xFF if d.rep == F8 else 0x3FF)
    assert ((backing & kmask) != 0).sum() > 0, "the backing level holds nothing: the case tests nothing"
    o = oracle_on(d, g)
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all()
    assert (v_g[f_g > 0] == v_o[f_o > 0].astype(v_g.dtype)).all()
    # removes and inserts in key order stay bit-exact from this state
    g.set_mode(pg.DETERMINISTIC)
    rm = np.concatenate([q[::7], q[-1000:]])
    assert (ops.remove(rm) == o.bulk_remove(rm)).all()
    again = np.concatenate([rm[::2], oracle.gen_keys(302, 0, 2000)])
    va = vals_for(d, again, 3)
    st_g, _ = ops.insert(again, va, with_delete=True)
    st_o = o.bulk_insert(again, va, with_delete=True)
    assert (st_g == st_o).all()
    for lvl in range(2):
        assert g.level_bytes(lvl) == o.level_bytes(lvl), lvl
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all() and (v_g == v_o.astype(v_g.dtype)).all()


@pytest.mark.parametrize("name", ["fp16_b8_tile4", "fp8_b16_tile2", "fp10_b8_tile4"])
def test_load_clear_and_upload_keep_the_index(name, tmp_path):
    d = CASES[name]
    g, ops, q = churned(d, 311)
    f, v = ops.query(q)
    path = str(tmp_path / "t.pgt")
    g.save(path)
    t2 = pg.Table.load(path)
    f2, v2 = Ops(t2).query(q)
    assert (f2 == f).all() and (v2 == v).all()
    # a slot image written through pg_level_ptr (level_upload syncs)
    t3 = pg.Table(d, seed=42)
    for lvl in range(2):
        t3.level_upload(lvl, g.level_bytes(lvl))
    f3, v3 = Ops(t3).query(q)
    assert (f3 == f).all() and (v3 == v).all()
    # clear empties slots and index; the table then fills again
    t2.clear()
    torch.cuda.synchronize()
    assert t2.get_fill() == 0
    f4, _ = Ops(t2).query(q)
    # only a 10/6 key whose stored bits are 0 matches an empty word
    low = q & np.uint64(0xFFFF)
    fp = np.where(low == 0, 1, np.where(low == 0xFFFF, 0xFFFE, low)).astype(np.uint64)
    assert (f4 == ((d.rep == F10) & ((fp & np.uint64(0x3FF)) == 0))).all()
    keys = q[: len(q) // 2]
    st, _ = Ops(t2).insert(keys, vals_for(d, keys, 4))
    f5, _ = Ops(t2).query(keys[st > 0])
    # equal to the reference's query over the same slot bytes; 10/6 slots may
    # hide a stored key (a key packing to the empty word, or an empty-looking
    # word in a full linear bucket, where the walk stops)
    f_o, _ = oracle_on(d, t2).bulk_query(keys[st > 0])
    assert (f5 == f_o).all()
    assert f5.all() or d.rep == F10


@pytest.mark.parametrize("rep", [F16, F8, F10])
def test_absent_queries_skip_an_empty_backing_level(rep):
    """Half full, filled in key order: no key leaves the primary level (as
    in the oracle), so the backing level is empty.  (With every key of this
    small table in one concurrent launch, a bounded claim may still pass a
    key on.)  Absent keys that reach the backing level examine no backing
    bucket where the index exists, one bucket each where it does not (10/6)."""
    d = desc(rep, 4, 8, nslots=1 << 16, backing=1 << 10)
    o, g = pair(d)
    g.set_mode(pg.DETERMINISTIC)
    ops = Ops(g)
    keys = oracle.gen_keys(321, 0, 1 << 15)
    vs = vals_for(d, keys, 1)
    st, _ = ops.insert(keys, vs)
    assert (st == o.bulk_insert(keys, vs)).all()
    assert (st == 1).all(), "a key left the primary level"
    absent = oracle.gen_keys(321, 1 << 42, 1 << 14)
    stats = torch.zeros(3, dtype=torch.int64, device="cuda")
    g.set_stats(stats)
    f, _ = ops.query(absent)
    g.set_stats(None)
    reached = int((f == 0).sum())   # the keys the primary level did not answer
    assert reached > 0.9 * len(absent)
    backing_read = int(stats[1].item())
    assert backing_read == (reached if rep == F10 else 0), (backing_read, reached)


def test_a_linear_primary_level_keeps_no_index():
    """Only backing levels keep the index: on a primary level every insert
    would pay its atomic OR.  An empty one-level linear 16/16 filter
    therefore reads the first probe bucket of every absent key (one bucket
    each: it is empty, so the walk stops there), where an indexed level
    would read none."""
    d = pg.TableDesc(F16, [pg.Level(1 << 12, pg.SCHEME_LINEAR_BUCKETS, 8, 4, pg.PROBE_LINEAR, pg.HASH_FMIX64,
                                    4000)])
    g = pg.Table(d, seed=42)
    absent = oracle.gen_keys(341, 0, 1 << 12)
    stats = torch.zeros(3, dtype=torch.int64, device="cuda")
    g.set_stats(stats)
    f, _ = Ops(g).query(absent)
    g.set_stats(None)
    assert not f.any()
    assert int(stats[0].item()) == len(absent)
