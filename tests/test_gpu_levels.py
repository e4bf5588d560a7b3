"""GPU parity of tables with more than one backing level (recursive
compositions, tables/bucketed_table.cuh:226-381; each level seeded with the
square of the one above, :170): three- and four-level filters and a
three-level key-value table, overfilled so keys reach the last level.

Deterministic mode: statuses, every level's bytes, queries, removes and
tombstone-reusing re-inserts are bit-exact with the oracle.  Concurrently:
every stored key is found and fill equals the stored count."""
import numpy as np
import pytest

import oracle
import poggers_amd as pg
from tests.gpu_util import Ops, pair

pytestmark = pytest.mark.gpu

M, D, L = pg.HASH_MURMUR64A, pg.PROBE_DOUBLE, pg.PROBE_LINEAR


def filter_desc(nslots, backing, tile=4, B=16):
    lv = [pg.Level(nslots, pg.SCHEME_SHORTCUT, B, tile, pg.PROBE_XOR_P2, M, 2)]
    for ns in backing:
        lv.append(pg.Level(ns, pg.SCHEME_LINEAR_BUCKETS, 8, tile, L, M, max(1, ns // 8)))
    return pg.TableDesc(pg.REP_FP16_V16, lv)


def kv_desc3(nslots, tile=4):
    return pg.TableDesc(pg.REP_KV64, [pg.Level(nslots, pg.SCHEME_BUCKET, 8, tile, D, M, 4),
                                      pg.Level(nslots // 8, pg.SCHEME_BUCKET, 8, tile, D, M, 8),
                                      pg.Level(nslots // 16, pg.SCHEME_BUCKET, 4, tile, L, M, nslots // 64)])


CASES = {
    "filter3": (filter_desc(1 << 12, [1 << 7, 1 << 7]), 1.06),
    "filter4": (filter_desc(1 << 12, [64, 64, 1 << 7], tile=2), 1.07),
    "kv3": (kv_desc3(1 << 12), 1.12),
}


def vals_for(keys, kv):
    return keys * np.uint64(3) if kv else ((keys >> np.uint64(30)) & np.uint64(0xFFFF)).astype(np.uint16)


@pytest.mark.parametrize("name", list(CASES))
def test_deterministic_bit_exact(name):
    desc, over = CASES[name]
    o, g = pair(desc)
    g.set_mode(pg.DETERMINISTIC)
    ops = Ops(g)
    kv = desc.rep == pg.REP_KV64
    n = int(over * desc.levels[0].nslots)
    keys = oracle.gen_keys(111, 0, n)
    vals = vals_for(keys, kv)
    st_o = o.bulk_insert(keys, vals)
    st_g, fails = ops.insert(keys, vals)
    assert (st_g == st_o).all() and fails == int((st_o == 0).sum())
    assert (st_o == len(desc.levels)).any(), "some keys must reach the last level"
    for lvl in range(len(desc.levels)):
        assert g.level_bytes(lvl) == o.level_bytes(lvl), lvl
    assert g.get_fill() == o.fill()
    q = np.concatenate([keys, oracle.gen_keys(111, 1 << 40, 2048)])
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all() and (v_g == v_o).all()
    rm = keys[::5]
    assert (ops.remove(rm) == o.bulk_remove(rm)).all()
    more = oracle.gen_keys(112, 0, n // 10)
    if kv:
        st_o = o.bulk_insert(more, vals_for(more, kv))
        st_g, _ = ops.insert(more, vals_for(more, kv))
    else:
        st_o = o.bulk_insert(more, vals_for(more, kv), with_delete=True)
        st_g, _ = ops.insert(more, vals_for(more, kv), with_delete=True)
    assert (st_g == st_o).all()
    for lvl in range(len(desc.levels)):
        assert g.level_bytes(lvl) == o.level_bytes(lvl), lvl
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all() and (v_g == v_o).all()


@pytest.mark.parametrize("name", list(CASES))
def test_concurrent_stored_keys_found(name):
    desc, over = CASES[name]
    g = pg.Table(desc)
    ops = Ops(g)
    kv = desc.rep == pg.REP_KV64
    n = int(0.97 * desc.levels[0].nslots)
    keys = oracle.gen_keys(113, 0, n)
    st, fails = ops.insert(keys, vals_for(keys, kv))
    ok = st > 0
    assert fails == int((~ok).sum()) and g.get_fill() == int(ok.sum())
    f, v = ops.query(keys)
    assert f[ok].all()
    if kv:
        assert (v[ok] == vals_for(keys, kv)[ok]).all()
