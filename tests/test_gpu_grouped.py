"""GPU parity of the 12-bit sub-word bucket (REP_FP12) and of cache-line-
grouped levels (SCHEME_GROUPED, blocked_bucket_insert<K>) against the oracle.

Bar: in deterministic mode the raw level bytes (slot words and the padding
of each 128-byte line), statuses, query results and fill are bit-exact with
the oracle's; removes, tombstone reuse and upserts too where the format has
them.  Concurrently: every stored key is found and fill equals the number
of stored keys.  The compositions are the reference's bucket benchmarks
(tests/src/bucket_speed_tests.cu:82-102) plus grouped 16/16, 8/8 and 10/6
levels."""
import ctypes

import numpy as np
import pytest
import torch

import oracle
import poggers_amd as pg
from tests.gpu_util import Ops, pair

pytestmark = pytest.mark.gpu

D, M = pg.PROBE_DOUBLE, pg.HASH_MURMUR64A
# (name, rep, tile, bucket, scheme, group, backing?)
CASES = [
    ("tcqf_twelve", pg.REP_FP12, 4, 16, pg.SCHEME_SHORTCUT, 0, False),
    ("tcqf_twelve_16", pg.REP_FP12, 4, 16, pg.SCHEME_GROUPED, 5, False),
    ("tcqf_twelve_8", pg.REP_FP12, 2, 8, pg.SCHEME_GROUPED, 10, False),
    ("tcqf_twelve_12", pg.REP_FP12, 4, 12, pg.SCHEME_GROUPED, 6, False),
    ("tcqf_twelve_32", pg.REP_FP12, 8, 32, pg.SCHEME_GROUPED, 2, False),
    ("grouped16_backed", pg.REP_FP16_V16, 4, 8, pg.SCHEME_GROUPED, 3, True),
    ("grouped88", pg.REP_FP8_V8, 4, 16, pg.SCHEME_GROUPED, 4, True),
    ("grouped106", pg.REP_FP10_V6, 1, 8, pg.SCHEME_GROUPED, 7, True),
]
IDS = [c[0] for c in CASES]


def desc_of(rep, tile, B, scheme, group, backed, nslots=1 << 12):
    levels = [pg.Level(nslots, scheme, B, tile, D, M, 2, group)]
    if backed:
        levels.append(pg.Level(nslots // 16, pg.SCHEME_LINEAR_BUCKETS, 8, tile, D, M, 20))
    return pg.TableDesc(rep, levels)


def vals_for(keys, rep):
    v = ((keys >> np.uint64(20)) & np.uint64(0xFFFF)).astype(np.uint16)
    return v.astype(np.uint8) if rep in (pg.REP_FP8_V8, pg.REP_FP10_V6) else v


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_deterministic_bit_exact(case):
    _, rep, tile, B, scheme, group, backed = case
    desc = desc_of(rep, tile, B, scheme, group, backed)
    o, g = pair(desc)
    g.set_mode(pg.DETERMINISTIC)
    ops = Ops(g)
    n = (1 << 12) + (1 << 8) if backed else (1 << 12)
    keys = oracle.gen_keys(21, 0, n)  # overfill: shortcut pass 2, spill, failures
    vals = vals_for(keys, rep)
    st_o = o.bulk_insert(keys, vals)
    st_g, fails = ops.insert(keys, vals)
    assert (st_g == st_o).all()
    assert fails == int((st_o == 0).sum())
    for lvl in range(len(desc.levels)):
        assert g.level_bytes(lvl) == o.level_bytes(lvl), f"level {lvl} bytes differ"
    assert g.get_fill() == o.fill()
    q = np.concatenate([keys, oracle.gen_keys(21, 1 << 40, 4096)])
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all() and (v_g.astype(np.uint64) == v_o.astype(np.uint64)).all()
    if rep == pg.REP_FP12:
        return
    rm = keys[::3]
    assert (ops.remove(rm) == o.bulk_remove(rm)).all()
    more = oracle.gen_keys(22, 0, 1024)
    st_o = o.bulk_insert(more, vals_for(more, rep), with_delete=True)
    st_g, _ = ops.insert(more, vals_for(more, rep), with_delete=True)
    assert (st_g == st_o).all()
    for lvl in range(len(desc.levels)):
        assert g.level_bytes(lvl) == o.level_bytes(lvl)
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all() and (v_g.astype(np.uint64) == v_o.astype(np.uint64)).all()


def test_grouped_upsert_bit_exact():
    desc = desc_of(pg.REP_FP16_V16, 4, 8, pg.SCHEME_GROUPED, 3, True)
    for delete in (False, True):
        o, g = pair(desc)
        g.set_mode(pg.DETERMINISTIC)
        ops = Ops(g)
        keys = oracle.gen_keys(23, 0, 3000)
        seq = np.concatenate([keys, keys[::2]])
        np.random.default_rng(2).shuffle(seq)
        vals = vals_for(seq, pg.REP_FP16_V16)
        st_o, fo_o, vo_o = o.bulk_insert_if_not_exists(seq, vals, delete=delete)
        st_g, fo_g, vo_g = ops.upsert(seq, vals, delete=delete)
        assert (st_g == st_o).all() and (fo_g == fo_o).all() and (vo_g == vo_o).all()
        for lvl in range(2):
            assert g.level_bytes(lvl) == o.level_bytes(lvl)


def test_twelve_bit_ops_without_reference_counterpart_are_refused():
    g = pg.Table(desc_of(pg.REP_FP12, 4, 16, pg.SCHEME_SHORTCUT, 0, False))
    keys = torch.zeros(16, dtype=torch.int64, device="cuda")
    out = torch.zeros(16, dtype=torch.uint8, device="cuda")
    L = pg.lib()
    assert L.pg_bulk_remove(g._h, keys.data_ptr(), 16, out.data_ptr(), None) == 2
    assert L.pg_bulk_insert_with_delete(g._h, keys.data_ptr(), None, 16, None, None, None) == 2
    assert L.pg_bulk_insert_if_not_exists(g._h, keys.data_ptr(), None, 16, None, out.data_ptr(), None,
                                          None, None) == 2
    torch.cuda.synchronize()


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_concurrent_stored_keys_found(case):
    _, rep, tile, B, scheme, group, backed = case
    nslots = 1 << 20
    desc = desc_of(rep, tile, B, scheme, group, backed, nslots=nslots)
    g = pg.Table(desc)
    ops = Ops(g)
    n = int(0.85 * nslots)
    keys = oracle.gen_keys(24, 0, n)
    vals = vals_for(keys, rep)
    st = np.zeros(n, dtype=np.uint8)
    step = (n + 63) // 64
    for a in range(0, n, step):
        st[a:a + step], _ = ops.insert(keys[a:a + step], vals[a:a + step])
    ok = st > 0
    assert ok.mean() > 0.97  # single-level 8-slot shapes drop ~0.6 % even in key order
    found, _ = ops.query(keys)
    if rep == pg.REP_FP12:
        # every slot holds a nonzero composite and there are no removes
        assert found[ok].all()
        assert g.get_fill() == int(ok.sum())
    elif rep == pg.REP_FP10_V6:
        low = (keys & np.uint64(0xFFFF)).astype(np.int64)
        low[low == 0] = 1
        low[low == 0xFFFF] = 0xFFFE
        normal = ((low & 0x3FF) != 0) & ((low & 0x3FF) != 0x3FF)
        assert found[(st == 1) & normal].all()
        assert g.get_fill() == int((ok & normal).sum())
    else:
        assert found[st == 1].all()
        assert g.get_fill() == int(ok.sum())
