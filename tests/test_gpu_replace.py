"""GPU parity of power_of_n_insert_shortcut_replace_scheme
(insert_schemes/power_of_n_shortcut_key_replace.cuh:39-523): the flat
shortcut scheme whose insert also takes the slot that holds the key and
replaces its value.

Deterministic mode: statuses, level bytes, fills and queries are bit-exact
with the oracle through inserts, re-inserts with new values, removes and
re-inserts after removes, for key_val_pair<u64,u64>, the 16/16 key-value
slot (with the flat TCF's bucket backing level) and key-only slots.
Concurrently: re-inserting every stored key with a new value takes no new
slot and every key then reads its new value, also through the serving
loop."""
import numpy as np
import pytest

import oracle
import poggers_amd as pg
from tests.gpu_util import Ops, pair

pytestmark = pytest.mark.gpu

M, D, R = pg.HASH_MURMUR64A, pg.PROBE_DOUBLE, pg.SCHEME_POWER_OF_N_SHORTCUT_REPLACE


def kv64(nslots, B=16, tile=4):
    return pg.TableDesc(pg.REP_KV64, [pg.Level(nslots, R, B, tile, D, M, 2)])


def key64(nslots, B=16, tile=4):
    return pg.TableDesc(pg.REP_KEY64, [pg.Level(nslots, R, B, tile, D, M, 2)])


def flat_replace(nslots, tile=4):
    d = pg.flat_tcf_desc(nslots, nslots // 32, tile=tile)
    d.levels[0].scheme = R
    return d


CASES = {
    "kv64_tile4": kv64(1 << 12),
    "kv64_tile1_b8": kv64(1 << 12, B=8, tile=1),
    "kv64_tile8_b8": kv64(1 << 12, B=8, tile=8),
    "kv16_flat_tile4": flat_replace(1 << 12),
    "kv16_flat_tile2": flat_replace(1 << 12, tile=2),
    "key64_tile4": key64(1 << 12),
    # a replace level backing a replace level: spilled keys are replaced there
    "kv64_two_replace_levels": pg.TableDesc(pg.REP_KV64, [pg.Level(1 << 12, R, 16, 4, D, M, 2),
                                                          pg.Level(1 << 9, R, 8, 4, D, M, 2)]),
}


def vals_for(desc, keys, salt):
    if desc.rep == pg.REP_KV64:
        return keys * np.uint64(5) + np.uint64(salt)
    return ((keys >> np.uint64(24)) + np.uint64(salt)).astype(np.uint64) & np.uint64(0xFFFF)


def same_state(o, g, desc, q, ops):
    for lvl in range(len(desc.levels)):
        assert g.level_bytes(lvl) == o.level_bytes(lvl), lvl
    assert g.get_fill() == o.fill()
    f_o, v_o = o.bulk_query(q)
    f_g, v_g = ops.query(q)
    assert (f_g == f_o).all() and (v_g == v_o).all()


@pytest.mark.parametrize("name", list(CASES))
def test_deterministic_bit_exact(name):
    desc = CASES[name]
    o, g = pair(desc, seed=42)
    g.set_mode(pg.DETERMINISTIC)
    ops = Ops(g)
    n = int(0.85 * desc.levels[0].nslots)
    keys = oracle.gen_keys(131, 0, n)
    q = np.concatenate([keys, oracle.gen_keys(131, 1 << 40, 2048)])
    steps = [
        (keys, 1),                                                   # fill to 85 %
        (keys[::2], 2),                                              # re-insert half with new values
        (np.concatenate([keys[1::4], oracle.gen_keys(132, 0, n // 10)]), 3),  # re-inserts and fresh keys
    ]
    for i, (ks, salt) in enumerate(steps):
        vs = vals_for(desc, ks, salt)
        st_o = o.bulk_insert(ks, vs.astype(o.val_dtype))
        st_g, fails = ops.insert(ks, vs.astype(np.uint64 if desc.rep == pg.REP_KV64 else np.uint16))
        assert (st_g == st_o).all() and fails == int((st_o == 0).sum()), i
        same_state(o, g, desc, q, ops)
    # removes leave tombstones; removed keys come back in fresh slots
    rm = keys[::5]
    assert (ops.remove(rm) == o.bulk_remove(rm)).all()
    back = keys[::10]
    vs = vals_for(desc, back, 4)
    st_o = o.bulk_insert(back, vs.astype(o.val_dtype))
    st_g, _ = ops.insert(back, vs.astype(np.uint64 if desc.rep == pg.REP_KV64 else np.uint16))
    assert (st_g == st_o).all()
    same_state(o, g, desc, q, ops)


@pytest.mark.parametrize("name", list(CASES))
def test_concurrent_reinsert_replaces_values(name):
    desc = CASES[name]
    g = pg.Table(desc, seed=42)
    ops = Ops(g)
    vdt = np.uint64 if desc.rep == pg.REP_KV64 else np.uint16
    n = int(0.8 * desc.levels[0].nslots)
    keys = oracle.gen_keys(133, 0, n)
    st, fails = ops.insert(keys, vals_for(desc, keys, 1).astype(vdt))
    ok = st > 0
    fill = g.get_fill()
    if desc.rep == pg.REP_KV16:
        # a 16/16 slot stores the key's low 16 bits: a key that meets an
        # earlier key with the same low bits in a probe bucket replaces that
        # key's value instead of taking a slot (the oracle does the same when
        # the earlier key lands first).  About 82 pairs of these 3 276 keys
        # share their low bits, and about 2 % of those share a probe bucket.
        assert 0 <= int(ok.sum()) - fill <= 8
    else:
        assert fill == int(ok.sum())
    kept = keys[ok]
    new = vals_for(desc, kept, 7).astype(vdt)
    st2, fails2 = ops.insert(kept, new)
    assert fails2 == 0 and (st2 > 0).all()
    # every key already held a slot: nothing new is taken
    assert g.get_fill() == fill
    f, v = ops.query(kept)
    assert f.all()
    if desc.rep == pg.REP_KV64:
        assert (v == new).all()
    elif desc.rep == pg.REP_KV16:
        # a 16-bit key can share its slot with another key of the same low bits
        assert (v == new).mean() > 0.99


def _drain(srv, tickets):
    done = []
    for _ in range(1_000_000):
        done += srv.progress()
        if set(tickets) <= set(done):
            return
    raise AssertionError("requests did not finish")


def test_server_serves_replace_tables():
    """A server holding a replace-scheme table (next to a plain one) serves
    its inserts with the replace kernel: a second insert of the same keys
    with new values takes no slot, and served queries read the new values."""
    import torch
    # the bucket_insert table stores every key twice: twice the slots
    rt, pt = pg.Table(kv64(1 << 14)), pg.Table(pg.kv_desc(1 << 15, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE,
                                                               max_probes=20))
    srv = pg.Server([rt, pt], budget_bytes=1 << 24, nstreams=2)
    n = 8000
    keys = oracle.gen_keys(141, 0, n)
    pin = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).view(dt)).pin_memory()
    kk = pin(keys, np.int64)
    flags = [torch.zeros(n, dtype=torch.uint8).pin_memory() for _ in range(4)]
    v1, v2 = pin(keys ^ np.uint64(0x1111), np.int64), pin(keys ^ np.uint64(0x2222), np.int64)
    for tab in (0, 1):
        _drain(srv, [srv.submit(tab, pg.OP_INSERT, kk, v1, None, flags[0])])
        assert flags[0].numpy().all(), tab
        _drain(srv, [srv.submit(tab, pg.OP_INSERT, kk, v2, None, flags[1])])
        assert flags[1].numpy().all(), tab
        out = torch.zeros(n, dtype=torch.int64).pin_memory()
        _drain(srv, [srv.submit(tab, pg.OP_QUERY, kk, None, out, flags[2])])
        assert flags[2].numpy().all()
        got = out.numpy().view(np.uint64)
        if tab == 0:     # replace: one slot per key, the new values
            assert rt.get_fill() == n and (got == (keys ^ np.uint64(0x2222))).all()
        else:            # bucket_insert: a second copy, the query finds the first
            assert pt.get_fill() == 2 * n and (got == (keys ^ np.uint64(0x1111))).all()
    srv.free()
