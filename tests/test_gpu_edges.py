"""Empty and ragged inputs across the entry points the other GPU tests cover
at ordinary sizes: zero-key calls on key-value, key-only and flat TCF tables, calls
without optional outputs, batches whose CSR offsets give some tables (first,
middle, last, or all) no keys, and served batches that mix zero-key requests
with others."""
import numpy as np
import pytest
import torch

import oracle
import poggers_amd as pg
from tests.gpu_util import Ops, dev_u64

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rep", [pg.REP_KV64, pg.REP_KEY64])
def test_zero_key_calls_and_missing_outputs(rep):
    levels = [pg.Level(1 << 12, pg.SCHEME_BUCKET, 8, 4, pg.PROBE_DOUBLE, pg.HASH_MURMUR64A, 20)]
    desc = pg.TableDesc(rep, levels)
    g = pg.Table(desc)
    ops = Ops(g)
    empty = np.zeros(0, dtype=np.uint64)
    st, fails = ops.insert(empty)
    assert len(st) == 0 and fails == 0
    assert len(ops.query(empty)[0]) == 0 and len(ops.remove(empty)) == 0
    keys = oracle.gen_keys(41, 0, 3000)
    kd = dev_u64(keys)
    vals = kd ^ 5 if rep == pg.REP_KV64 else None
    g.bulk_insert(kd, vals)                     # no status, no fail counter
    g.bulk_query(kd)                            # no outputs at all
    torch.cuda.synchronize()
    # concurrently a key whose double-hashing step is 0 mod 512 buckets
    # probes one bucket 20 times and can find it full (the oracle, in key
    # order, stores all 3000): membership is exact for what was stored
    fill = g.get_fill()
    assert fill >= 2990
    f, v = ops.query(keys)
    assert int(f.sum()) == fill
    if rep == pg.REP_KV64:
        assert (v[f > 0] == (keys ^ np.uint64(5))[f > 0]).all() and not v[f == 0].any()
    stored = f > 0
    g.bulk_remove(kd[:1000])                    # no found flags
    torch.cuda.synchronize()
    f, _ = ops.query(keys)
    assert not f[:1000].any() and (f[1000:] > 0).tolist() == stored[1000:].tolist()
    assert g.get_fill() == int(stored[1000:].sum())


def test_flat_tcf_zero_key_calls_and_missing_outputs():
    """The flat TCF (16/16 key-value slots): zero-key calls, and calls
    without status, counter, value or found outputs.  16-bit keys can
    collide, so a remove may take a colliding key's slot: fill then follows
    the removes that landed."""
    g = pg.Table(pg.flat_tcf_desc(1 << 12, 1 << 7))
    ops = Ops(g)
    empty = np.zeros(0, dtype=np.uint64)
    st, fails = ops.insert(empty)
    assert len(st) == 0 and fails == 0
    assert len(ops.query(empty)[0]) == 0 and len(ops.remove(empty)) == 0
    keys = oracle.gen_keys(42, 0, 3000)
    kd = dev_u64(keys)
    g.bulk_insert(kd, (kd & 0x7FFF).to(torch.int16))   # no status, no fail counter
    g.bulk_query(kd)                                    # no outputs at all
    torch.cuda.synchronize()
    assert g.get_fill() == 3000                         # 73 % of level 0: every key fits
    f, v = ops.query(keys)
    assert f.all() and (v == (keys & np.uint64(0x7FFF)).astype(np.uint16)).mean() > 0.998
    g.bulk_remove(kd[:1000])                            # no found flags
    torch.cuda.synchronize()
    assert 1998 <= g.get_fill() <= 2002
    f, _ = ops.query(keys)
    assert f[1000:].mean() > 0.998 and f[:1000].mean() < 0.002


@pytest.mark.parametrize("mode", [pg.CONCURRENT, pg.DETERMINISTIC])
def test_batch_with_empty_tables(mode):
    sizes = np.array([4096, 1 << 13, 5000, 4096, 1 << 12, 6000], dtype=np.uint64)
    counts = [0, 6000, 0, 3000, 0, 0]            # first, middle and last tables get no keys
    b = pg.Batch(pg.tcf_desc(1 << 12, 1 << 6, bucket_size=16, tile=4), sizes)
    b.set_mode(mode)
    ins, off = [], [0]
    for t, c in enumerate(counts):
        ins.append(oracle.gen_keys(50 + t, 0, c))
        off.append(off[-1] + c)
    ins = np.concatenate(ins)
    off = np.array(off, dtype=np.uint64)
    st = torch.zeros(len(ins), dtype=torch.uint8, device="cuda")
    fails = torch.zeros(1, dtype=torch.int64, device="cuda")
    b.insert(dev_u64(ins), dev_u64(off), None, st, fails)
    found = torch.zeros(len(ins), dtype=torch.uint8, device="cuda")
    b.query(dev_u64(ins), dev_u64(off), None, found)
    torch.cuda.synchronize()
    assert int(fails.item()) == 0 and bool((st > 0).all()) and bool(found.all())
    assert [b.fill(t) for t in range(len(sizes))] == counts
    # a call with no keys at all (every table empty)
    zoff = dev_u64(np.zeros(len(sizes) + 1, dtype=np.uint64))
    nokeys = torch.zeros(0, dtype=torch.int64, device="cuda")
    b.insert(nokeys, zoff, None, None, fails)
    b.query(nokeys, zoff, None, torch.zeros(0, dtype=torch.uint8, device="cuda"))
    torch.cuda.synchronize()
    assert int(fails.item()) == 0 and [b.fill(t) for t in range(len(sizes))] == counts
    b.free()


def test_served_zero_key_requests():
    tabs = [pg.Table(pg.tcf_desc(1 << 14, bucket_size=16, tile=4)),
            pg.Table(pg.kv_desc(1 << 14, bucket_size=8, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20))]
    srv = pg.Server(tabs, budget_bytes=1 << 22, nstreams=2)
    keys = oracle.gen_keys(61, 0, 4000)
    kd = dev_u64(keys)
    empty = torch.zeros(0, dtype=torch.int64, device="cuda")
    st0 = torch.zeros(4000, dtype=torch.uint8, device="cuda")
    st1 = torch.zeros(4000, dtype=torch.uint8, device="cuda")
    f0 = torch.zeros(4000, dtype=torch.uint8, device="cuda")
    f1 = torch.zeros(4000, dtype=torch.uint8, device="cuda")
    v1 = torch.zeros(4000, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    tickets = [srv.submit(0, pg.OP_QUERY, empty, None, None, None),
               srv.submit(0, pg.OP_INSERT, kd, None, None, st0),
               srv.submit(1, pg.OP_INSERT, empty, None, None, None),
               srv.submit(1, pg.OP_INSERT, kd, kd ^ 9, None, st1),
               srv.submit(1, pg.OP_REMOVE, empty, None, None, None)]
    done = []
    for _ in range(1_000_000):
        done += srv.progress()
        if len(done) == len(tickets):
            break
    assert sorted(done) == sorted(tickets)
    tickets = [srv.submit(0, pg.OP_QUERY, kd, None, None, f0), srv.submit(1, pg.OP_QUERY, empty, None, None, None),
               srv.submit(1, pg.OP_QUERY, kd, None, v1, f1)]
    done = []
    for _ in range(1_000_000):
        done += srv.progress()
        if len(done) == len(tickets):
            break
    assert sorted(done) == sorted(tickets)
    torch.cuda.synchronize()
    assert bool((st0 > 0).all()) and bool((st1 > 0).all()) and bool(f0.all()) and bool(f1.all())
    assert bool((v1 == (kd ^ 9)).all())
    assert tabs[0].get_fill() == 4000 and tabs[1].get_fill() == 4000
    srv.free()
