"""Bulk calls whose one-tile-per-key grid would pass HIP's work-item limit for
one grid dimension (below 2^32): single-table calls are cut into launches of
at most 2^31 work-items, and a served batch runs a capped grid in which each
tile takes every tiles-th key.  Every key of such a call must be processed
exactly once, with its own value and status: the ragged sizes below end the
last launch mid-way."""
import numpy as np
import pytest
import torch

import oracle
import poggers_amd as pg
from tests.gpu_util import dev_u64

pytestmark = pytest.mark.gpu

GOLD = 0x9E3779B97F4A7C15 - (1 << 64)


def test_kv_calls_past_2p32_work_items():
    n = (1 << 26) + 7                              # x 64 lanes = 2^32 + 448 work-items
    t = pg.Table(pg.kv_desc(1 << 27, bucket_size=16, tile=64, prober=pg.PROBE_DOUBLE, max_probes=20))
    keys = torch.empty(n, dtype=torch.int64, device="cuda")
    pg.gen_keys(17, 0, n, keys)
    vals = keys ^ torch.tensor(GOLD, dtype=torch.int64, device="cuda")
    st = torch.zeros(n, dtype=torch.uint8, device="cuda")
    fails = torch.zeros(1, dtype=torch.int64, device="cuda")
    t.bulk_insert(keys, vals, st, fails)
    torch.cuda.synchronize()
    assert int(fails.item()) == 0 and bool((st == 1).all())
    assert t.get_fill() == n
    out = torch.zeros(n, dtype=torch.int64, device="cuda")
    found = torch.zeros(n, dtype=torch.uint8, device="cuda")
    t.bulk_query(keys, out, found)
    torch.cuda.synchronize()
    assert bool((found == 1).all()) and bool((out == vals).all())
    rm = torch.zeros(n, dtype=torch.uint8, device="cuda")
    t.bulk_remove(keys, rm)
    torch.cuda.synchronize()
    assert bool((rm == 1).all()) and t.get_fill() == 0


def test_tcf_query_past_2p32_work_items():
    """Each key's answer in the 2^27-key launch equals its answer in a small
    launch (a filter may answer a key with the value of another key whose
    fingerprint shares its bucket, so values are compared with that)."""
    t = pg.Table(pg.tcf_desc(1 << 16, bucket_size=32, tile=32))
    base = oracle.gen_keys(18, 0, 1 << 14)
    vals = ((base >> np.uint64(7)) & np.uint64(0xFFFF)).astype(np.uint16)
    t.bulk_insert(dev_u64(base), torch.from_numpy(vals.view(np.int16)).cuda())
    absent = oracle.gen_keys(18, 1 << 40, 1 << 14)
    pool = dev_u64(np.concatenate([base, absent]))
    f1 = torch.zeros(pool.numel(), dtype=torch.uint8, device="cuda")
    v1 = torch.zeros(pool.numel(), dtype=torch.int16, device="cuda")
    t.bulk_query(pool, v1, f1)
    n = (1 << 27) + 5                              # x 32 lanes = 2^32 + 160 work-items
    idx = torch.arange(n, device="cuda", dtype=torch.int64) % pool.numel()
    q = pool[idx]
    found = torch.full((n,), 7, dtype=torch.uint8, device="cuda")
    out = torch.zeros(n, dtype=torch.int16, device="cuda")
    t.bulk_query(q, out, found)
    torch.cuda.synchronize()
    assert bool(f1[:base.size].all())
    assert bool((found == f1[idx]).all()) and bool((out == v1[idx]).all())
    del q, idx, found, out


def test_served_batch_past_2p31_work_items():
    t = pg.Table(pg.kv_desc(1 << 16, bucket_size=16, tile=16, prober=pg.PROBE_DOUBLE, max_probes=20))
    base = oracle.gen_keys(19, 0, 1 << 15)
    kb = dev_u64(base)
    t.bulk_insert(kb, kb ^ torch.tensor(GOLD, dtype=torch.int64, device="cuda"))
    n = (1 << 27) + 3                              # x 16 lanes = 2^31 + 48 work-items
    idx = torch.arange(n, device="cuda", dtype=torch.int64) % kb.numel()
    q = kb[idx]
    out = torch.zeros(n, dtype=torch.int64, device="cuda")
    found = torch.zeros(n, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    srv = pg.Server([t], budget_bytes=1 << 20, nstreams=2)
    srv.submit(0, pg.OP_QUERY, q, None, out, found)
    done = []
    for _ in range(10_000_000):
        done += srv.progress()
        if done:
            break
    assert len(done) == 1
    assert bool((found == 1).all()) and bool((out == (q ^ torch.tensor(GOLD, dtype=torch.int64, device="cuda"))).all())
    srv.free()
