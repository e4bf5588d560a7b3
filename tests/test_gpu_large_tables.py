"""Tables past 32-bit slot indexes, as a 288 GB part holds them: a 2^34-slot
filter (64 GiB, 2^30 buckets) and a 2^32-slot key-value table (64 GiB).
Sparse fills keep the runs short.  Besides membership, each test copies four
64 MiB slices of the table, at 0, 1/4, 1/2 and 3/4 of its bytes, and counts
the occupied slots in each: a slot index truncated to 32 bits would leave
every slice but the first empty."""
import ctypes

import numpy as np
import pytest
import torch

import poggers_amd as pg

pytestmark = pytest.mark.gpu

SLICE = 64 << 20


def occupied_per_quarter(table, dtype, occupied):
    p, n = ctypes.c_void_p(), ctypes.c_uint64()
    pg.check(pg.lib().pg_level_ptr(table._h, 0, ctypes.byref(p), ctypes.byref(n)), "pg_level_ptr")
    torch.cuda.synchronize()
    counts = []
    for q in range(4):
        buf = ctypes.create_string_buffer(SLICE)
        pg._hip_memcpy(buf, p.value + (n.value // 4) * q, SLICE, 2)
        counts.append(int(occupied(np.frombuffer(buf.raw, dtype=dtype)).sum()))
    return n.value, counts


def test_tcf_2p34_slots():
    nslots = 1 << 34
    t = pg.Table(pg.tcf_desc(nslots, nslots // 100, bucket_size=16, tile=2, prober=pg.PROBE_DOUBLE,
                             hasher=pg.HASH_FMIX64, backing_probes=4000))
    n = 1 << 25
    keys = torch.empty(n, dtype=torch.int64, device="cuda")
    pg.gen_keys(91, 0, n, keys)
    st = torch.zeros(n, dtype=torch.uint8, device="cuda")
    t.bulk_insert(keys, None, st)
    found = torch.zeros(n, dtype=torch.uint8, device="cuda")
    t.bulk_query(keys, None, found)
    absent = torch.empty(n, dtype=torch.int64, device="cuda")
    pg.gen_keys(91, 1 << 40, n, absent)
    fa = torch.zeros(n, dtype=torch.uint8, device="cuda")
    t.bulk_query(absent, None, fa)
    torch.cuda.synchronize()
    assert bool((st == 1).all()) and bool(found.all()) and t.get_fill() == n
    assert float(fa.float().mean()) < 1e-4      # two sparse buckets: ~2 * 16 * 2^-9 / 2^16 plus noise
    nbytes, counts = occupied_per_quarter(t, np.uint32, lambda w: (w & 0xFFFF) != 0)
    assert nbytes == nslots * 4
    expect = n * SLICE / nbytes                    # 2^25 keys spread over 64 GiB
    assert all(abs(c - expect) < 6 * expect ** 0.5 for c in counts), (counts, expect)
    t.free_on_device()


def test_kv_2p32_slots():
    nslots = 1 << 32
    t = pg.Table(pg.kv_desc(nslots, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20))
    n = 1 << 25
    keys = torch.empty(n, dtype=torch.int64, device="cuda")
    pg.gen_keys(92, 0, n, keys)
    vals = keys ^ 0x5A5A
    st = torch.zeros(n, dtype=torch.uint8, device="cuda")
    t.bulk_insert(keys, vals, st)
    out = torch.zeros(n, dtype=torch.int64, device="cuda")
    found = torch.zeros(n, dtype=torch.uint8, device="cuda")
    t.bulk_query(keys, out, found)
    torch.cuda.synchronize()
    assert bool((st == 1).all()) and bool(found.all()) and bool((out == vals).all()) and t.get_fill() == n
    nbytes, counts = occupied_per_quarter(t, np.uint64, lambda w: w.reshape(-1, 2)[:, 0] != 0)
    assert nbytes == nslots * 16
    expect = n * SLICE / nbytes
    assert all(abs(c - expect) < 6 * expect ** 0.5 for c in counts), (counts, expect)
    t.free_on_device()
