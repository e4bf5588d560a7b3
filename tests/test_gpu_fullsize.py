"""Full-size GPU property tests: the 2^28-slot filter at 90 % load and the
2^30-slot key-value table at 0.9 load, checked through properties that do
not depend on size (no false negatives, fill = successful inserts, FP rate
against its closed form, exact values, delete / re-query consistency)."""
import pytest
import torch

import poggers_amd as pg

pytestmark = pytest.mark.gpu


def test_tcf_2p28_at_90_percent():
    nslots = 1 << 28
    t = pg.Table(pg.tcf_desc(nslots, tile=2), seed=42)
    n = int(0.9 * nslots)
    keys = torch.empty(n, dtype=torch.int64, device="cuda")
    pg.gen_keys(77, 0, n, keys)
    fails = torch.zeros(1, dtype=torch.int64, device="cuda")
    status = torch.empty(n, dtype=torch.uint8, device="cuda")
    for a in range(0, n, 1 << 24):
        t.bulk_insert(keys[a:a + (1 << 24)], None, status[a:a + (1 << 24)], fails)
    torch.cuda.synchronize()
    assert int(fails.item()) == 0
    in_backing = int((status == 2).sum().item())
    assert t.get_fill() == n
    # every inserted key is found (sampled across the insert order)
    sample = keys[::4].contiguous()
    found = torch.empty(sample.numel(), dtype=torch.uint8, device="cuda")
    t.bulk_query(sample, None, found)
    torch.cuda.synchronize()
    assert int((found == 0).sum().item()) == 0
    # FP rate of never-inserted keys vs (2 * items / nb + backing fill) * q
    nq = 1 << 26
    absent = torch.empty(nq, dtype=torch.int64, device="cuda")
    pg.gen_keys(77, 1 << 40, nq, absent)
    fa = torch.empty(nq, dtype=torch.uint8, device="cuda")
    t.bulk_query(absent, None, fa)
    torch.cuda.synchronize()
    got = float(fa.float().mean().item())
    q = (65532 + 8) / 2.0 ** 32
    backing_buckets = (nslots // 100 + 7) // 8
    expect = (2 * (n - in_backing) / (nslots // 16) + in_backing / backing_buckets) * q
    sigma = (expect / nq) ** 0.5
    assert abs(got - expect) < 5 * sigma, (got, expect)
    t.free_on_device()


def test_kv_2p30_double_hashing_at_0_9():
    m = 1 << 30
    t = pg.Table(pg.kv_desc(m, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20))
    n = int(0.9 * m)
    keys = torch.empty(n, dtype=torch.int64, device="cuda")
    pg.gen_keys(3, 0, n, keys)
    vals = keys * 7 + 1
    status = torch.empty(n, dtype=torch.uint8, device="cuda")
    fails = torch.zeros(1, dtype=torch.int64, device="cuda")
    for a in range(0, n, 1 << 25):
        t.bulk_insert(keys[a:a + (1 << 25)], vals[a:a + (1 << 25)], status[a:a + (1 << 25)], fails)
    torch.cuda.synchronize()
    ok = status > 0
    nok = int(ok.sum().item())
    assert int(fails.item()) == n - nok and n - nok < n // 100_000
    assert t.get_fill() == nok
    sample = torch.arange(0, n, 7, device="cuda")
    qk = keys[sample]
    found = torch.empty(qk.numel(), dtype=torch.uint8, device="cuda")
    out = torch.empty(qk.numel(), dtype=torch.int64, device="cuda")
    t.bulk_query(qk, out, found)
    torch.cuda.synchronize()
    assert bool((found.bool() == ok[sample]).all().item())
    assert bool(((out == vals[sample]) | ~ok[sample]).all().item())
    # delete every tenth key, re-query: deleted and failed keys absent, the rest exact
    dels = keys[::10].contiguous()
    rf = torch.empty(dels.numel(), dtype=torch.uint8, device="cuda")
    t.bulk_remove(dels, rf)
    torch.cuda.synchronize()
    assert bool((rf.bool() == ok[::10]).all().item())
    t.bulk_query(qk, out, found)
    torch.cuda.synchronize()
    live = ok[sample] & (sample % 10 != 0)
    assert bool((found.bool() == live).all().item())
    assert bool(((out == vals[sample]) | ~live).all().item())
    assert t.get_fill() == nok - int(ok[::10].sum().item())
    t.free_on_device()
