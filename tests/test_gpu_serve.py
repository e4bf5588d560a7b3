"""GPU tests of the continuous-batching server (pg_server_*): every request's
result is the same whether it runs alone (per-request plain launch) or in
the served mix, with buffers used in place (pinned host memory) or staged
(pageable memory); admission of staged requests respects the HBM budget;
requests submitted while others are in flight complete; with per-table
ordering, dependent requests submitted at once give their one-by-one
results."""
import numpy as np
import pytest
import torch

import oracle
import poggers_amd as pg
from tests.gpu_util import dev_u64

pytestmark = pytest.mark.gpu

GOLD = np.uint64(0x9E3779B97F4A7C15)


def host_u64(a, pinned=True):
    t = torch.empty(len(a), dtype=torch.int64, pin_memory=pinned)
    t.numpy()[:] = np.ascontiguousarray(a, dtype=np.uint64).view(np.int64)
    return t


def host_u16(a, pinned=True):
    t = torch.empty(len(a), dtype=torch.int16, pin_memory=pinned)
    t.numpy()[:] = np.ascontiguousarray(a, dtype=np.uint16).view(np.int16)
    return t


def build_tables():
    tabs = [pg.Table(pg.tcf_desc(1 << 19, tile=4)), pg.Table(pg.tcf_desc(1 << 18, tile=4)),
            pg.Table(pg.kv_desc(1 << 19, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20)),
            pg.Table(pg.kv_desc(1 << 19, bucket_size=1, tile=4, prober=pg.PROBE_LINEAR)),
            pg.Table(pow_kv_desc(1 << 19))]
    return tabs


def pow_kv_desc(nslots):
    """key_val_pair<u64,u64> slots under power_of_n_insert_scheme: 8-slot
    buckets, three double-hashing probes (tests/src/test_components.cu:54),
    at the served tile width."""
    return pg.TableDesc(pg.REP_KV64, [pg.Level(nslots, pg.SCHEME_POWER_OF_N, 8, 4, pg.PROBE_DOUBLE,
                                                pg.HASH_MURMUR64A, 3)])


def populate(tabs, base):
    for i, t in enumerate(tabs):
        k = base[i]
        dk = torch.from_numpy(k.view(np.int64)).cuda()
        if t.is_kv:
            t.bulk_insert(dk, torch.from_numpy((k ^ GOLD).view(np.int64)).cuda())
        else:
            t.bulk_insert(dk, torch.from_numpy((k & np.uint64(0xFFFF)).astype(np.uint16).view(np.int16)).cuda())
    torch.cuda.synchronize()


def make_requests(rng, base, ntables, nreq):
    """Mixed requests whose answers do not depend on their order: queries of
    pre-populated keys (first half of each base set), deletes of keys from
    the second quarter, inserts of fresh keys."""
    reqs = []
    fresh_start = 1 << 44
    for r in range(nreq):
        t = int(rng.integers(ntables))
        n = int(np.exp(rng.uniform(np.log(1000), np.log(20000))))
        op = rng.choice([pg.OP_QUERY, pg.OP_QUERY, pg.OP_INSERT, pg.OP_REMOVE])
        b = base[t]
        if op == pg.OP_QUERY:
            keys = b[rng.integers(0, len(b) // 2, n)]
        elif op == pg.OP_REMOVE:
            lo = len(b) // 2 + (r * 997) % (len(b) // 4)
            keys = b[lo:lo + min(n, len(b) // 2 - (lo - len(b) // 2))]
            keys = np.unique(keys)
        else:
            keys = oracle.gen_keys(1000 + r, fresh_start, n)
        reqs.append((t, int(op), np.ascontiguousarray(keys, dtype=np.uint64)))
    return reqs


def buffers(tab, op, keys, pinned=True):
    n = len(keys)
    k = host_u64(keys, pinned)
    vals = out = None
    if op == pg.OP_INSERT:
        vals = (host_u64(keys ^ GOLD, pinned) if tab.is_kv
                else host_u16((keys & np.uint64(0xFFFF)).astype(np.uint16), pinned))
    if op == pg.OP_QUERY:
        out = torch.zeros(n, dtype=torch.int64 if tab.is_kv else torch.int16, pin_memory=pinned)
    flags = torch.zeros(n, dtype=torch.uint8, pin_memory=pinned)
    return k, vals, out, flags


@pytest.mark.parametrize("pinned", [True, False], ids=["in-place", "staged"])
def test_results_alone_equal_results_in_mix(pinned):
    rng = np.random.default_rng(4)
    tabs_mix, tabs_alone = build_tables(), build_tables()
    base = [oracle.gen_keys(50 + i, 0, int(0.5 * t.get_num_slots())) for i, t in enumerate(tabs_mix)]
    populate(tabs_mix, base)
    populate(tabs_alone, base)
    reqs = make_requests(rng, base, len(tabs_mix), 120)
    # deletes of one key by two requests would race: keep each key in one delete
    seen = set()
    clean = []
    for t, op, k in reqs:
        if op == pg.OP_REMOVE:
            k = np.array([x for x in k if (t, int(x)) not in seen], dtype=np.uint64)
            seen.update((t, int(x)) for x in k)
            if len(k) == 0:
                continue
        clean.append((t, op, k))
    reqs = clean
    srv_mix = pg.Server(tabs_mix, budget_bytes=64 << 20, nstreams=4)
    bufs_mix = [buffers(tabs_mix[t], op, k, pinned) for t, op, k in reqs]
    tickets = {}
    done = set()
    for i, ((t, op, k), (kk, vv, oo, ff)) in enumerate(zip(reqs, bufs_mix)):
        tickets[srv_mix.submit(t, op, kk, vv, oo, ff)] = i
        if i % 7 == 0:
            done.update(srv_mix.progress())  # requests keep arriving while batches run
    for _ in range(1_000_000):
        done.update(srv_mix.progress())
        if len(done) == len(reqs):
            break
    assert len(done) == len(reqs)
    srv_alone = pg.Server(tabs_alone, budget_bytes=64 << 20, nstreams=1)
    bufs_alone = [buffers(tabs_alone[t], op, k, pinned) for t, op, k in reqs]
    for (t, op, k), (kk, vv, oo, ff) in zip(reqs, bufs_alone):
        srv_alone.run_plain(t, op, kk, vv, oo, ff)
    for i, ((t, op, k), bm, ba) in enumerate(zip(reqs, bufs_mix, bufs_alone)):
        fm, fa = bm[3].numpy(), ba[3].numpy()
        if op == pg.OP_QUERY:
            assert fm.all() and fa.all(), i          # pre-populated keys: no false negatives
            om, oa = bm[2].numpy(), ba[2].numpy()
            if tabs_mix[t].is_kv:
                assert (om == oa).all() and (om.view(np.uint64) == (k ^ GOLD)).all(), i
            else:
                assert (om == oa).mean() > 0.995, i   # fingerprint collisions may pick another slot
        elif op == pg.OP_INSERT:
            # the level a key lands in (status 1 or 2) depends on insertion order
            assert (fm > 0).all() and (fa > 0).all(), i
        else:
            assert (fm == fa).all() and fm.all(), (i, op)
    srv_mix.free()
    srv_alone.free()


def test_admission_respects_budget_and_drains_queue():
    """Pageable buffers are staged through the arena: the budget bounds what
    is in flight, and a request that can never fit is refused."""
    tabs = build_tables()[:1]
    srv = pg.Server(tabs, budget_bytes=1 << 20, nstreams=2)
    big = oracle.gen_keys(1, 0, 1 << 17)       # 1 MiB of keys alone: can never fit
    kk, _, _, ff = buffers(tabs[0], pg.OP_INSERT, big, pinned=False)
    with pytest.raises(pg.PgError):
        srv.submit(0, pg.OP_INSERT, kk, None, None, ff)
    reqs = [oracle.gen_keys(2, i * 20000, 20000) for i in range(20)]  # 4.4 MB staged in total, 76 % load
    bufs = [buffers(tabs[0], pg.OP_INSERT, k, pinned=False) for k in reqs]
    for (kk, vv, oo, ff) in bufs:
        srv.submit(0, pg.OP_INSERT, kk, vv, oo, ff)
    q, f = srv.pending()
    assert q == len(reqs)
    done = []
    max_inflight = 0
    for _ in range(1_000_000):
        done += srv.progress()
        max_inflight = max(max_inflight, srv.pending()[1])
        if len(done) == len(reqs):
            break
    assert sorted(done) == sorted(set(done)) and len(done) == len(reqs)
    assert max_inflight * 20000 * 11 <= (1 << 20)  # keys + vals + flags of in-flight requests fit
    for (kk, vv, oo, ff) in bufs:
        assert ff.numpy().all()
    assert tabs[0].get_fill() == 20 * 20000
    srv.free()


def test_in_place_requests_take_no_staging():
    """Pinned buffers are used in place: a request far larger than the
    staging budget is admitted, and its results land in the caller's
    buffers."""
    tabs = build_tables()[:1]
    srv = pg.Server(tabs, budget_bytes=1 << 20, nstreams=2)
    keys = oracle.gen_keys(3, 0, 1 << 18)      # 2 MiB of keys > the 1 MiB budget
    kk, vv, _, ff = buffers(tabs[0], pg.OP_INSERT, keys)
    srv.submit(0, pg.OP_INSERT, kk, vv, None, ff)
    done = []
    for _ in range(1_000_000):
        done += srv.progress()
        if done:
            break
    assert len(done) == 1 and ff.numpy().all()
    qk, _, qo, qf = buffers(tabs[0], pg.OP_QUERY, keys)
    srv.submit(0, pg.OP_QUERY, qk, None, qo, qf)
    done = []
    for _ in range(1_000_000):
        done += srv.progress()
        if done:
            break
    assert qf.numpy().all()
    assert (qo.numpy().view(np.uint16) == (keys & np.uint64(0xFFFF)).astype(np.uint16)).mean() > 0.99
    srv.free()


def test_ordered_mode_applies_writes_in_submission_order():
    """ORDER_TABLE: two tables (bucket_insert and power-of-n key-value
    tables) each get a dependent chain, submitted at once
    and interleaved: insert, query, remove half, query, re-insert that half
    with new values, query.  Every result is the one-by-one result; with
    the default ordering the same burst would land in one batch."""
    tabs = [pg.Table(pg.kv_desc(1 << 16, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20)),
            pg.Table(pow_kv_desc(1 << 16))]
    srv = pg.Server(tabs, budget_bytes=1 << 24, nstreams=4)
    srv.set_order(pg.ORDER_TABLE)
    n = 4000
    chains = []
    for t in range(2):
        k = dev_u64(oracle.gen_keys(70 + t, 0, n))
        v1, v2 = k ^ 0x1111, k ^ 0x2222
        half = k[::2].contiguous()
        outs = [(torch.zeros(n, dtype=torch.int64, device="cuda"), torch.zeros(n, dtype=torch.uint8, device="cuda"))
                for _ in range(3)]
        steps = [(pg.OP_INSERT, k, v1, None, None), (pg.OP_QUERY, k, None) + outs[0],
                 (pg.OP_REMOVE, half, None, None, None), (pg.OP_QUERY, k, None) + outs[1],
                 (pg.OP_INSERT, half, v2[::2].contiguous(), None, None), (pg.OP_QUERY, k, None) + outs[2]]
        chains.append((k, v1, v2, outs, steps))
    torch.cuda.synchronize()
    tickets = []
    for i in range(6):
        for t in range(2):
            op, keys, vals, vo, fl = chains[t][4][i]
            tickets.append(srv.submit(t, op, keys, vals, vo, fl))
    done = []
    for _ in range(1_000_000):
        done += srv.progress()
        if len(done) == len(tickets):
            break
    assert sorted(done) == sorted(tickets)
    torch.cuda.synchronize()
    even = torch.arange(n, device="cuda") % 2 == 0
    for k, v1, v2, outs, _ in chains:
        (o1, f1), (o2, f2), (o3, f3) = outs
        assert bool(f1.all()) and bool((o1 == v1).all())
        assert not bool(f2[even].any()) and bool(f2[~even].all()) and bool((o2[~even] == v1[~even]).all())
        assert bool(f3.all()) and bool((o3[even] == v2[even]).all()) and bool((o3[~even] == v1[~even]).all())
    assert tabs[0].get_fill() == n and tabs[1].get_fill() == n
    srv.free()


def test_progress_hands_over_a_large_batch_in_parts():
    """3000 small requests form one batch; with room for 100 tickets per
    engine step every ticket is still reported, exactly once."""
    t = pg.Table(pg.kv_desc(1 << 16, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20))
    srv = pg.Server([t], budget_bytes=1 << 22, nstreams=1)
    keys = dev_u64(oracle.gen_keys(75, 0, 3000 * 4))
    flags = torch.zeros(3000 * 4, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    tickets = [srv.submit(0, pg.OP_INSERT, keys[4 * i:4 * i + 4], None, None, flags[4 * i:4 * i + 4])
               for i in range(3000)]
    done, steps = [], 0
    for _ in range(1_000_000):
        got = srv.progress(cap=100)
        assert len(got) <= 100
        done += got
        steps += 1
        if len(done) >= len(tickets):
            break
    assert sorted(done) == sorted(tickets) and steps >= 30
    torch.cuda.synchronize()
    assert bool((flags == 1).all()) and t.get_fill() == 12000
    srv.free()


def test_flat_tcf_served_like_plain_calls():
    """The reference's flat TCF (16/16 key-value slots): insert requests,
    then queries of every key and removes of the last quarter, served and
    as per-request plain calls on a twin table.  Every key is stored and
    found both ways, every remove lands, and both tables end with the same
    fill.  16-bit keys can collide, and the two layouts differ in slot order,
    so a query of a colliding key may read the other key's slot (the values
    here are the keys' low 16 bits, which such keys share)."""
    tabs = [pg.Table(pg.flat_tcf_desc(1 << 18, 1 << 12)) for _ in range(2)]
    srv = pg.Server([tabs[0]], budget_bytes=64 << 20, nstreams=4)
    plain = pg.Server([tabs[1]], budget_bytes=64 << 20, nstreams=1)
    keys = oracle.gen_keys(80, 0, 200000)          # 76 % of level 0
    chunks = [keys[a:a + 20000] for a in range(0, len(keys), 20000)]
    n_rm = len(keys) // 4

    def drain(s, want):
        done = []
        for _ in range(1_000_000):
            done += s.progress()
            if len(done) == want:
                return done
        raise AssertionError("server did not drain")

    results = []
    for s, run in ((srv, False), (plain, True)):
        tab = tabs[1] if run else tabs[0]
        ins = [buffers(tab, pg.OP_INSERT, k) for k in chunks]
        qry = [buffers(tab, pg.OP_QUERY, k) for k in chunks]
        rm = buffers(tab, pg.OP_REMOVE, keys[-n_rm:])
        phases = [[(pg.OP_INSERT, b) for b in ins], [(pg.OP_QUERY, b) for b in qry], [(pg.OP_REMOVE, rm)]]
        for phase in phases:
            for op, (kk, vv, oo, ff) in phase:
                if run:
                    s.run_plain(0, op, kk, vv, oo, ff)
                else:
                    s.submit(0, op, kk, vv, oo, ff)
            if not run:
                drain(s, len(phase))
        results.append((ins, qry, rm))
    (ins_s, qry_s, rm_s), (ins_p, qry_p, rm_p) = results
    for a, b in zip(ins_s, ins_p):
        assert (a[3].numpy() == 1).all() and (b[3].numpy() == 1).all()
    same = 0
    for a, b, k in zip(qry_s, qry_p, chunks):
        assert a[3].numpy().all() and b[3].numpy().all()
        want = (k & np.uint64(0xFFFF)).astype(np.uint16).view(np.int16)
        assert (a[2].numpy() == want).mean() > 0.999 and (b[2].numpy() == want).mean() > 0.999
        same += int((a[2].numpy() == b[2].numpy()).sum())
    assert same >= len(keys) - 20
    assert rm_s[3].numpy().all() and rm_p[3].numpy().all()
    assert tabs[0].get_fill() == tabs[1].get_fill() == len(keys) - n_rm
    srv.free()
    plain.free()


def test_binding_keeps_request_tensors_until_done():
    """The server reads and writes request buffers in place after submit
    returns: the binding holds each request's tensors until progress()
    reports its ticket, so temporaries passed to submit / submit_many stay
    valid, and the results land in them."""
    import gc
    tabs = build_tables()[:1]
    srv = pg.Server(tabs, budget_bytes=1 << 22, nstreams=2)
    keys = oracle.gen_keys(91, 0, 20000)
    kk, vv, _, ff = buffers(tabs[0], pg.OP_INSERT, keys)
    t0 = srv.submit(0, pg.OP_INSERT, kk, vv, None, ff)
    flags_ref = ff
    del kk, vv, ff
    gc.collect()
    assert t0 in srv._live
    done = []
    for _ in range(1_000_000):
        done += srv.progress()
        if t0 in done:
            break
    assert t0 not in srv._live and flags_ref.numpy().all()
    # submit_many: the packed array keeps its tensors; the tickets keep the array
    qk, _, qo, qf = buffers(tabs[0], pg.OP_QUERY, keys)
    packed = pg.Server.pack([(0, pg.OP_QUERY, qk, None, qo, qf)], tabs)
    out_ref, found_ref = qo, qf
    del qk, qo, qf
    tickets = srv.submit_many(packed)
    del packed
    gc.collect()
    assert all(tk in srv._live for tk in tickets)
    done = []
    for _ in range(1_000_000):
        done += srv.progress()
        if set(tickets) <= set(done):
            break
    assert not srv._live and found_ref.numpy().all()
    srv.free()
