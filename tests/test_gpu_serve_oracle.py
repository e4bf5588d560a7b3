"""The serving loop (pg_server_*) against the oracle, on the serving
workload's request stream: 64 resident tables (32 filters, 32 key-value
tables), op mix 60 % query / 30 % insert / 10 % delete, request sizes
log-uniform, the table of each request chosen Zipf(1.1), Poisson arrivals.

* Scaled tables (2^17-2^21 slots), ordered per table (ORDER_TABLE): the
  stream is replayed through one oracle table per resident table in
  submission order (tables/bucketed_table.cuh:271-381 per key), and every
  served result is compared with the replay: key-value found flags, values,
  remove flags and insert statuses exactly; filter answers exactly up to the
  rare deletes that take another key's slot with the same fingerprint
  (their placement, hence which key loses its slot, differs between the
  concurrent and the key-order layouts).
* The workload's full table sizes (32 filters of 2^26-2^28 slots, 32
  key-value tables of 2^24-2^27 slots), unordered, requests of 1K-4M keys:
  every stored key is found (filters: unless a delete of a key with the same
  fingerprint in one of its buckets explains it), key-value answers carry
  key ^ 0x5A5A, every table's fill equals its stored-key count.
* Inserts and queries of the same keys in one batch, unordered: a query that
  finds a key-value key returns the value inserted with it (the slot is
  claimed busy, the value published, then the key)."""
import os
import time

import numpy as np
import pytest
import torch

import oracle
import poggers_amd as pg
from poggers_amd.workloads import keys_at

pytestmark = pytest.mark.gpu

KV_XOR = np.uint64(0x5A5A)


def pinned(a, dtype):
    t = torch.empty(len(a), dtype=dtype, pin_memory=True)
    if len(a):
        t.numpy()[:] = a
    return t


def zipf_weights(n, s=1.1):
    w = 1.0 / np.arange(1, n + 1) ** s
    return w / w.sum()


class Stream:
    """A request stream over resident tables with per-table key bookkeeping:
    inserts take fresh keys, queries mix stored keys (70 %) with keys never
    inserted, removes take distinct stored keys (plus 10 % never-inserted
    keys).  Table ranks for the Zipf choice go to the largest tables.  Each
    request carries a mask of its keys that are stored when it is
    submitted."""

    def __init__(self, kinds, npop, seed):
        self.kinds = kinds
        self.rng = np.random.default_rng(seed)
        self.live = [np.arange(n, dtype=np.int64) for n in npop]  # key counters stored in each table
        self.next = list(npop)
        self.absent = [0] * len(kinds)

    @staticmethod
    def key(t, c):
        return keys_at(700 + t, np.asarray(c, dtype=np.int64))

    def make(self, nreq, lo, hi):
        w = zipf_weights(len(self.kinds))
        out = []
        for _ in range(nreq):
            t = int(self.rng.choice(len(self.kinds), p=w))
            n = int(np.exp(self.rng.uniform(np.log(lo), np.log(hi))))
            u = self.rng.random()
            live = self.live[t]
            if u < 0.6:
                op = pg.OP_QUERY
                m = min(int(0.7 * n), len(live))
                c = live[self.rng.choice(len(live), m, replace=False)] if m else live[:0]
            elif u < 0.9:
                op, m = pg.OP_INSERT, n
                c = np.arange(self.next[t], self.next[t] + n, dtype=np.int64)
                self.next[t] += n
                self.live[t] = np.concatenate([live, c])
            else:
                op = pg.OP_REMOVE
                m = min(int(0.9 * n), len(live))
                pick = self.rng.choice(len(live), m, replace=False) if m else np.zeros(0, np.int64)
                c = live[pick]
                keep = np.ones(len(live), dtype=bool)
                keep[pick] = False
                self.live[t] = live[keep]
            keys = np.concatenate([self.key(t, c), self.fresh_absent(t, n - m)])
            stored = np.zeros(n, dtype=bool)
            stored[:m] = op != pg.OP_INSERT
            perm = self.rng.permutation(n)
            out.append((t, op, np.ascontiguousarray(keys[perm], dtype=np.uint64), stored[perm]))
        return out

    def fresh_absent(self, t, n):
        c = np.arange(self.absent[t], self.absent[t] + n, dtype=np.int64) + (1 << 50)
        self.absent[t] += n
        return self.key(t, c)


def close_count(g, o, rel=0.05):
    """Two event counts of the same process under different placements."""
    return abs(int(g) - int(o)) <= rel * o + 4.0 * np.sqrt(o + 1.0) + 2


def buffers(kind, op, keys):
    kv = kind == "kv"
    k = pinned(keys.view(np.int64), torch.int64)
    vals = out = None
    if op == pg.OP_INSERT:
        vals = (pinned((keys ^ KV_XOR).view(np.int64), torch.int64) if kv
                else pinned((keys & np.uint64(0xFFFF)).astype(np.uint16).view(np.int16), torch.int16))
    if op == pg.OP_QUERY:
        out = torch.zeros(len(keys), dtype=torch.int64 if kv else torch.int16, pin_memory=True)
    flags = torch.zeros(len(keys), dtype=torch.uint8, pin_memory=True)
    return k, vals, out, flags


def serve_poisson(srv, reqs, bufs, rate, seed):
    """Open-loop Poisson arrivals; every engine step first hands over the
    requests that have arrived.  Returns when every request is done."""
    arrive = np.cumsum(np.random.default_rng(seed).exponential(1.0 / rate, len(reqs)))
    packed = pg.Server.pack([(t, op) + b for (t, op, _), b in zip(reqs, bufs)])
    i = done = 0
    t0 = time.perf_counter()
    deadline = t0 + 120.0
    while done < len(reqs):
        now = time.perf_counter() - t0
        j = int(np.searchsorted(arrive, now, side="right"))
        if j > i:
            srv.submit_many(packed, i, j - i)
            i = j
        done += len(srv.progress())
        assert time.perf_counter() < deadline, "server did not drain"


def tcf_buckets(keys, nb, seed=42):
    """SPLIT64 probe buckets of fmix64 (level 0 of tcf_desc's filters)."""
    with np.errstate(over="ignore"):
        k = keys ^ np.uint64(seed)
        k ^= k >> np.uint64(33)
        k *= np.uint64(0xff51afd7ed558ccd)
        k ^= k >> np.uint64(33)
        k *= np.uint64(0xc4ceb9fe1a85ec53)
        k ^= k >> np.uint64(33)
    nb = np.uint64(nb)
    return (k & np.uint64(0xFFFFFFFF)) % nb, (k >> np.uint64(32)) % nb


def fp16(keys):
    f = (keys & np.uint64(0xFFFF)).astype(np.int64)
    f[f == 0] = 1
    f[f == 0xFFFF] = 0xFFFE
    return f


# PG_SERVE_SEEDS (comma-separated, default 9) runs the stream under more seeds
@pytest.mark.parametrize("seed", [int(x) for x in os.environ.get("PG_SERVE_SEEDS", "9").split(",")])
def test_serving_stream_matches_oracle_in_submission_order(seed):
    rng = np.random.default_rng(5)
    tcf_sizes = np.exp(rng.uniform(np.log(2.0 ** 19), np.log(2.0 ** 21), 32)).astype(np.int64)
    kv_sizes = np.exp(rng.uniform(np.log(2.0 ** 17), np.log(2.0 ** 20), 32)).astype(np.int64)
    specs = sorted([("tcf", int(s)) for s in tcf_sizes] + [("kv", int(s)) for s in kv_sizes],
                   key=lambda x: -(x[1] * (4 if x[0] == "tcf" else 1)))
    descs = [pg.tcf_desc(s, tile=4) if k == "tcf" else
             pg.kv_desc(s, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20) for k, s in specs]
    kinds = [k for k, _ in specs]
    npop = [int(0.2 * s) for _, s in specs]
    gpu = [pg.Table(d) for d in descs]
    ora = [oracle.Table(d.rep, [lv.as_dict() for lv in d.levels], seed=42) for d in descs]
    stream = Stream(kinds, npop, seed=seed)
    for t in range(len(gpu)):
        keys = Stream.key(t, np.arange(npop[t]))
        vals = keys ^ KV_XOR if kinds[t] == "kv" else (keys & np.uint64(0xFFFF)).astype(np.uint16)
        ora[t].bulk_insert(keys, vals)
        dk = torch.from_numpy(keys.view(np.int64)).cuda()
        dv = torch.from_numpy(vals.view(np.int64 if kinds[t] == "kv" else np.int16)).cuda()
        gpu[t].bulk_insert(dk, dv)
    torch.cuda.synchronize()
    stream_reqs = stream.make(600, 256, 16 << 10)
    reqs = [(t, op, k) for t, op, k, _ in stream_reqs]
    bufs = [buffers(kinds[t], op, k) for t, op, k in reqs]
    print(f"{len(reqs)} requests, {sum(len(k) for _, _, k in reqs)} keys over {len(gpu)} tables", flush=True)
    srv = pg.Server(gpu, budget_bytes=256 << 20, nstreams=4)
    srv.set_order(pg.ORDER_TABLE)
    serve_poisson(srv, reqs, bufs, rate=20000.0, seed=abs(seed - 6))
    srv.free()
    torch.cuda.synchronize()
    # filters: counts of stored keys not found (deletes that took a key's
    # slot through a shared fingerprint), of answers that differ on stored
    # keys, of false positives on never-inserted keys, of removes that found
    # nothing
    c = dict(miss_g=0, miss_o=0, diff=0, fp_g=0, fp_o=0, rm_g=0, rm_o=0)
    for i, ((t, op, keys, stored), (_, _, out, flags)) in enumerate(zip(stream_reqs, bufs)):
        o = ora[t]
        f = flags.numpy()
        kv = kinds[t] == "kv"
        if op == pg.OP_INSERT:
            st = o.bulk_insert(keys, keys ^ KV_XOR if kv else (keys & np.uint64(0xFFFF)).astype(np.uint16))
            assert (st > 0).all() and (f > 0).all(), (i, t)
            if kv:
                assert (f == st).all(), (i, t)
        elif op == pg.OP_REMOVE:
            fo = o.bulk_remove(keys)
            if kv:
                assert (f == fo).all() and (f == stored).all(), (i, t)
            else:
                c["rm_g"] += int((f == 0).sum())
                c["rm_o"] += int((fo == 0).sum())
        else:
            fo, vo = o.bulk_query(keys)
            if kv:
                assert (f == fo).all() and (f == stored).all(), (i, t)
                assert (out.numpy().view(np.uint64)[fo == 1] == vo[fo == 1]).all(), (i, t)
            else:
                vg = out.numpy().view(np.uint16)
                both = (f == 1) & (fo == 1)
                assert (vg[both] == vo[both]).all(), (i, t)
                c["miss_g"] += int((stored & (f == 0)).sum())
                c["miss_o"] += int((stored & (fo == 0)).sum())
                c["diff"] += int((stored & (f != fo)).sum())
                c["fp_g"] += int((~stored & (f == 1)).sum())
                c["fp_o"] += int((~stored & (fo == 1)).sum())
    print(f"filters (gpu/oracle): stored keys missed {c['miss_g']}/{c['miss_o']}, differing answers "
          f"{c['diff']}, false positives {c['fp_g']}/{c['fp_o']}, removes that found nothing "
          f"{c['rm_g']}/{c['rm_o']}", flush=True)
    assert c["diff"] <= c["miss_g"] + c["miss_o"]
    assert close_count(c["miss_g"], c["miss_o"]) and close_count(c["rm_g"], c["rm_o"])
    assert close_count(c["fp_g"], c["fp_o"])
    for t in range(len(gpu)):
        g, o = gpu[t].get_fill(), ora[t].fill()
        assert g == o if kinds[t] == "kv" else close_count(g, o, rel=0.0), (t, g, o)


def test_served_key_value_queries_never_see_a_key_without_its_value():
    """Inserts of fresh keys and queries of those same keys in one batch, with
    no ordering: a query may or may not find a key whose insert runs
    alongside it, but a key it finds carries the inserted value."""
    tabs = [pg.Table(pg.kv_desc(1 << 20, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20)),
            pg.Table(pg.kv_desc(1 << 20, bucket_size=1, tile=4, prober=pg.PROBE_LINEAR)),
            pg.Table(pg.TableDesc(pg.REP_KV64, [pg.Level(1 << 20, pg.SCHEME_POWER_OF_N, 8, 4, pg.PROBE_DOUBLE,
                                                         pg.HASH_MURMUR64A, 3)]))]
    srv = pg.Server(tabs, budget_bytes=64 << 20, nstreams=4)
    seen = 0
    stored = [0] * len(tabs)
    per_round = 150_000
    for rnd in range(4):
        reqs, bufs = [], []
        for t in range(len(tabs)):
            keys = keys_at(900 + t, np.arange(rnd * per_round, (rnd + 1) * per_round))
            for a in range(0, len(keys), 2000):
                k = keys[a:a + 2000]
                for op in (pg.OP_INSERT, pg.OP_QUERY, pg.OP_QUERY):
                    reqs.append((t, op, k))
                    bufs.append(buffers("kv", op, k))
        packed = pg.Server.pack([(t, op) + b for (t, op, _), b in zip(reqs, bufs)])
        srv.submit_many(packed, 0, len(reqs))
        done = 0
        deadline = time.perf_counter() + 60
        while done < len(reqs):
            done += len(srv.progress())
            assert time.perf_counter() < deadline
        status = {}
        for (t, op, k), (_, _, out, flags) in zip(reqs, bufs):
            f = flags.numpy()
            if op == pg.OP_INSERT:
                status.setdefault(t, np.zeros(256, np.int64))[:] += np.bincount(f, minlength=256)
            else:
                v = out.numpy().view(np.uint64)
                assert (v[f == 1] == (k ^ KV_XOR)[f == 1]).all(), (rnd, t)
                seen += int(f.sum())
        for t, c in status.items():
            # single-level tables: 1 = stored, 0 = no free slot on the probe path (double hashing
            # over a power-of-two bucket count can revisit one bucket, as in the reference)
            assert c[1] + c[0] == per_round and c[0] <= per_round // 10_000, (rnd, t, c[:2])
            stored[t] += int(c[1])
    print(f"{seen} of the concurrent queries found their key", flush=True)
    for t, tab in enumerate(tabs):
        assert tab.get_fill() == stored[t], (t, tab.get_fill(), stored[t])
    srv.free()


def test_serving_workload_full_size_properties():
    rng = np.random.default_rng(2024)
    tcf_sizes = np.exp(rng.uniform(np.log(2.0 ** 26), np.log(2.0 ** 28), 32)).astype(np.int64)
    kv_sizes = np.exp(rng.uniform(np.log(2.0 ** 24), np.log(2.0 ** 27), 32)).astype(np.int64)
    specs = sorted([("tcf", int(s)) for s in tcf_sizes] + [("kv", int(s)) for s in kv_sizes],
                   key=lambda x: -(x[1] * (4 if x[0] == "tcf" else 1)))
    kinds = [k for k, _ in specs]
    tabs = []
    npop = []
    tmp = torch.empty(1 << 26, dtype=torch.int64, device="cuda")
    for i, (kind, s) in enumerate(specs):
        t = pg.Table(pg.tcf_desc(s, tile=4) if kind == "tcf" else
                     pg.kv_desc(s, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20))
        n = int(0.4 * s)
        for a in range(0, n, 1 << 26):
            m = min(1 << 26, n - a)
            pg.gen_keys(1000 + i, a, m, tmp[:m])
            t.bulk_insert(tmp[:m], (tmp[:m] ^ 0x5A5A) if kind == "kv" else None)
        tabs.append(t)
        npop.append(n)
    del tmp
    torch.cuda.synchronize()
    gib = sum(t.host_bytes_in_use() for t in tabs) / 2 ** 30
    print(f"{len(tabs)} tables resident, {gib:.1f} GiB", flush=True)
    # queries: stored keys of the first half of each populated range (never
    # deleted); removes: distinct keys of the second half; inserts: fresh keys
    w = zipf_weights(len(tabs))
    r = np.random.default_rng(17)
    next_del = [n // 2 for n in npop]
    fresh = [0] * len(tabs)
    reqs = []
    for _ in range(200):
        t = int(r.choice(len(tabs), p=w))
        n = int(np.exp(r.uniform(np.log(1000), np.log(4 << 20))))
        u = r.random()
        if u < 0.6:
            op, keys = pg.OP_QUERY, keys_at(1000 + t, r.integers(0, npop[t] // 2, n))
        elif u < 0.9:
            op, keys = pg.OP_INSERT, keys_at(5000 + t, np.arange(fresh[t], fresh[t] + n) + (1 << 40))
            fresh[t] += n
        else:
            n = min(n, npop[t] - next_del[t])
            op, keys = pg.OP_REMOVE, keys_at(1000 + t, np.arange(next_del[t], next_del[t] + n))
            next_del[t] += n
        reqs.append((t, op, keys))
    bufs = [buffers(kinds[t], op, k) for t, op, k in reqs]
    nkeys = sum(len(k) for _, _, k in reqs)
    print(f"{len(reqs)} requests, {nkeys} keys", flush=True)
    srv = pg.Server(tabs, budget_bytes=4 << 30, nstreams=4)
    serve_poisson(srv, reqs, bufs, rate=2000.0, seed=5)
    srv.free()
    torch.cuda.synchronize()
    print("served; checking", flush=True)
    stored = [np.int64(n) for n in npop]
    deleted = [[] for _ in tabs]
    for (t, op, keys), (_, _, out, flags) in zip(reqs, bufs):
        f = flags.numpy()
        if op == pg.OP_INSERT:
            assert (f > 0).all(), t
            stored[t] += len(keys)
        elif op == pg.OP_REMOVE:
            deleted[t].append(keys)
            if kinds[t] == "kv":
                assert f.all(), t
            else:
                assert (f == 0).sum() <= 1e-4 * len(keys) + 2, t
            stored[t] -= int(f.sum())
    unexplained = 0
    shared_of = {}
    for (t, op, keys), (_, _, out, flags) in zip(reqs, bufs):
        if op != pg.OP_QUERY:
            continue
        f = flags.numpy()
        if kinds[t] == "kv":
            assert f.all(), t
            assert (out.numpy().view(np.uint64) == (keys ^ KV_XOR)).all(), t
            continue
        miss = keys[f == 0]
        if len(miss) == 0:
            continue
        # a filter delete tombstones the lowest slot with its fingerprint in its
        # buckets: a miss is explained by a deleted key sharing a bucket and
        # the fingerprint with the missing key
        nb = (specs[t][1] + 15) // 16
        if t not in shared_of:
            gone = np.concatenate(deleted[t]) if deleted[t] else np.zeros(0, np.uint64)
            g0, g1 = tcf_buckets(gone, nb)
            gfp = fp16(gone)
            shared_of[t] = np.unique(np.concatenate([g0.astype(np.int64) * 65536 + gfp,
                                                     g1.astype(np.int64) * 65536 + gfp]))
        shared = shared_of[t]
        m0, m1 = tcf_buckets(miss, nb)
        mfp = fp16(miss)
        ok = np.isin(m0.astype(np.int64) * 65536 + mfp, shared) | np.isin(m1.astype(np.int64) * 65536 + mfp, shared)
        unexplained += int((~ok).sum())
    assert unexplained == 0
    for t in range(len(tabs)):
        assert tabs[t].get_fill() == stored[t], (t, tabs[t].get_fill(), stored[t])
