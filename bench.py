"""Benchmark: Two-Choice Filter bulk insert + query at 90 % load on one MI355X.

Workload (one "step"): a 2^28-slot TCF (16-bit fingerprints + 16-bit values,
16-slot buckets = 64 B, two choices from one fmix64 hash split into 32-bit
halves) with a 1 % linear backing level is cleared, filled to 90 % of its
primary slots (241 591 910 distinct uniform uint64 keys, batches of 2^24),
then queried with 2^26 inserted and 2^26 never-inserted keys.  All inputs are
generated on the GPU before timing and stay resident in HBM.

Prints one JSON line (the bench contract).  `value` is whole-step throughput
in Mops/s (inserts + queries per second).  `roofline` is for the dominant
kernel (insert), with algorithmic bytes counted by the kernels' own
accounting counters in an untimed pass; `cpu_baseline` times the CPU oracle
on a bounded sample of the same workload.  `serving` is the served request
loop at one operating point: 64 resident tables, requests of 1K-64K keys,
Poisson arrivals, against per-request plain launches.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--log2-slots", type=int, default=28)
    ap.add_argument("--tile", type=int, default=2)
    ap.add_argument("--query-items", type=int, default=0,
                    help="query launch shape: keys per tile (1, 2, 4; 0 = library default)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--no-roofline", action="store_true", help="skip the random-sector microbenchmark")
    ap.add_argument("--no-serve", action="store_true",
                    help="skip the serving leg of the default line (64 resident tables, 1K-64K-key requests)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a 1 GiB streaming read+write kernel and 2^24 random 64 B reads "
                         "of a 1 GiB buffer (known byte counts that scale the PMC byte counters)")
    ap.add_argument("--summarize-profiles", nargs="+", metavar="DIR",
                    help="write profiles/ summaries from rocprofv3 csv output (kernel stats, "
                         "FETCH_SIZE run, WRITE_SIZE run, and optionally the TCC counter run) and exit")
    ap.add_argument("--summarize-kernel-stats", nargs=2, metavar=("STATS_DIR", "NAME"),
                    help="write profiles/NAME_kernel_stats.md from a rocprofv3 --stats csv directory "
                         "(NAME: kv_2p30, flat_tcf_2p28, batched_4096, serve_64_tables) and exit")
    ap.add_argument("--requests", type=int, default=400, help="serve: requests per experiment")
    ap.add_argument("--workload", choices=["tcf", "batched", "kv", "serve", "flat"], default="tcf",
                    help="tcf: the headline 2^28-slot filter; batched: 4096 uneven filters in one graph; "
                         "kv: 2^30-slot double-hashing key-value table; flat: the flat TCF "
                         "(16/16 key-value slots, 2^28 + 1 %% slots)")
    ap.add_argument("--tables", type=int, default=4096, help="batched: number of filters")
    ap.add_argument("--kv-tile", type=int, default=4, help="kv: tile width (4 is hash_table_64's)")
    ap.add_argument("--flat-tile", type=int, default=4, help="flat: tile width (4 is the reference's)")
    ap.add_argument("--kv-scheme", choices=["bucket", "pow"], default="bucket",
                    help="kv: bucket = hash_table_64's bucket_insert (B 16, 20 probes); pow = "
                         "power_of_n_insert_scheme (B 8, 3 probes, tile 8: the shape of the reference's p2_table)")
    return ap.parse_args()


def run_kv(args):
    """Key-value table of 2^30 16 B slots (16 GiB), double hashing, 16-slot
    buckets, tile 4, 20 probes (the hash_table_64 shape), filled to 0.9 with
    distinct uniform keys (value = key ^ golden ratio); a query stream of 2^28
    keys, half drawn Zipf(0.99) over the inserted keys and half absent; then
    10 % of the keys deleted and the stream re-queried."""
    import torch
    import poggers_amd as pg

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    log2 = args.log2_slots if args.log2_slots != 28 else 30
    m = 1 << log2
    n = int(0.9 * m)
    nq = 1 << (log2 - 2)
    if args.kv_scheme == "pow":
        table = pg.Table(pg.TableDesc(pg.REP_KV64, [pg.Level(m, pg.SCHEME_POWER_OF_N, 8, 8, pg.PROBE_DOUBLE,
                                                             pg.HASH_FMIX64, 3)]))
        shape = "power_of_n_insert_scheme, B=8, tile 8, 3 probes"
    else:
        table = pg.Table(pg.kv_desc(m, bucket_size=16, tile=args.kv_tile, prober=pg.PROBE_DOUBLE,
                                    hasher=pg.HASH_FMIX64, max_probes=20))
        shape = f"bucket_insert, B=16, tile {args.kv_tile}, 20 probes"
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    pg.gen_keys(99, 0, n, keys)
    vals = keys ^ torch.tensor(0x9E3779B97F4A7C15 - (1 << 64), dtype=torch.int64, device=dev)
    # bounded Zipf(s) ranks by inverse transform of the continuous CDF
    s = 0.99
    gen = torch.Generator(device=dev)
    gen.manual_seed(5)
    u = torch.rand(nq // 2, dtype=torch.float64, device=dev, generator=gen)
    x = (1.0 + u * (float(n) ** (1.0 - s) - 1.0)) ** (1.0 / (1.0 - s))
    ranks = torch.clamp(x.floor().long() - 1, 0, n - 1)
    absent = torch.empty(nq - nq // 2, dtype=torch.int64, device=dev)
    pg.gen_keys(99, 1 << 40, absent.numel(), absent)
    q = torch.cat([keys[ranks], absent])
    found = torch.empty(nq, dtype=torch.uint8, device=dev)
    out = torch.empty(nq, dtype=torch.int64, device=dev)
    fails = torch.zeros(1, dtype=torch.int64, device=dev)
    status = torch.empty(n, dtype=torch.uint8, device=dev)
    ev = lambda: torch.cuda.Event(enable_timing=True)
    batch = 1 << 24
    dels = keys[::10].contiguous()  # every tenth key: 10 % spread over all Zipf ranks
    ndel = dels.numel()

    def timed(fn):
        e0, e1 = ev(), ev()
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    res = {"insert": 0.0, "query": 0.0, "delete": 0.0, "requery": 0.0}
    for step in range(args.warmup + args.steps):
        table.clear()
        fails.zero_()
        ti = timed(lambda: [table.bulk_insert(keys[a:a + batch], vals[a:a + batch], status[a:a + batch], fails)
                            for a in range(0, n, batch)])
        tq = timed(lambda: table.bulk_query(q, out, found))
        hits1 = int(found[:nq // 2].sum().item())
        ok_vals = bool(((out[:nq // 2] == vals[ranks]) | (found[:nq // 2] == 0)).all().item())
        fp1 = int(found[nq // 2:].sum().item())
        td = timed(lambda: table.bulk_remove(dels))
        tr = timed(lambda: table.bulk_query(q, out, found))
        if step >= args.warmup:
            for k, v in zip(res, (ti, tq, td, tr)):
                res[k] += v / args.steps
    inserted = status > 0
    hits_expected = int(inserted[ranks].sum().item())
    live_hits = int((inserted[ranks] & (ranks % 10 != 0)).sum().item())
    hits2 = int(found[:nq // 2].sum().item())
    out_line = {
        "metric": "KV double-hashing table: insert/query/delete Mops/s at 0.9 load",
        "value": round((n + 2 * nq + ndel) / (sum(res.values()) / 1e3) / 1e6, 1),
        "unit": "Mops/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(sum(res.values()), 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64", "data": "synthetic (splitmix64 keys, Zipf(0.99) hits)",
        "config": {"workload": f"KV 2^{log2} slots x 16 B ({m * 16 >> 30} GiB), double hashing, {shape}; "
                               f"{n} inserts, {nq} queries (50% Zipf(0.99) hits, 50% absent), "
                               f"delete {ndel}, re-query"},
        "insert_mops": round(n / res["insert"] / 1e3, 1), "query_mops": round(nq / res["query"] / 1e3, 1),
        "delete_mops": round(ndel / res["delete"] / 1e3, 1), "requery_mops": round(nq / res["requery"] / 1e3, 1),
        "insert_failures": int(fails.item()),
        "hit_queries_found": hits1, "hit_queries": nq // 2, "hit_queries_of_inserted_keys": hits_expected,
        "hit_values_exact": ok_vals,
        "absent_found": fp1, "requery_hits": hits2, "requery_hits_expected": live_hits,
        "ms": {k: round(v, 3) for k, v in res.items()},
    }
    print(json.dumps(out_line), flush=True)


def run_flat(args):
    """The flat TCF (16/16 key-value slots, 16-slot buckets, two double-hash
    choices with the 75 % shortcut, tile 4, a 1 % backing level of 4-slot
    buckets with 20 probes) at 2^28 primary slots: fill to 0.9 of the primary
    in batches of 2^24 keys, query 2^26 inserted and 2^26 absent keys, remove
    10 % of the keys, re-query."""
    import torch
    import poggers_amd as pg

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    m = 1 << args.log2_slots
    n = int(0.9 * m)
    nq = 1 << (args.log2_slots - 2)
    table = pg.Table(pg.flat_tcf_desc(m, m // 100, tile=args.flat_tile))
    keys = torch.empty(n, dtype=torch.int64, device=dev)
    pg.gen_keys(98, 0, n, keys)
    vals = ((keys >> 20) & 0xFFFF).to(torch.int16)
    absent = torch.empty(nq // 2, dtype=torch.int64, device=dev)
    pg.gen_keys(98, 1 << 40, absent.numel(), absent)
    q = torch.cat([keys[: nq // 2], absent])
    found = torch.empty(nq, dtype=torch.uint8, device=dev)
    out = torch.empty(nq, dtype=torch.int16, device=dev)
    fails = torch.zeros(1, dtype=torch.int64, device=dev)
    status = torch.empty(n, dtype=torch.uint8, device=dev)
    dels = keys[::10].contiguous()
    ev = lambda: torch.cuda.Event(enable_timing=True)
    batch = 1 << 24

    def timed(fn):
        e0, e1 = ev(), ev()
        e0.record()
        fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    res = {"insert": 0.0, "query": 0.0, "delete": 0.0, "requery": 0.0}
    for step in range(args.warmup + args.steps):
        table.clear()
        fails.zero_()
        ti = timed(lambda: [table.bulk_insert(keys[a:a + batch], vals[a:a + batch], status[a:a + batch], fails)
                            for a in range(0, n, batch)])
        tq = timed(lambda: table.bulk_query(q, out, found))
        hits = int(found[: nq // 2].sum().item())
        fp = int(found[nq // 2:].sum().item())
        td = timed(lambda: table.bulk_remove(dels))
        tr = timed(lambda: table.bulk_query(q, out, found))
        if step >= args.warmup:
            for k, v in zip(res, (ti, tq, td, tr)):
                res[k] += v / args.steps
    inserted = int((status > 0).sum().item())
    in_backing = int((status == 2).sum().item())
    out_line = {
        "metric": "flat TCF: insert/query/delete Mops/s at 0.9 load",
        "value": round((n + 2 * nq + dels.numel()) / (sum(res.values()) / 1e3) / 1e6, 1),
        "unit": "Mops/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(sum(res.values()), 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32", "data": "synthetic (splitmix64 keys)",
        "config": {"workload": f"flat TCF 2^{args.log2_slots} + 1 % slots x 4 B (16/16 key_val_pair), "
                               f"B 16, tile {args.flat_tile}, double hashing, shortcut; {n} inserts, {nq} queries "
                               f"(50 % inserted, 50 % absent), delete {dels.numel()}, re-query"},
        "insert_mops": round(n / res["insert"] / 1e3, 1), "query_mops": round(nq / res["query"] / 1e3, 1),
        "delete_mops": round(dels.numel() / res["delete"] / 1e3, 1),
        "requery_mops": round(nq / res["requery"] / 1e3, 1),
        "insert_failures": int(fails.item()), "inserted": inserted, "in_backing": in_backing,
        "hit_queries_found": hits, "hit_queries": nq // 2, "fp_rate": fp / (nq // 2),
        "ms": {k: round(v, 3) for k, v in res.items()},
    }
    print(json.dumps(out_line), flush=True)


class ServeWorkload:
    """The serving workload: 64 resident tables (32 TCFs of 2^26-2^28 slots,
    32 double-hashing KV tables of 2^24-2^27 slots, tile 4), each 40 % full;
    requests with an op mix of 60 % query / 30 % insert / 10 % delete, the
    table of each chosen Zipf(1.1), keys in pinned host memory (the server
    uses them in place)."""

    def __init__(self, log):
        import numpy as np
        import torch
        import poggers_amd as pg
        self.np, self.torch, self.pg = np, torch, pg
        dev = torch.device("cuda", 0)
        rng = np.random.default_rng(2024)
        tcf_sizes = np.exp(rng.uniform(np.log(2.0 ** 26), np.log(2.0 ** 28), 32)).astype(np.int64)
        kv_sizes = np.exp(rng.uniform(np.log(2.0 ** 24), np.log(2.0 ** 27), 32)).astype(np.int64)
        specs = [("tcf", int(s)) for s in tcf_sizes] + [("kv", int(s)) for s in kv_sizes]
        # Zipf ranks go to the tables with the most free room (largest first)
        specs.sort(key=lambda x: -(x[1] * (4 if x[0] == "tcf" else 1)))
        self.specs = specs
        self.tables, self.npop = [], []
        tmp = torch.empty(1 << 27, dtype=torch.int64, device=dev)
        for i, (kind, s) in enumerate(specs):
            if kind == "tcf":
                t = pg.Table(pg.tcf_desc(s, tile=4))
            else:
                t = pg.Table(pg.kv_desc(s, bucket_size=16, tile=4, prober=pg.PROBE_DOUBLE, max_probes=20))
            n = int(0.4 * s)
            for a in range(0, n, 1 << 27):
                m = min(1 << 27, n - a)
                pg.gen_keys(1000 + i, a, m, tmp[:m])
                t.bulk_insert(tmp[:m], (tmp[:m] ^ 0x5A5A) if kind == "kv" else None)
            self.tables.append(t)
            self.npop.append(n)
        del tmp
        torch.cuda.synchronize()
        self.total_bytes = sum(t.host_bytes_in_use() for t in self.tables)
        self.fresh = [0] * len(self.tables)
        w = 1.0 / np.arange(1, len(self.tables) + 1) ** 1.1
        self.zipf_w = w / w.sum()
        log(f"serve: {len(self.tables)} tables, {self.total_bytes / 2**30:.1f} GiB resident")

    def describe(self):
        return (f"64 resident tables (32 TCF 2^26-2^28 slots, 32 KV 2^24-2^27 slots; "
                f"{self.total_bytes / 2**30:.1f} GiB), 40% full; op mix 60/30/10 query/insert/delete, "
                f"tables Zipf(1.1), keys in pinned host memory used in place; 4 streams")

    def make(self, nreq, lo, hi, seed):
        """nreq requests of log-uniform [lo, hi] keys: (table, op, keys, vals, out, flags, n)."""
        np, torch, pg = self.np, self.torch, self.pg
        from poggers_amd.workloads import keys_at
        r = np.random.default_rng(seed)
        reqs = []
        for _ in range(nreq):
            t = int(r.choice(len(self.tables), p=self.zipf_w))
            n = int(np.exp(r.uniform(np.log(lo), np.log(hi))))
            u = r.random()
            op = pg.OP_QUERY if u < 0.6 else (pg.OP_INSERT if u < 0.9 else pg.OP_REMOVE)
            if op == pg.OP_INSERT:
                keys = keys_at(5000 + t, np.arange(self.fresh[t], self.fresh[t] + n) + (1 << 40))
                self.fresh[t] += n
            else:
                keys = keys_at(1000 + t, r.integers(0, self.npop[t], n))
            kt = torch.empty(n, dtype=torch.int64, pin_memory=True)
            kt.numpy()[:] = keys.view(np.int64)
            kv = self.specs[t][0] == "kv"
            vals = out = None
            if op == pg.OP_INSERT:
                vals = torch.empty(n, dtype=torch.int64 if kv else torch.int16, pin_memory=True)
                vals.numpy()[:] = ((keys ^ np.uint64(0x5A5A)).view(np.int64) if kv
                                   else (keys & np.uint64(0xFFFF)).astype(np.uint16).view(np.int16))
            if op == pg.OP_QUERY:
                out = torch.empty(n, dtype=torch.int64 if kv else torch.int16, pin_memory=True)
            flags = torch.empty(n, dtype=torch.uint8, pin_memory=True)
            reqs.append((t, op, kt, vals, out, flags, n))
        return reqs

    def packed_of(self, reqs):
        return self.pg.Server.pack([(t, op, kt, vals, out, flags) for (t, op, kt, vals, out, flags, n) in reqs])

    def saturated(self, srv, reqs):
        """All requests offered at once, handed over 64 at a time while the
        engine runs (submission overlaps execution).  Returns (s, keys)."""
        packed = self.packed_of(reqs)
        self.torch.cuda.synchronize()
        t0 = time.perf_counter()
        done = 0
        for s0 in range(0, len(reqs), 64):
            srv.submit_many(packed, s0, min(64, len(reqs) - s0))
            done += len(srv.progress())
        while done < len(reqs):
            done += len(srv.progress())
        return time.perf_counter() - t0, sum(r[6] for r in reqs)

    @staticmethod
    def plain(srv, reqs):
        """Per-request plain launches: one kernel on the request's buffers, then a
        wait.  Returns (seconds, keys, each request's own time in s: its service
        time with no queueing, the plain path's best case latency)."""
        lat = []
        t0 = time.perf_counter()
        for (t, op, kt, vals, out, flags, n) in reqs:
            a = time.perf_counter()
            srv.run_plain(t, op, kt, vals, out, flags)
            lat.append(time.perf_counter() - a)
        return time.perf_counter() - t0, sum(r[6] for r in reqs), lat

    @staticmethod
    def plain_poisson(srv, reqs, rate, seed):
        """The same open-loop Poisson arrivals as poisson(), served by plain
        launches one request at a time in arrival order (a request waits for
        its arrival time, then for the requests before it).  Returns each
        request's latency in s."""
        import numpy as np
        r = np.random.default_rng(seed)
        arrive = np.cumsum(r.exponential(1.0 / rate, len(reqs)))
        lat = np.zeros(len(reqs))
        t0 = time.perf_counter()
        for q, (t, op, kt, vals, out, flags, n) in enumerate(reqs):
            while time.perf_counter() - t0 < arrive[q]:
                pass
            srv.run_plain(t, op, kt, vals, out, flags)
            lat[q] = time.perf_counter() - t0 - arrive[q]
        return lat

    def poisson(self, srv, reqs, rate, seed):
        """Open-loop Poisson arrivals at `rate` requests/s; each engine step first
        hands over every request that has arrived.  Returns (latencies in s,
        seconds from the first arrival to the last completion)."""
        np = self.np
        r = np.random.default_rng(seed)
        arrive = np.cumsum(r.exponential(1.0 / rate, len(reqs)))
        lat = np.zeros(len(reqs))
        packed = self.packed_of(reqs)
        tick2i = {}
        i = done = 0
        last = 0.0
        t0 = time.perf_counter()
        while done < len(reqs):
            now = time.perf_counter() - t0
            j = int(np.searchsorted(arrive, now, side="right"))
            if j > i:
                for k, tk in enumerate(srv.submit_many(packed, i, j - i)):
                    tick2i[tk] = i + k
                i = j
            fin = srv.progress()
            if fin:
                last = time.perf_counter() - t0
                for tk in fin:
                    q = tick2i[tk]
                    lat[q] = last - arrive[q]
                done += len(fin)
        return lat, last - arrive[0]


def serving_point(log, nreq=4000, lo=1000, hi=64 << 10):
    """The serving leg of the default bench line, on the serving workload with
    requests of 1K-64K keys: per-request plain launches, the saturated served
    rate, and open-loop Poisson arrivals at fractions of the saturated request
    rate.  The operating point is the highest offered rate whose p99 latency
    is at most 2 ms; its served rate (keys completed from the first arrival to
    the last completion) is compared with the plain launches."""
    import numpy as np
    import poggers_amd as pg
    w = ServeWorkload(log)
    srv = pg.Server(w.tables, budget_bytes=4 << 30, nstreams=4)
    reqs = w.make(nreq, lo, hi, 7 + hi)
    sizes = np.array([r[6] for r in reqs])
    w.saturated(srv, reqs[: nreq // 10])  # warm up
    ts, keys = w.saturated(srv, reqs)
    tp, _, plat = w.plain(srv, reqs)
    plain_mops = keys / tp / 1e6
    sat_rate = len(reqs) / ts
    sweep = []
    for frac in (0.5, 0.7, 0.85, 1.0):
        lat, span = w.poisson(srv, reqs, frac * sat_rate, 11)
        served = keys / span / 1e6
        sweep.append({"offered_frac_of_saturated": frac, "rate_req_per_s": round(frac * sat_rate, 1),
                      "served_mops": round(served, 1), "speedup_vs_plain": round(served / plain_mops, 2),
                      "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
                      "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3)})
        log(f"serve poisson {frac}: {sweep[-1]}")
    ok = [p for p in sweep if p["p99_ms"] <= 2.0]
    point = max(ok, key=lambda p: p["served_mops"]) if ok else min(sweep, key=lambda p: p["p99_ms"])
    # latency at a load the plain path sustains: 70 % of its request rate,
    # the same arrivals served both ways
    prate = 0.7 * len(reqs) / tp
    lat_p = w.plain_poisson(srv, reqs, prate, 13)
    lat_s, _ = w.poisson(srv, reqs, prate, 13)
    at_plain = {"rate_req_per_s": round(prate, 1),
                "plain_p50_ms": round(float(np.percentile(lat_p, 50)) * 1e3, 3),
                "plain_p99_ms": round(float(np.percentile(lat_p, 99)) * 1e3, 3),
                "served_p50_ms": round(float(np.percentile(lat_s, 50)) * 1e3, 3),
                "served_p99_ms": round(float(np.percentile(lat_s, 99)) * 1e3, 3)}
    log(f"serve at 70% of plain capacity: {at_plain}")
    srv.free()
    out = {
        "workload": w.describe() + f"; {len(reqs)} requests of log-uniform {lo}-{hi} keys "
                                   f"(mean {sizes.mean():.0f})",
        "plain_mops": round(plain_mops, 1),
        # each plain request timed alone (no queueing): the plain path's best case
        "plain_p50_ms": round(float(np.percentile(plat, 50)) * 1e3, 3),
        "plain_p99_ms": round(float(np.percentile(plat, 99)) * 1e3, 3),
        "saturated_mops": round(keys / ts / 1e6, 1),
        "served_mops": point["served_mops"],
        "speedup_vs_plain": point["speedup_vs_plain"],
        "p50_ms": point["p50_ms"],
        "p99_ms": point["p99_ms"],
        "rate_req_per_s": point["rate_req_per_s"],
        "at_70pct_plain_capacity": at_plain,
        "sweep": sweep,
    }
    del reqs, w
    return out


def run_serve(args):
    """The serving workload (ServeWorkload).  For request sizes log-uniform in
    [1K, 4M] and in [1K, 64K]: saturated throughput of the served loop (all
    requests queued at once) against per-request plain launches, and p50/p99
    latency under Poisson arrivals at half the served throughput; then both
    again with per-table request ordering."""
    import numpy as np
    import torch
    import poggers_amd as pg

    torch.cuda.set_device(0)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    w = ServeWorkload(log)
    results = {}
    srv = pg.Server(w.tables, budget_bytes=4 << 30, nstreams=4)
    for name, lo, hi, nreq in (("mix_1K_4M", 1000, 4 << 20, args.requests),
                               ("small_1K_64K", 1000, 64 << 10, 4 * args.requests)):
        reqs = w.make(nreq, lo, hi, 7 + hi)
        w.saturated(srv, reqs[: max(1, nreq // 10)])  # warm up
        ts, keys = w.saturated(srv, reqs)
        tp, _, plat = w.plain(srv, reqs)
        served = keys / ts / 1e6
        plain_mops = keys / tp / 1e6
        rate = 0.5 * (len(reqs) / ts)
        lat, _ = w.poisson(srv, reqs, rate, 11)
        sizes = np.array([r[6] for r in reqs])
        small = sizes <= 64 * 1024
        results[name] = {
            "requests": len(reqs), "keys": int(keys), "mean_keys": round(float(sizes.mean()), 1),
            "served_mops": round(served, 1), "plain_mops": round(plain_mops, 1),
            "plain_p50_ms": round(float(np.percentile(plat, 50)) * 1e3, 3),
            "plain_p99_ms": round(float(np.percentile(plat, 99)) * 1e3, 3),
            "speedup": round(served / plain_mops, 2),
            "poisson_rate_req_per_s": round(rate, 1),
            "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "p99_ms_requests_le_64K": round(float(np.percentile(lat[small], 99)) * 1e3, 3) if small.any() else None,
        }
        # the same requests with per-table ordering (writes in submission order)
        srv.set_order(pg.ORDER_TABLE)
        to, _ = w.saturated(srv, reqs)
        rate_o = 0.5 * (len(reqs) / to)
        lat_o, _ = w.poisson(srv, reqs, rate_o, 11)
        srv.set_order(pg.ORDER_NONE)
        results[name]["ordered"] = {
            "served_mops": round(keys / to / 1e6, 1), "speedup": round(keys / to / 1e6 / plain_mops, 2),
            "poisson_rate_req_per_s": round(rate_o, 1),
            "p50_ms": round(float(np.percentile(lat_o, 50)) * 1e3, 3),
            "p99_ms": round(float(np.percentile(lat_o, 99)) * 1e3, 3),
            "p99_ms_requests_le_64K": round(float(np.percentile(lat_o[small], 99)) * 1e3, 3) if small.any() else None,
        }
        log(f"serve {name}: {results[name]}")
        del reqs
    srv.free()
    out = {
        "metric": "served Mops/s and p50/p99 latency vs per-request plain launches",
        "value": results["mix_1K_4M"]["served_mops"], "unit": "Mops/s",
        "n_gpus": 1, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32/u64",
        "data": "synthetic (splitmix64 keys in pinned host memory)",
        "config": {"workload": w.describe() + ", 4 GiB staging budget"},
        "results": results,
    }
    print(json.dumps(out), flush=True)


PLAIN_TILE, PLAIN_BUCKET = 4, 16   # the reference's TCF shape (tests/src/delete_tests.cu:138-152)


def run_batched(args):
    """4096 independent TCFs, slots log-uniform in [2^12, 2^22], each filled to
    85 % and queried with as many keys (half present, half absent), all in one
    hipGraph replay of a persistent insert kernel and a persistent query
    kernel; tile width, bucket size, work chunk and keys per tile in flight
    picked by the autotuner on a 512-table sample.  To separate the gain of
    batching from that of the shape, the graph and plain per-table launches
    (a clear, an insert and a query launch per table) are both measured at
    the tuned shape and at the reference's TCF shape (4-lane tiles, 16-slot
    buckets)."""
    import numpy as np
    import torch
    import poggers_amd as pg
    from poggers_amd.autotune import tune_batch
    from poggers_amd.workloads import absent_mask, batched_workload, log_uniform_sizes

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    sizes = log_uniform_sizes(args.tables, 12, 22, seed=11)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    sample = sizes[:min(512, len(sizes))]
    s_ins, s_off, s_q, s_qoff, s_hoff, s_npres = batched_workload(sample, 0.85, 77, dev)
    tune = tune_batch(sample, s_ins, s_off, s_q, s_qoff, chunks=(512, 2048, 8192), reps=2,
                      absent=absent_mask(s_hoff, s_npres, dev), log=log)
    del s_ins, s_off, s_q, s_qoff
    best = tune["best"]
    ins, off, q, qoff, hoff, npres = batched_workload(sizes, 0.85, 77, dev)
    total = ins.numel()
    present = ~absent_mask(hoff, npres, dev)
    found = torch.zeros(total, dtype=torch.uint8, device=dev)
    fails = torch.zeros(1, dtype=torch.int64, device=dev)

    def quality():
        return {"false_negatives": int((found[present] == 0).sum().item()),
                "fp_rate": float(found[~present].float().mean().item()),
                "insert_failures": int(fails.item())}

    def graph(tile, bucket, chunk, items):
        desc = pg.tcf_desc(1 << 12, 1 << 6, bucket_size=bucket, tile=tile)
        batch = pg.Batch(desc, sizes)
        batch.set_launch(chunk)
        batch.set_items(items)
        stream = torch.cuda.Stream()
        batch.capture(ins, off, q, qoff, None, found, fails, True, stream)
        for _ in range(args.warmup):
            batch.replay(stream)
        stream.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        for _ in range(args.steps):
            fails.zero_()
            batch.replay(stream)
        e1.record(stream)
        stream.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        out = {"ms": round(ms, 3), "gpu_ms_per_replay": round(e0.elapsed_time(e1) / args.steps, 3), **quality()}
        batch.free()
        return out

    def plain(tile, bucket):
        tables = [pg.Table(pg.tcf_desc(int(s), max(int(s) // 100, 8), bucket_size=bucket, tile=tile))
                  for s in sizes]
        torch.cuda.synchronize()

        def once():
            fails.zero_()
            for t, tb in enumerate(tables):
                a, z = int(hoff[t]), int(hoff[t + 1])
                tb.clear()
                tb.bulk_insert(ins[a:z], None, None, fails)
                tb.bulk_query(q[a:z], None, found[a:z])
        once()
        torch.cuda.synchronize()
        reps = max(1, args.steps // 2)
        p0 = time.perf_counter()
        for _ in range(reps):
            once()
        torch.cuda.synchronize()
        out = {"ms": round((time.perf_counter() - p0) * 1e3 / reps, 3), **quality()}
        for tb in tables:
            tb.free_on_device()
        return out

    shapes = {"tuned": (best["tile"], best["bucket"]), "reference": (PLAIN_TILE, PLAIN_BUCKET)}
    results = {}
    for name, (tile, bucket) in shapes.items():
        g = graph(tile, bucket, best["chunk"], best["items"] if name == "tuned" else 1)
        p = plain(tile, bucket)
        results[name] = {"tile": tile, "bucket": bucket, "graph": g, "plain_launches": p,
                         "speedup_of_batching": round(p["ms"] / g["ms"], 2)}
        log(f"batched {name}: {results[name]}")
    head = results["tuned"]["graph"]
    ops = 2 * total
    out = {
        "metric": "batched TCF insert+query Mops/s (4096 uneven filters, one hipGraph)",
        "value": round(ops / (head["ms"] / 1e3) / 1e6, 1),
        "unit": "Mops/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": head["ms"],
        "gpu_ms_per_replay": head["gpu_ms_per_replay"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (splitmix64 uint64 keys generated on the GPU)",
        "config": {"workload": f"{len(sizes)} TCFs, slots log-uniform in [2^12, 2^22] "
                               f"({int(sizes.sum())} slots), insert to 85% ({total} keys), query {total} keys "
                               f"(half present, half absent)",
                   "tile": best["tile"], "bucket": best["bucket"], "chunk": best["chunk"],
                   "items": best["items"]},
        "false_negatives": head["false_negatives"], "fp_rate": head["fp_rate"],
        "insert_failures": head["insert_failures"],
        "shapes": results,
        "speedup_vs_plain_launches_same_shape": results["tuned"]["speedup_of_batching"],
        "speedup_vs_plain_launches_reference_shape": round(results["reference"]["plain_launches"]["ms"] / head["ms"], 2),
        "autotune": tune,
    }
    print(json.dumps(out), flush=True)


CAL_COPY_BYTES = 1 << 30          # --calibrate-copy: bytes read and written by the copy
CAL_RANDOM_READS = 1 << 24        # --calibrate-copy: random 64 B reads
_PMC = "python3 bench.py --steps 1 --warmup 0 --no-cpu --no-roofline --no-serve --calibrate-copy"
PROFILE_CMDS = {
    "stats": "rocprofv3 --kernel-trace --stats -f csv -d gpurun_out/prof_stats -o tcf -- "
             "python3 bench.py --steps 3 --warmup 1 --no-cpu --no-roofline --no-serve",
    "fetch": f"rocprofv3 --pmc FETCH_SIZE -f csv -d gpurun_out/pmc_fetch -o tcf -- {_PMC}",
    "write": f"rocprofv3 --pmc WRITE_SIZE -f csv -d gpurun_out/pmc_write -o tcf -- {_PMC}",
    "tcc": "rocprofv3 --pmc TCC_HIT_sum TCC_MISS_sum TCC_ATOMIC_sum TCC_EA0_ATOMIC_sum -f csv "
           f"-d gpurun_out/pmc_tcc -o tcf -- {_PMC}",
    "rdreq": "rocprofv3 --pmc TCC_EA0_RDREQ_sum TCC_EA0_RDREQ_32B_sum TCC_EA0_RDREQ_64B_sum "
             f"TCC_EA0_RDREQ_128B_sum -f csv -d gpurun_out/pmc_rdreq -o tcf -- {_PMC}",
    "wrreq": "rocprofv3 --pmc TCC_EA0_WRREQ_sum TCC_EA0_WRREQ_64B_sum TCC_EA0_RDREQ_DRAM_sum "
             f"TCC_EA0_WRREQ_DRAM_sum -f csv -d gpurun_out/pmc_wrreq -o tcf -- {_PMC}",
    "summary": "python3 bench.py --summarize-profiles gpurun_out/prof_stats gpurun_out/pmc_fetch "
               "gpurun_out/pmc_write gpurun_out/pmc_tcc gpurun_out/pmc_rdreq gpurun_out/pmc_wrreq",
}


def _csv_rows(directory, suffix):
    import csv
    import glob
    paths = sorted(glob.glob(os.path.join(directory, "**", "*" + suffix), recursive=True))
    if not paths:
        raise SystemExit(f"no *{suffix} under {directory}")
    with open(paths[0], newline="") as f:
        return list(csv.DictReader(f))


def _short(name, n=110):
    return name if len(name) <= n else name[:n - 3] + "..."


def summarize_profiles(stats_dir, fetch_dir, write_dir, tcc_dir, rdreq_dir, wrreq_dir):
    """Condense rocprofv3 csv output into the committed text summaries:
    profiles/tcf_2p28_kernel_stats.md (per-kernel launch statistics and
    per-dispatch memory counters) and profiles/tcf_2p28_traffic.json (the
    insert kernel's bytes between L2 and memory per launch, read by the bench
    line's roofline.traffic).  The bytes come from the request counters by
    size (TCC_EA0_RDREQ_32B/64B/128B, TCC_EA0_WRREQ/_64B); the runs include
    kernels of known traffic (a 1 GiB streaming read + write, 2^24 random
    64 B reads) that check them, and FETCH_SIZE / WRITE_SIZE are reported
    beside them with the factor the copy kernel gives them."""
    import re
    stats = _csv_rows(stats_dir, "_kernel_stats.csv")
    runs = {k: _csv_rows(d, "_counter_collection.csv")
            for k, d in (("fetch", fetch_dir), ("write", write_dir), ("tcc", tcc_dir), ("rd", rdreq_dir),
                         ("wr", wrreq_dir))}
    avg_ns = {r["Name"]: float(r["AverageNs"]) for r in stats}

    def per_dispatch(rows, pred):
        """{counter: mean value per dispatch} over dispatches of kernels matching pred."""
        per, names, disp = {}, set(), {}
        for r in rows:
            n = r["Kernel_Name"]
            if pred(n):
                names.add(n)
                per.setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
                disp[r["Counter_Name"]] = disp.get(r["Counter_Name"], 0) + 1
        return {c: sum(v) / len(v) for c, v in per.items()}, names, max(disp.values()) if disp else 0

    kinds = {
        # the timed kernels: fourth template argument false (true marks the
        # accounting variant)
        "insert": lambda n: re.search(r"k_tcf_insert<\d+, \d+, \d+, false", n) is not None,
        "query": lambda n: re.search(r"k_tcf_query<\d+, \d+, \d+, false", n) is not None,
        "copy 1 GiB": lambda n: "neg" in n and "double" in n,
        "random 64 B x 2^24": lambda n: "k_random_sector" in n,
    }
    m = {}
    for kind, pred in kinds.items():
        vals, names, nd = {}, set(), 0
        for run in runs.values():
            v, nm, d = per_dispatch(run, pred)
            vals.update(v)
            names |= nm
            nd = max(nd, d)
        if names:
            m[kind] = {"values": vals, "names": sorted(names), "dispatches": nd}
    if "insert" not in m or "copy 1 GiB" not in m:
        raise SystemExit("the insert kernel and the calibration copy must be in the PMC runs")

    def rd_bytes(v):
        return 32 * v["TCC_EA0_RDREQ_32B_sum"] + 64 * v["TCC_EA0_RDREQ_64B_sum"] + 128 * v["TCC_EA0_RDREQ_128B_sum"]

    def wr_bytes(v):
        return 32 * (v["TCC_EA0_WRREQ_sum"] - v["TCC_EA0_WRREQ_64B_sum"]) + 64 * v["TCC_EA0_WRREQ_64B_sum"]

    ins = m["insert"]
    name = ins["names"][0]
    tile = int(re.search(r"k_tcf_insert<(\d+),", name).group(1))
    grid = [int(r["Grid_Size"]) for r in runs["fetch"] if r["Kernel_Name"] == name]
    keys_per_launch = sum(grid) / len(grid) / tile
    cp = m["copy 1 GiB"]["values"]
    fetch_factor = CAL_COPY_BYTES / (cp["FETCH_SIZE"] * 1024.0)
    write_factor = CAL_COPY_BYTES / (cp["WRITE_SIZE"] * 1024.0)
    table = {}
    for kind, d in m.items():
        v = d["values"]
        rb, wb = rd_bytes(v), wr_bytes(v)
        hit, miss = v.get("TCC_HIT_sum", 0.0), v.get("TCC_MISS_sum", 0.0)
        ns = next((avg_ns[n] for n in d["names"] if n in avg_ns), None)
        table[kind] = {
            "kernel_name": d["names"][0], "dispatches": d["dispatches"],
            "read_requests": {sz: int(v[f"TCC_EA0_RDREQ_{sz}_sum"]) for sz in ("32B", "64B", "128B")},
            "read_requests_total": int(v["TCC_EA0_RDREQ_sum"]),
            "read_bytes": int(rb), "write_bytes": int(wb),
            "read_requests_to_dram": int(v["TCC_EA0_RDREQ_DRAM_sum"]),
            "write_requests_to_dram": int(v["TCC_EA0_WRREQ_DRAM_sum"]),
            "fetch_size_kib": round(v["FETCH_SIZE"], 1), "write_size_kib": round(v["WRITE_SIZE"], 1),
            "fetch_write_scaled_by_copy_bytes": int(v["FETCH_SIZE"] * 1024 * fetch_factor +
                                                    v["WRITE_SIZE"] * 1024 * write_factor),
            "tcc_hit_rate": round(hit / (hit + miss), 4) if hit + miss else None,
            "atomics_per_launch": int(v.get("TCC_ATOMIC_sum", 0.0)),
            "atomics_to_memory_per_launch": int(v.get("TCC_EA0_ATOMIC_sum", 0.0)),
            "atomics_per_s": round(v.get("TCC_ATOMIC_sum", 0.0) / (ns * 1e-9), 1) if ns else None,
            "average_us": round(ns / 1e3, 2) if ns else None,
        }
    t = table["insert"]
    traffic = {
        "kernel": "insert", "kernel_name": name, "dispatches": t["dispatches"],
        "keys_per_launch": int(keys_per_launch),
        "hbm_bytes_per_launch": t["read_bytes"] + t["write_bytes"],
        "read_bytes_per_launch": t["read_bytes"], "write_bytes_per_launch": t["write_bytes"],
        "bytes_per_key": round((t["read_bytes"] + t["write_bytes"]) / keys_per_launch, 2),
        "copy_check": {"read_bytes": table["copy 1 GiB"]["read_bytes"],
                       "write_bytes": table["copy 1 GiB"]["write_bytes"], "expected_each": CAL_COPY_BYTES},
        "fetch_size_factor_from_copy": round(fetch_factor, 4), "write_size_factor_from_copy": round(write_factor, 4),
        "hbm_bytes_per_launch_fetch_write_scaled": t["fetch_write_scaled_by_copy_bytes"],
        "commands": [PROFILE_CMDS[k] for k in ("fetch", "write", "tcc", "rdreq", "wrreq", "summary")],
        "tcc": {"insert": {k: table["insert"][k] for k in ("tcc_hit_rate", "atomics_per_launch",
                                                           "atomics_to_memory_per_launch", "atomics_per_s")},
                "query": ({k: table["query"][k] for k in ("tcc_hit_rate", "atomics_per_launch",
                                                         "atomics_to_memory_per_launch", "atomics_per_s")}
                          if "query" in table else None)},
        "kernels": table,
    }
    lines = ["# TCF 2^28 slots, 90 % load: rocprofv3 summaries", "",
             "Commands (run from the repository root on one MI355X):", ""]
    lines += [f"    {PROFILE_CMDS[k]}" for k in ("stats", "fetch", "write", "tcc", "rdreq", "wrreq", "summary")]
    lines += ["", "## Kernel statistics (`--kernel-trace --stats`)", "",
              "Kernels whose fourth template argument is `true` are the accounting variants,",
              "run once, untimed, to count bucket reads; the timed kernels are the `false` ones.",
              "A sixth argument `true` marks the plain insert built for 8 waves per SIMD.", "",
              "| kernel | calls | average us | total ms | % |", "|---|---|---|---|---|"]
    for r in sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))[:12]:
        lines.append(f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['Percentage']):.2f} |")
    lines += ["", "## Traffic between L2 (TCC) and memory, per dispatch", "",
              "Bytes from the request counters by size: read = 32 x RDREQ_32B + 64 x RDREQ_64B + "
              "128 x RDREQ_128B, write = 32 x (WRREQ - WRREQ_64B) + 64 x WRREQ_64B.  Requests 'to DRAM' go to "
              "the local memory controllers (no counter separates the memory-side Infinity Cache's hits "
              "from DRAM).  FETCH_SIZE / WRITE_SIZE (KiB) are shown with the factor that makes the 1 GiB "
              "copy read 1 GiB.", "",
              "| kernel | dispatches | read req 32/64/128 B | read bytes | write bytes | rd / wr req to DRAM | "
              "FETCH_SIZE x factor + WRITE_SIZE x factor | L2 hit rate | atomics (to memory) |",
              "|---|---|---|---|---|---|---|---|---|"]
    for kind, d in table.items():
        rq = d["read_requests"]
        lines.append(f"| {kind} `{_short(d['kernel_name'], 50)}` | {d['dispatches']} | "
                     f"{rq['32B']:,} / {rq['64B']:,} / {rq['128B']:,} | {d['read_bytes']:,} | {d['write_bytes']:,} | "
                     f"{d['read_requests_to_dram']:,} / {d['write_requests_to_dram']:,} | "
                     f"{d['fetch_write_scaled_by_copy_bytes']:,} | {d['tcc_hit_rate']} | "
                     f"{d['atomics_per_launch']:,} ({d['atomics_to_memory_per_launch']:,}) |")
    ab = traffic["hbm_bytes_per_launch"]
    lines += ["", f"Factors from the copy: FETCH_SIZE x {fetch_factor:.3f}, WRITE_SIZE x {write_factor:.3f}.",
              "", f"Insert kernel, {int(keys_per_launch):,} keys per launch: {ab:,} bytes between L2 and memory "
              f"per launch ({traffic['bytes_per_key']} B per key); the bench line's `roofline.traffic`.", ""]
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    with open(os.path.join(ROOT, "profiles", "tcf_2p28_kernel_stats.md"), "w") as f:
        f.write("\n".join(lines))
    with open(os.path.join(ROOT, "profiles", "tcf_2p28_traffic.json"), "w") as f:
        json.dump(traffic, f, indent=1)
    print("\n".join(lines))


WORKLOAD_PROFILES = {
    # name: (title, rocprofv3 command); summarised with --summarize-kernel-stats DIR NAME
    "kv_2p30": ("KV table 2^30 slots x 16 B, 0.9 load, double hashing, B 16, tile 4",
                "rocprofv3 --kernel-trace --stats -f csv -d gpurun_out/prof_kv -o kv -- "
                "python3 bench.py --workload kv --steps 2 --warmup 1"),
    "batched_4096": ("4096 batched TCFs (one hipGraph per step; the autotuner's sweep runs first, "
                     "so every swept shape appears; the shape it picked for the bench line is listed "
                     "after the top kernels)",
                     "rocprofv3 --kernel-trace --stats -f csv -d gpurun_out/prof_batch -o batch -- "
                     "python3 bench.py --workload batched --steps 3 --warmup 1 > gpurun_out/prof_batch.log"),
    "flat_tcf_2p28": ("flat TCF 2^28 + 1 % slots (16/16 key-value slots, B 16, tile 4), 0.9 load",
                      "rocprofv3 --kernel-trace --stats -f csv -d gpurun_out/prof_flat -o flat -- "
                      "python3 bench.py --workload flat --steps 2 --warmup 1"),
    "serve_64_tables": ("serving loop over 64 resident tables against per-request plain launches",
                        "rocprofv3 --kernel-trace --stats -f csv -d gpurun_out/prof_serve -o serve -- "
                        "python3 bench.py --workload serve --requests 200"),
}


def summarize_kernel_stats(stats_dir, name):
    """profiles/<name>_kernel_stats.md from a `rocprofv3 --kernel-trace
    --stats` csv directory of one of the WORKLOAD_PROFILES runs."""
    title, cmd = WORKLOAD_PROFILES[name]
    stats = _csv_rows(stats_dir, "_kernel_stats.csv")
    lines = [f"# {title}: rocprofv3 kernel statistics", "",
             "Commands (run from the repository root on one MI355X):", "", f"    {cmd}",
             f"    python3 bench.py --summarize-kernel-stats {stats_dir} {name}", "",
             "| kernel | calls | average us | total ms | % |", "|---|---|---|---|---|"]
    ranked = sorted(stats, key=lambda r: -float(r["TotalDurationNs"]))
    row = lambda r: (f"| `{_short(r['Name'])}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.1f} | "
                     f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['Percentage']):.2f} |")
    lines += [row(r) for r in ranked[:14]]
    if name == "batched_4096":
        # the shape the autotuner picked: the bench line (in the run's
        # log, next to the csv directory) names it
        import re
        shape = None
        for path in (os.path.join(stats_dir, "..", "prof_batch.log"),):
            try:
                with open(path) as f:
                    for line in f:
                        if line.startswith("{"):
                            c = json.loads(line)["config"]
                            shape = (c["tile"], c["bucket"])
            except (OSError, ValueError, KeyError):
                pass
        if shape:
            pat = re.compile(rf"k_batch_persistent<{shape[0]}, {shape[1]}, ")
            lines += ["", f"The picked shape (tile {shape[0]}, bucket {shape[1]}), sweep replays and "
                          "timed replays together:", "",
                      "| kernel | calls | average us | total ms | % |", "|---|---|---|---|---|"]
            lines += [row(r) for r in ranked if pat.search(r["Name"])]
    lines.append("")
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    with open(os.path.join(ROOT, "profiles", f"{name}_kernel_stats.md"), "w") as f:
        f.write("\n".join(lines))
    print("\n".join(lines))


def load_traffic(path):
    """HBM bytes per launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), committed as a small JSON summary; None when absent."""
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return None


RR_OPS = (("read", 0, 64), ("read_write", 2, 128), ("read_cas", 1, 128))
RR_SHAPES = (0, 1, 2)          # 4 lanes x 16 B, 1 lane x 64 B, 16 lanes x 4 B
RR_SECTORS = (1, 2, 4, 8)      # sectors in flight per tile
RR_WAVES = (1, 2, 3, 4, 6, 0)  # waves per SIMD of a grid-striding grid; 0 = a grid covering the work


def random_roofline(pg, torch, dev, stream, footprint, n_reads=1 << 26, reps=3):
    """Random 64 B-sector bandwidth over a buffer of `footprint` bytes: the best
    variant per op kind, and for the best tile shape the whole
    sectors-in-flight x waves-per-SIMD grid (GB/s), so the position of the
    best variant in the sweep is on record."""
    buf = torch.zeros(footprint // 8, dtype=torch.int64, device=dev)
    sink = torch.zeros(1, dtype=torch.int64, device=dev)
    ev = lambda: torch.cuda.Event(enable_timing=True)
    out = {"best_variant": {}, "grid": {}}
    for name, wop, per_op in RR_OPS:
        best, grids = 0.0, {}
        for shape in RR_SHAPES:
            grid = {}
            for sec in RR_SECTORS:
                for w in RR_WAVES:
                    v = sec.bit_length()  # 1 << (v - 1) sectors
                    op = wop | (v << 4) | (shape << 9) | (w << 11)
                    pg.random_sector_bench(buf, n_reads, op, 7, sink, stream)
                    e0, e1 = ev(), ev()
                    e0.record(stream)
                    for r in range(reps):
                        pg.random_sector_bench(buf, n_reads, op, 8 + r, sink, stream)
                    e1.record(stream)
                    torch.cuda.synchronize()
                    gbs = n_reads * per_op / (e0.elapsed_time(e1) / reps / 1e3) / 1e9
                    grid[f"s{sec}_w{w}"] = round(gbs, 1)
                    if gbs > best:
                        best = gbs
                        out["best_variant"][name] = {"tile_shape": shape, "sectors_in_flight": sec,
                                                     "waves_per_simd": w or "grid"}
            grids[shape] = grid
        out[name] = round(best, 1)
        out["grid"][name] = grids[out["best_variant"][name]["tile_shape"]]
    del buf
    return out


def cpu_baseline(log2_slots=28, runs=3, sample_shift=4, query_log2=22):
    """The CPU oracle on the headline workload, single-threaded and pinned to
    one core, on a stratified sample that keeps the run near 30 s.

    Workload: the 2^28-slot TCF + 1 % backing filled to 90 % in batches of
    2^24, then 2^26 present + 2^26 absent queries.  In every insert batch the
    multi-threaded oracle (every core this process may use) inserts the first
    15/16 of the keys untimed, then the single-threaded oracle, pinned to one
    core with sched_setaffinity, inserts the last 1/16 timed, so the timed
    inserts cover every load level of the fill; at 90 % load it then times
    2^22 present + 2^22 absent queries.  The per-op times scale to the full
    op counts, and the rate is the whole workload's ops over that time; the
    median of `runs` runs.  `multi_thread` is the untimed part's rate."""
    import numpy as np
    import oracle
    oracle.build()
    nslots = 1 << log2_slots
    levels = [oracle.level(nslots, oracle.SCHEME_SHORTCUT, 16),
              oracle.level(nslots // 100, oracle.SCHEME_LINEAR_BUCKETS, 8, prober=oracle.PROBE_LINEAR,
                           max_probes=4000)]
    t = oracle.Table(oracle.REP_FP16_V16, levels, seed=42)
    n = int(0.9 * nslots)
    batch = 1 << 24
    nq_full = 1 << (log2_slots - 2)
    nq = 1 << query_log2
    keys = oracle.gen_keys(1234, 0, n)
    absent = oracle.gen_keys(1234, 1 << 40, nq)
    present = keys[::max(1, n // nq)][:nq].copy()   # present keys sampled across the insert order
    allowed = sorted(os.sched_getaffinity(0))
    threads = len(allowed)
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = max(1, min(threads, int(os.environ["OMP_NUM_THREADS"])))
    core = allowed[-1]
    rates, ins_rates, q_rates, mt = [], [], [], []
    used = 1
    for _ in range(runs):
        t.clear()
        t_ins = t_mt = 0.0
        n_ins = n_mt = 0
        for s0 in range(0, n, batch):
            s1 = min(n, s0 + batch)
            cut = s1 - max(1, (s1 - s0) >> sample_shift)
            m0 = time.perf_counter()
            _, used = t.bulk_insert_mt(keys[s0:cut], threads=threads)
            t_mt += time.perf_counter() - m0
            n_mt += cut - s0
            os.sched_setaffinity(0, {core})
            c0 = time.perf_counter()
            t.bulk_insert(keys[cut:s1])
            t_ins += time.perf_counter() - c0
            os.sched_setaffinity(0, allowed)
            n_ins += s1 - cut
        os.sched_setaffinity(0, {core})
        c0 = time.perf_counter()
        t.bulk_query(present)
        t.bulk_query(absent)
        t_q = time.perf_counter() - c0
        os.sched_setaffinity(0, allowed)
        per_ins, per_q = t_ins / n_ins, t_q / (2 * nq)
        total = per_ins * n + per_q * 2 * nq_full
        rates.append((n + 2 * nq_full) / total / 1e6)
        ins_rates.append(1 / per_ins / 1e6)
        q_rates.append(1 / per_q / 1e6)
        mt.append(n_mt / t_mt / 1e6)
    med = lambda v: float(np.median(v))
    return {
        "value": round(med(rates), 3),
        "unit": "Mops/s",
        "cores": 1,
        "kind": "port",
        "sample": f"TCF 2^{log2_slots} slots + 1% backing filled to 90% in batches of 2^24 (the headline "
                  f"workload); single-threaded oracle pinned to one core timed on the last 1/{1 << sample_shift} "
                  f"of every insert batch ({n_ins} inserts, the rest inserted untimed by the multi-threaded "
                  f"oracle) and on 2^{query_log2} present + 2^{query_log2} absent queries at 90% load, "
                  f"per-op times scaled to {n} inserts + {2 * nq_full} queries; median of {runs} runs",
        "insert_mops": round(med(ins_rates), 3),
        "query_mops": round(med(q_rates), 3),
        "runs": [round(r, 3) for r in rates],
        "pinned_core": core,
        "multi_thread": {"insert_mops": round(med(mt), 3), "cores": used,
                         "host_threads_online": oracle.hw_threads()},
        "cpu_model": _cpu_model(),
    }


def _cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or "" where it is not readable."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return ""


def main():
    args = parse()
    if args.summarize_profiles:
        if len(args.summarize_profiles) != 6:
            raise SystemExit("--summarize-profiles STATS FETCH WRITE TCC RDREQ WRREQ (rocprofv3 csv directories)")
        summarize_profiles(*args.summarize_profiles)
        return
    if args.summarize_kernel_stats:
        summarize_kernel_stats(*args.summarize_kernel_stats)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU: run with --gpus 1 (multi-GPU sharding "
                                   "is a design note only)"}))
        sys.exit(2)
    if args.workload == "batched":
        run_batched(args)
        return
    if args.workload == "flat":
        return run_flat(args)
    if args.workload == "kv":
        run_kv(args)
        return
    if args.workload == "serve":
        run_serve(args)
        return

    import torch
    import poggers_amd as pg

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    stream = torch.cuda.current_stream()

    nslots = 1 << args.log2_slots
    backing = nslots // 100
    desc = pg.tcf_desc(nslots, backing, bucket_size=16, tile=args.tile, prober=pg.PROBE_SPLIT64,
                       hasher=pg.HASH_FMIX64)
    table = pg.Table(desc, seed=42)
    table.set_items(0, args.query_items)
    n_ins = int(0.9 * nslots)
    batch = 1 << 24 if args.log2_slots >= 26 else max(1, n_ins // 8)
    nq = 1 << (args.log2_slots - 2)
    seed = 1234
    keys = torch.empty(n_ins, dtype=torch.int64, device=dev)
    absent = torch.empty(nq, dtype=torch.int64, device=dev)
    pg.gen_keys(seed, 0, n_ins, keys)
    pg.gen_keys(seed, 1 << 40, nq, absent)
    # present-key queries sample the whole insert order (every stride-th key):
    # early keys sit in their first choice far more often than late ones
    stride = max(1, n_ins // nq)
    present = keys[::stride][:nq].contiguous()
    found_p = torch.empty(nq, dtype=torch.uint8, device=dev)
    found_a = torch.empty(nq, dtype=torch.uint8, device=dev)
    vals_p = torch.empty(nq, dtype=torch.int16, device=dev)
    vals_a = torch.empty(nq, dtype=torch.int16, device=dev)
    fails = torch.zeros(1, dtype=torch.int64, device=dev)
    batches = [(s, min(s + batch, n_ins)) for s in range(0, n_ins, batch)]
    torch.cuda.synchronize()

    ev = lambda: torch.cuda.Event(enable_timing=True)
    acc = {"insert": 0.0, "query_present": 0.0, "query_absent": 0.0, "clear": 0.0}
    launches = {"insert": 0, "query_present": 0, "query_absent": 0}

    def step(timed):
        e = [ev() for _ in range(2 * len(batches) + 6)]
        k = 0
        e[k].record(stream); table.clear(stream); e[k + 1].record(stream); k += 2
        for (s, t) in batches:
            e[k].record(stream)
            table.bulk_insert(keys[s:t], None, None, fails, stream)
            e[k + 1].record(stream)
            k += 2
        e[k].record(stream); table.bulk_query(present, vals_p, found_p, stream); e[k + 1].record(stream)
        e[k + 2].record(stream); table.bulk_query(absent, vals_a, found_a, stream); e[k + 3].record(stream)
        return e

    def collect(e):
        torch.cuda.synchronize()
        acc["clear"] += e[0].elapsed_time(e[1])
        k = 2
        for _ in batches:
            acc["insert"] += e[k].elapsed_time(e[k + 1])
            launches["insert"] += 1
            k += 2
        acc["query_present"] += e[k].elapsed_time(e[k + 1])
        acc["query_absent"] += e[k + 2].elapsed_time(e[k + 3])
        launches["query_present"] += 1
        launches["query_absent"] += 1

    for _ in range(args.warmup):
        step(False)
    torch.cuda.synchronize()
    fails.zero_()
    for key in acc:
        acc[key] = 0.0
    events = []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        events.append(step(True))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for e in events:
        collect(e)
    ms_per_step = (t1 - t0) * 1e3 / args.steps

    # correctness of the last timed step
    false_neg = int((found_p == 0).sum().item())
    fp_rate = float(found_a.float().mean().item())
    insert_fails = int(fails.item())

    # accounting pass (untimed): algorithmic bucket examinations per phase
    stats = torch.zeros(3, dtype=torch.int64, device=dev)
    table.set_stats(stats)
    table.clear(stream)
    for (s, t) in batches:
        table.bulk_insert(keys[s:t], None, None, None, stream)
    torch.cuda.synchronize()
    st_ins = [int(x) for x in stats.tolist()]
    stats.zero_()
    table.bulk_query(present, vals_p, found_p, stream)
    torch.cuda.synchronize()
    st_qp = [int(x) for x in stats.tolist()]
    stats.zero_()
    table.bulk_query(absent, vals_a, found_a, stream)
    torch.cuda.synchronize()
    st_qa = [int(x) for x in stats.tolist()]
    table.set_stats(None)

    b0, b1, wsec = 64, 32, 64  # bytes per primary bucket, backing bucket, dirtied sector
    bytes_ins = st_ins[0] * b0 + st_ins[1] * b1 + st_ins[2] * wsec
    bytes_qp = st_qp[0] * b0 + st_qp[1] * b1
    bytes_qa = st_qa[0] * b0 + st_qa[1] * b1

    def kinfo(name, nops, nbytes):
        ms = acc[name] / args.steps
        gbs = nbytes / (ms / 1e3) / 1e9
        return {"ms": round(ms, 4), "mops": round(nops / (ms / 1e3) / 1e6, 1),
                "algorithmic_bytes": nbytes, "bytes_per_op": round(nbytes / nops, 2),
                "achieved_gbs": round(gbs, 1), "frac_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}

    kernels = {
        "insert": kinfo("insert", n_ins, bytes_ins),
        "query_present": kinfo("query_present", nq, bytes_qp),
        "query_absent": kinfo("query_absent", nq, bytes_qa),
    }

    # measured random 64 B-sector roofline over a buffer of the table's
    # footprint: best over the microbenchmark's tile shapes, sectors in flight
    # per tile (1-8) and waves per SIMD (1-6 grid-striding, or a grid covering
    # the work as the table kernels launch).  read = 64 B per op; read_write
    # and read_cas = 64 B read + the 64 B sector the store or CAS dirties (the
    # same accounting as the insert kernel).
    rand = None
    if not args.no_roofline:
        rand = random_roofline(pg, torch, dev, stream, nslots * 4)
        kernels["insert"]["frac_random_roofline"] = round(kernels["insert"]["achieved_gbs"] / rand["read_cas"], 4)
        for q in ("query_present", "query_absent"):
            kernels[q]["frac_random_roofline"] = round(kernels[q]["achieved_gbs"] / rand["read"], 4)

    if args.calibrate_copy:
        # a streaming kernel of known traffic (1 GiB read, 1 GiB written), so
        # PMC byte counters can be scaled; an elementwise kernel rather than
        # copy_(), which becomes a DMA transfer with no kernel to count ...
        a = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        b = torch.empty_like(a)
        torch.neg(a, out=b)
        torch.cuda.synchronize()
        del a, b
        # ... and random 64 B reads of known count (read-only, one sector in
        # flight per lane), the request size of the table kernels
        buf = torch.zeros(1 << 27, dtype=torch.int64, device=dev)
        sink = torch.zeros(1, dtype=torch.int64, device=dev)
        pg.random_sector_bench(buf, CAL_RANDOM_READS, 0, 5, sink, stream)
        torch.cuda.synchronize()
        del buf

    dom = max(("insert", "query_present", "query_absent"), key=lambda k: acc[k])
    traffic_info = load_traffic(os.path.join(ROOT, "profiles", "tcf_2p28_traffic.json"))
    traffic = None
    if traffic_info and traffic_info.get("kernel") == dom:
        traffic = traffic_info.get("hbm_bytes_per_launch")
    roofline = {
        "bound": "hbm",
        "achieved": kernels[dom]["achieved_gbs"],
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": kernels[dom]["frac_hbm_peak"],
        "traffic": traffic,
        "kernel": dom,
        "algorithmic_bytes_per_launch": kernels[dom]["algorithmic_bytes"] // max(1, launches[dom] // args.steps),
        "avg_launch_ms": round(acc[dom] / max(1, launches[dom]), 4),
        "measured_random_64B_gbs": rand,
    }
    tcc_dom = ((traffic_info or {}).get("tcc") or {}).get("insert" if dom == "insert" else "query")
    if tcc_dom:
        roofline["tcc_hit_rate"] = tcc_dom["tcc_hit_rate"]
        roofline["atomics_per_s"] = tcc_dom["atomics_per_s"]

    ops_per_step = n_ins + 2 * nq
    value = ops_per_step / (ms_per_step / 1e3) / 1e6
    out = {
        "metric": "TCF insert/query Mops/s at 90% load",
        "value": round(value, 1),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic (splitmix64 uint64 keys generated on the GPU)",
        "config": {
            "workload": f"TCF 2^{args.log2_slots} slots (16-bit fp, 16-slot buckets, split-fmix64 two choices, "
                        f"tile {args.tile}) + 1% linear backing; insert {n_ins} keys (90%) in batches of "
                        f"{batch}, query {nq} present + {nq} absent",
            "slots": nslots, "backing_slots": backing, "inserts": n_ins, "queries": 2 * nq,
        },
        "insert_mops": kernels["insert"]["mops"],
        "query_mops": round(2 * nq / ((acc["query_present"] + acc["query_absent"]) / args.steps / 1e3) / 1e6, 1),
        "false_negatives": false_neg,
        "fp_rate": fp_rate,
        "insert_failures": insert_fails,
        "kernels": kernels,
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_serve:
        del keys, absent, present, table
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        out["serving"] = serving_point(lambda m: print(m, file=sys.stderr, flush=True))
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
