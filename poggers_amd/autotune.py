"""Launch-shape autotuner for batched filters.

The reference tunes tile width ("Partition_Size") and bucket size by hand:
its tests recompile and time every pair (tests/src/power_2_tests.cu:63-100,
354-446; test_cg_variations.cu:878-1027; bucket_speed_tests.cu:1363-1390).
Here the shapes are run-time choices of one library, so the sweep is a
function: for each (tile, bucket) the batch is built, the captured
clear + insert + query graph is replayed for each work-chunk size, the
fastest shapes are then tried with 2 and 4 keys per tile in flight, and the
fastest combination wins.  Bucket size changes the filter's layout (and its
false-positive rate, about proportional to B), so callers that need a fixed
format pass a single bucket size.
"""
from __future__ import annotations

import torch

from . import Batch, tcf_desc


def _time_replays(batch: Batch, stream, reps: int) -> float:
    batch.replay(stream)
    stream.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        batch.replay(stream)
    e1.record(stream)
    stream.synchronize()
    return e0.elapsed_time(e1) / reps


def tune_batch(sizes, ins_keys, ins_off, q_keys, q_off, tiles=(1, 2, 4, 8, 16, 32, 64), buckets=(8, 16, 32),
               chunks=(256, 1024, 4096), items=(2, 4), refine=3, reps=3, seed=42, absent=None, fp_slack=0.05,
               log=None):
    """Sweep tile width x bucket size x chunk on the given workload (device
    tensors) with one key per tile in flight, then try each `items` count
    (keys per tile in flight) on the `refine` fastest shapes.

    Concurrency changes where keys land, not whether they are found: when
    many keys of one small table are in flight at once (small chunks, more
    keys per tile, narrow tiles over 8-slot buckets), stale bucket fills send
    more keys past their first choice and into the 1 % backing level, which
    can overflow (failed inserts) and lengthens absent-key walks (a higher
    false-positive rate).  So a shape qualifies only if its replays failed
    no insert and, when `absent` (a bool tensor over the queries) is given,
    its false-positive rate is within `fp_slack` of the lowest measured for
    its bucket size.  The fastest qualifying shape wins (when none qualifies,
    the fastest lossless one, then the fastest).

    Returns {"best": {...}, "results": [...]}: per shape the time in ms of
    one clear + insert + query replay, the inserts each replay failed and
    the false-positive rate (None without `absent`)."""
    stream = torch.cuda.Stream()
    found = torch.empty(q_keys.numel(), dtype=torch.uint8, device=q_keys.device)
    fails = torch.zeros(1, dtype=torch.int64, device=q_keys.device)
    results = []

    def run(batch, P, B, chunk, it):
        batch.set_launch(chunk)
        batch.set_items(it)
        fails.zero_()
        torch.cuda.synchronize()
        batch.capture(ins_keys, ins_off, q_keys, q_off, None, found, fails, True, stream)
        ms = _time_replays(batch, stream, reps)
        dropped = int(fails.item()) // (reps + 1)
        fp = float(found[absent].float().mean().item()) if absent is not None else None
        results.append({"tile": P, "bucket": B, "chunk": chunk, "items": it, "ms": round(ms, 4),
                        "insert_failures": dropped, "fp_rate": fp})
        if log:
            extra = f", fp {fp:.3e}" if fp is not None else ""
            log(f"tune tile={P} bucket={B} chunk={chunk} items={it}: {ms:.3f} ms, {dropped} failed inserts{extra}")

    def qualified(rs):
        ok = [r for r in rs if r["insert_failures"] == 0]
        if absent is not None and ok:
            floor = {}
            for r in ok:
                floor[r["bucket"]] = min(floor.get(r["bucket"], 1.0), r["fp_rate"])
            ok = [r for r in ok if r["fp_rate"] <= (1.0 + fp_slack) * floor[r["bucket"]]]
        return ok or [r for r in rs if r["insert_failures"] == 0] or rs

    for B in buckets:
        for P in tiles:
            batch = Batch(tcf_desc(1 << 12, 1 << 6, bucket_size=B, tile=P), sizes, seed=seed)
            for chunk in chunks:
                run(batch, P, B, chunk, 1)
            batch.free()
            del batch
    for r in sorted(qualified(list(results)), key=lambda r: r["ms"])[:refine]:
        batch = Batch(tcf_desc(1 << 12, 1 << 6, bucket_size=r["bucket"], tile=r["tile"]), sizes, seed=seed)
        for it in items:
            run(batch, r["tile"], r["bucket"], r["chunk"], it)
        batch.free()
        del batch
    best = min(qualified(results), key=lambda r: r["ms"])
    return {"best": best, "results": results}
