"""Launch-shape autotuner for batched filters.

The reference tunes tile width ("Partition_Size") and bucket size by hand:
its tests recompile and time every pair (tests/src/power_2_tests.cu:63-100,
354-446; test_cg_variations.cu:878-1027; bucket_speed_tests.cu:1363-1390).
Here the shapes are run-time choices of one library, so the sweep is a
function: for each (tile, bucket) the batch is built, the captured
clear + insert + query graph is replayed for each work-chunk size, and the
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
               chunks=(256, 1024, 4096), reps=3, seed=42, log=None):
    """Sweep tile width x bucket size x chunk on the given workload (device
    tensors).  Returns {"best": {...}, "results": [...]} with times in ms for
    one clear + insert + query replay and the inserts each replay failed;
    the fastest shape with no failed insert wins."""
    stream = torch.cuda.Stream()
    found = torch.empty(q_keys.numel(), dtype=torch.uint8, device=q_keys.device)
    fails = torch.zeros(1, dtype=torch.int64, device=q_keys.device)
    results = []
    for B in buckets:
        for P in tiles:
            desc = tcf_desc(1 << 12, 1 << 6, bucket_size=B, tile=P)
            batch = Batch(desc, sizes, seed=seed)
            for chunk in chunks:
                batch.set_launch(chunk)
                fails.zero_()
                torch.cuda.synchronize()
                batch.capture(ins_keys, ins_off, q_keys, q_off, None, found, fails, True, stream)
                ms = _time_replays(batch, stream, reps)
                dropped = int(fails.item()) // (reps + 1)
                results.append({"tile": P, "bucket": B, "chunk": chunk, "ms": round(ms, 4),
                                "insert_failures": dropped})
                if log:
                    log(f"tune tile={P} bucket={B} chunk={chunk}: {ms:.3f} ms, {dropped} failed inserts")
            batch.free()
            del batch
    # a shape that drops keys is only chosen when every shape does
    complete = [r for r in results if r["insert_failures"] == 0] or results
    best = min(complete, key=lambda r: r["ms"])
    return {"best": best, "results": results}
