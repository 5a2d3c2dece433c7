"""Autotuner for the projector's launch shapes and size-class threshold.

Times candidate settings on a particle sample with HIP events on torch's
current stream -- projectors run on it, batches fork from and join back
into it, so the events bracket the GPU work -- and keeps the fastest.  Single map: persistent workgroups
per CU of the two footprint kernels, the footprint size that sends a
particle to the workgroup kernel and, optionally, the group size of the
grouped medium-class kernel.  Batch: the per-map pipeline, or particles per
work item and workgroups per CU of the persistent kernel.
"""
from __future__ import annotations

import itertools

import torch


def _time(fn, reps: int = 2) -> float:
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def tune_projector(projector, parts, medium=(4, 8, 16), large=(2, 4, 8), large_pix=(16384, 65536, 262144),
                   group=(None,), reps: int = 2):
    """Returns (best settings dict, {settings: ms}) and applies the best.

    group: candidate group sizes K of the grouped medium-class kernel
    (Projector.set_grouping; None keeps the current one)."""
    results = {}

    def run():
        projector.reset()
        projector.render(parts)

    for k, m, l, lp in itertools.product(group, medium, large, large_pix):
        if k is not None:
            projector.set_grouping(k)
        projector.set_tuning(m, l, lp)
        results[(m, l, lp, k)] = _time(run, reps)
    best = min(results, key=results.get)
    if best[3] is not None:
        projector.set_grouping(best[3])
    projector.set_tuning(*best[:3])
    projector.reset()
    return {"medium_blocks_per_cu": best[0], "large_blocks_per_cu": best[1], "large_pix": best[2],
            "group": best[3]}, results


def tune_batch(batch, parts, chunks=(0, 16, 64, 256), blocks=(4, 8, 16), reps: int = 2):
    """chunk 0 is the per-map pipeline (blocks does not apply to it);
    chunk >= 1 the persistent launch with that many particles per item."""
    results = {}
    for c, b in itertools.product(chunks, blocks):
        if c == 0 and b != blocks[0]:
            continue
        batch.set_tuning(c, b)
        results[(c, b)] = _time(lambda: batch.render(parts, use_graph=True), reps)
    best = min(results, key=results.get)
    batch.set_tuning(*best)
    return {"chunk": best[0], "blocks_per_cu": best[1]}, results
