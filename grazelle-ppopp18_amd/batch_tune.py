"""Autotuner of the batched pull sweep.

Searches vectors per lane x workgroup size (one value each for the whole
batch, since the persistent sweep is one launch) and, for every size bucket,
the chunk length, by timing the batched pull sweep over that bucket's graphs
on the device (gz_batch_time_pull, best of `reps` runs after a warm-up).  The
pick minimises the sum over buckets of each bucket's best time.

    python batch_tune.py                  # tune the 1024-graph batch and print JSON
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import grazelle as gz  # noqa: E402

VPLS = (1, 2, 4)
WORKGROUPS = (64, 128, 256, 512, 1024)
CHUNK_STEPS = (1, 2, 4, 8, 16, 32)  # chunk = 64 * vpl * step


def tune(reps: int = 3, vpls=VPLS, workgroups=WORKGROUPS, steps=CHUNK_STEPS) -> dict:
    """Tune the attached batch; sets and returns the chosen configuration."""
    table = []
    best = None
    for vpl in vpls:
        for wg in workgroups:
            total = 0.0
            chunks = []
            for b in range(gz.BATCH_BUCKETS):
                bt, bc = None, None
                for s in steps:
                    c = 64 * vpl * s
                    t = gz.batch_time_pull(vpl, wg, (c,) * gz.BATCH_BUCKETS, b, reps)
                    table.append({"vpl": vpl, "workgroup": wg, "bucket": b, "chunk": c, "ms": t})
                    if t >= 0 and (bt is None or t < bt):
                        bt, bc = t, c
                total += bt
                chunks.append(bc)
            if best is None or total < best[0]:
                best = (total, vpl, wg, chunks)
    _, vpl, wg, chunks = best
    gz.batch_set_config(vectors_per_lane=vpl, workgroup_size=wg, chunk=tuple(chunks))
    return {"config": gz.batch_get_config(), "pull_ms_sum_of_buckets": best[0], "table": table}


def main() -> int:
    gz.init(0)
    gz.load_batch_device(gz.batch_specs())
    res = tune()
    print(json.dumps({k: v for k, v in res.items() if k != "table"}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
