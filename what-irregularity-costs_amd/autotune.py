"""Autotuner for the batched integration launch shape.

Searches, on the caller's real workload (volumes, frames, step count), the
knobs that change speed but not results:
  * allocate: workgroups (each loops over 16x16 pixel tiles), LDS key-set
    entries per workgroup, and the probe scheme (reference multiply-xor hash
    over a power-of-two table, or a 64-bit mix with multiply-shift range);
  * update: workgroup size (waves per workgroup, one block per wave) and
    the number of workgroups.
Coordinate descent: the allocate knobs are tried with the default update,
then the update knobs with the best allocate, each update shape in two
interleaved rounds scored by their mean; each trial integrates the same
steps into fresh volumes and is timed per stage with HIP events. Every trial
and the choice are written to a JSON log.
"""
from __future__ import annotations

import json
import time

ALLOC_GRIDS = (0, 1024, 4096)
LDS_SETS = (0, 256, 1024)
HASHES = (0, 1)
UPDATE_THREADS = (128, 256, 512)
# 2048 workgroups leave too few waves for the tail of larger frames: on the
# first frames they can look best, and later lose ~10 %
UPDATE_GRIDS = (4096, 8192, 16384)
UPDATE_ROUNDS = 2  # update shapes differ by a few percent: each is timed twice

DEFAULT = dict(alloc_grid=0, lds_set=1024, hash_kind=0, update_threads=256, update_grid=4096)


def _trial(H, make_volumes, depths, frames, warmup, steps, knobs):
    vols = make_volumes(knobs["hash_kind"])
    b = H.Batch(vols)
    if b.tune(knobs["alloc_grid"], knobs["update_grid"], knobs["update_threads"],
              knobs["lds_set"]):
        raise RuntimeError(f"bad knobs {knobs}")
    if b.set_schedule(depths, frames):
        raise RuntimeError("set_schedule failed")
    b.run(warmup)
    b.synchronize()
    b.read_timing()
    b.set_timing(True)
    b.run(steps)
    ms, calls = b.read_timing()
    b.close()
    for v in vols:
        v.close()
    per = [float(ms[i] / max(calls[i], 1)) for i in range(3)]
    return {"allocate_ms": per[0], "assign_ms": per[1], "update_ms": per[2]}


# one scene at a time has a few thousand blocks per frame: smaller grids
PLAIN_UPDATE_GRIDS = (512, 1024, 2048, 4096)


def tune(H, make_volumes, depths, frames, warmup=10, steps=20, log_path=None,
         update_grids=UPDATE_GRIDS, alloc_grids=ALLOC_GRIDS):
    """make_volumes(hash_kind) -> fresh volumes for one trial; depths/frames as
    for Batch.set_schedule, at least warmup + steps long. Returns the chosen
    knobs (a dict usable with Batch.tune and the volume config)."""
    t0 = time.time()
    trials = []
    best = dict(DEFAULT)

    def run(knobs):
        r = _trial(H, make_volumes, depths, frames, warmup, steps, knobs)
        trials.append({"knobs": dict(knobs), **r})
        return r

    # allocate knobs (the update does not depend on them)
    scores = []
    for g in alloc_grids:
        for s in LDS_SETS:
            for h in HASHES:
                k = dict(best, alloc_grid=g, lds_set=s, hash_kind=h)
                scores.append((run(k)["allocate_ms"], k))
    best = min(scores, key=lambda x: x[0])[1]
    # update knobs
    shapes = [dict(best, update_threads=t, update_grid=g)
              for t in UPDATE_THREADS for g in update_grids]
    sums = [0.0] * len(shapes)
    for _ in range(UPDATE_ROUNDS):
        for i, k in enumerate(shapes):
            sums[i] += run(k)["update_ms"]
    best = shapes[min(range(len(shapes)), key=lambda i: sums[i])]
    log = {"chosen": best, "default": DEFAULT, "trials": trials,
           "seconds": round(time.time() - t0, 2), "warmup": warmup, "steps": steps}
    if log_path:
        with open(log_path, "w") as f:
            json.dump(log, f, indent=1)
    return best, log
