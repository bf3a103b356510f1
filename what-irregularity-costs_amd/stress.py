"""Allocator stress: insert-only throughput at high table load factors.

Workload (the allocate-only stress configuration): a universe of 2^18 distinct
block coordinates drawn uniformly in [-2^10, 2^10)^3; a stream of 2^22
coordinates made of every universe key once plus 2^22 - 2^18 draws from a
Zipf(s = 1.1) law over the universe (ranks mapped to keys by a seeded
permutation), shuffled. Drawing every key once makes the realised distinct
count exactly 2^18, so the tables below run at their nominal load factors:
ceil(2^18 / LF) slots for LF 0.5 / 0.7 / 0.9, which are not powers of two,
hence the mixed hash with multiply-shift range reduction; the reference
multiply-xor hash is also measured on the 2^19-slot power-of-two table.

The key set oracle is the closed form: the universe itself.

Run:  python what-irregularity-costs_amd/stress.py [--log2-keys 18] [--log2-stream 22]
Prints one JSON object: per table, insert time, inserts/s, compare-exchange
attempts and attempts/s, probe-depth histogram summary and the wave/lane step
ratio, plus the device's uncontended compare-exchange ceiling.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)


def pack_keys(c: np.ndarray) -> np.ndarray:
    c = c.astype(np.int64)
    b = np.int64(1 << 20)
    return ((c[:, 0] + b) << 42) | ((c[:, 1] + b) << 21) | (c[:, 2] + b)


def make_stream(seed: int = 7, log2_keys: int = 18, log2_stream: int = 22, s: float = 1.1):
    """Returns (universe int32 (K,3), stream int32 (N,3))."""
    rng = np.random.default_rng(seed)
    K = 1 << log2_keys
    N = 1 << log2_stream
    uni = np.empty((0, 3), np.int32)
    while len(uni) < K:
        more = rng.integers(-(1 << 10), 1 << 10, size=(2 * K, 3), dtype=np.int32)
        allc = np.concatenate([uni, more])
        _, idx = np.unique(pack_keys(allc), return_index=True)
        uni = allc[np.sort(idx)][:K]
    ranks = np.arange(1, K + 1, dtype=np.float64)
    p = ranks ** -s
    p /= p.sum()
    draws = rng.choice(K, size=N - K, p=p)
    perm = rng.permutation(K)  # rank -> key
    stream = np.concatenate([uni, uni[perm[draws]]])
    stream = stream[rng.permutation(N)]
    return uni, np.ascontiguousarray(stream)


TABLES = [  # (label, load factor, hash kind)
    ("lf0.5-mix", 0.5, 1),
    ("lf0.7-mix", 0.7, 1),
    ("lf0.9-mix", 0.9, 1),
    ("lf0.5-reference-hash", 0.5, 0),
]


def capacity(n_keys: int, lf: float, hash_kind: int) -> int:
    c = math.ceil(n_keys / lf)
    if hash_kind == 0:
        c = 1 << (c - 1).bit_length()
    return c


def run(log2_keys=18, log2_stream=22, seed=7, reps=3):
    import torch
    import osn_tsdf_hip as H
    uni, stream = make_stream(seed, log2_keys, log2_stream)
    K = len(uni)
    want = np.sort(pack_keys(uni))
    d = torch.from_numpy(stream.ravel()).cuda()
    out = {"keys": K, "stream": len(stream), "tables": {}}
    for label, lf, hk in TABLES:
        cap = capacity(K, lf, hk)
        rec = {"slots": cap, "load_factor": round(K / cap, 4), "hash": "mix" if hk else "reference"}
        times = []
        for r in range(reps + 1):
            v = H.HipVolume(0.01, K, -1.0, 0.0, storage=H.STORE_F16, table_slots=cap,
                            hash_kind=hk)
            v.set_timing(True)
            if v.insert_blocks(d):
                raise RuntimeError("insert failed")
            ms, calls = v.read_timing()
            if r > 0:
                times.append((ms[0], ms[1]))
            if r == reps:
                got = np.sort(v.keys())
                rec["key_set_exact"] = bool(np.array_equal(got, want))
                rec["block_count"] = v.block_count
                rec["drop_count"] = v.drop_count
            v.close()
        ins = float(np.mean([t[0] for t in times]))
        asg = float(np.mean([t[1] for t in times]))
        rec["insert_ms"] = round(ins, 4)
        rec["assign_ms"] = round(asg, 4)
        rec["inserts_per_s"] = round(len(stream) / (ins * 1e-3), 0)
        # compare-exchange attempts (counting mode)
        v = H.HipVolume(0.01, K, -1.0, 0.0, storage=H.STORE_F16, table_slots=cap, hash_kind=hk)
        v.set_timing(True)
        H.insert_blocks_mode(v, d, 1)
        ms_c, _ = v.read_timing()
        cas = H.cas_attempts(v)
        v.close()
        rec["cas_attempts"] = int(cas)
        rec["cas_per_s"] = round(cas / (ms_c[0] * 1e-3), 0)
        # probe depth
        v = H.HipVolume(0.01, K, -1.0, 0.0, storage=H.STORE_F16, table_slots=cap, hash_kind=hk)
        H.insert_blocks_mode(v, d, 2)
        hist, lane, wave, lookups = v.probe_stats()
        v.close()
        depth = np.arange(1, 65)
        cum = np.cumsum(hist)
        rec["probe"] = {
            "lookups": int(lookups), "lane_steps": int(lane), "wave_steps": int(wave),
            "mean_depth": round(float(lane) / max(int(lookups), 1), 3),
            "wave_over_lane": round(float(wave) / max(int(lane), 1), 3),
            "p50_depth": int(depth[np.searchsorted(cum, 0.5 * cum[-1])]),
            "p99_depth": int(depth[np.searchsorted(cum, 0.99 * cum[-1])]),
            "depth_64_or_more": int(hist[-1]),
            "hist": [int(x) for x in hist],
        }
        out["tables"][label] = rec
    ceil_ms = H.cas_ceiling(22, reps=5)
    out["cas_ceiling"] = {"ops": 1 << 22, "ms": round(ceil_ms, 4),
                          "cas_per_s": round((1 << 22) / (ceil_ms * 1e-3), 0)}
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--log2-keys", type=int, default=18)
    ap.add_argument("--log2-stream", type=int, default=22)
    ap.add_argument("--seed", type=int, default=7)
    a = ap.parse_args()
    print(json.dumps(run(a.log2_keys, a.log2_stream, a.seed)))
