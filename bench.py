"""Benchmark: frames/s of depth-frame TSDF integration at 640x480, 1 cm voxels.

Workload ("precise" scene): a seeded 6x6x3 m room with 20 spheres and 10 boxes,
640x480 planar depth frames rendered on the GPU with N(0, 0.0012 z^2) noise and
1% dropouts, fx=fy=320, cx=320, cy=240; voxel 1 cm, truncation 4 cm, table
2^22 slots, pool 2^21 blocks, weight cap 32. A step is one frame of every
scene: allocate (FP32 back-projection + band walk + insert), index
assignment, projective update. All depth frames are resident in HBM before
the timed region. Each scene first integrates `--preroll` frames (default 40)
and `--warmup` frames untimed, so the timed frames meet a volume in steady
state; the autotuner picks its launch shapes on the pre-roll frames only.

Reported:
  value          frames/s of the batched FP16-storage pipeline (FP32
                 arithmetic, FP16 normalised TSDF + u16 weight), 16 scenes per
                 step, replayed from a captured graph;
  plain_*        one scene at a time in FP32 / FP16 storage: eagerly (three
                 launches per frame) and replayed from a one-scene graph tuned
                 on a single scene; the headline ratio uses the faster plain
                 FP32 figure; `attribution` splits the ratio into batching
                 and FP16 storage;
  roofline       the update kernel: algorithmic bytes (touched blocks x 512
                 voxels x (read + write) x bytes/voxel + depth image + block
                 list) / its average duration from HIP events that its own
                 launches record (start and end of the dispatch, on the batch
                 stream), against 8 TB/s;
  alloc_roofline the allocate kernel: distinct-key table lookups and counted
                 compare-and-swap attempts per launch, per second, against an
                 uncontended 64-bit CAS ceiling measured in the same run;
                 probe-depth percentiles and the wave/lane probe-step ratio;
  chain_*        the same frames with a device-resident mesh extraction of
                 every scene after every frame (allocate -> assign -> update
                 -> extract), batched fp16 against plain fp32;
  batched_workload  24 independent scenes of uneven size (320x240, 640x480,
                 1280x720, tables 2^16..2^22), frame counts scaled to a
                 bounded run, tuned on frames before the timed ones; timed
                 again with every scene's mesh extracted in the graph after
                 every step;
  precision      max / RMS |dtsdf| against the FP64 oracle kernel after the
                 full 500-frame trajectory of one scene, in truncation units
                 and in voxels, and the meshes through the mesh gate;
  cpu_baseline   the reference CPU arm (compiled from the reference sources,
                 oracle/_ref) with all host threads, on a bounded sample of the
                 same frames, plus its extraction time and the small workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "what-irregularity-costs_amd"))

HBM_PEAK_GBS = 8000.0

# HBM bytes per update launch from FETCH_SIZE and WRITE_SIZE counter runs of
# `bench.py --profile-only` (rocprofv3 --pmc, one counter per run), each
# scaled by what the counter reads for a 256 MiB device copy (FETCH x 2.0,
# WRITE x 1.0). Counters cannot run inside this timed benchmark, so the
# launches they were measured on are named beside them.
UPDATE_TRAFFIC = {"hbm_bytes_per_launch": 335.7e6, "read": 239.2e6, "write": 96.5e6,
                  "algorithmic_bytes_per_launch": 435.0e6,
                  "launches": "k_proj_update_batch<fp16>, 16 scenes, the timed steps 45..64 "
                              "(after 40 pre-roll and 5 warmup frames), 101.1k touched blocks "
                              "per launch; voxels frozen at the weight cap are read, not written",
                  "summary": "profiles/k_proj_update_batch__precise16_640x480.md"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def render_frames(H, S, cfg, seed, n):
    import torch
    scene, poses = S.build(dict(cfg, frames=n), seed)
    W, Hh = cfg["width"], cfg["height"]
    fx, fy, cx, cy = S.intrinsics(W, Hh)
    sc = scene.to_c()
    depth = torch.empty((n, Hh * W), dtype=torch.float32, device="cuda")
    frames = []
    for k, (R, t) in enumerate(poses):
        fr = H.make_frame(W, Hh, fx, fy, cx, cy, R, t)
        if S.render_depth(sc, fr, cfg["noise_k"], cfg["dropout"], seed, k, depth[k]) != 0:
            raise RuntimeError("render failed")
        frames.append(fr)
    torch.cuda.synchronize()
    return frames, depth


KNOBS = {}        # batched launch shape chosen by the autotuner (empty = library defaults)
PLAIN_KNOBS = {}  # the one-scene launch shape


def make_volume(H, cfg, storage, hash_kind=None, pool=None, knobs=None):
    """Table cfg["table"] slots; pool table/2 blocks unless given (the tuner's
    short trials use a smaller pool over the same table)."""
    k = KNOBS if knobs is None else knobs
    hk = k.get("hash_kind", 0) if hash_kind is None else hash_kind
    return H.HipVolume(cfg["voxel"], pool or cfg["table"] // 2, cfg["trunc"], 32.0,
                       storage=storage, table_slots=cfg["table"],
                       hash_kind=hk, lds_set=k.get("lds_set", -1),
                       alloc_grid=k.get("alloc_grid", 0),
                       update_threads=k.get("update_threads", 0),
                       update_grid=k.get("update_grid", 0))


def apply_knobs(b, k):
    if k and b.tune(k["alloc_grid"], k["update_grid"], k["update_threads"], k["lds_set"]):
        raise RuntimeError("batch tune failed")


def run_plain(H, cfg, storage, scenes, untimed, steps, timing=False):
    """One scene at a time, one frame per call: three launches per frame.
    With timing, also the per-stage kernel times (events around each stage's
    launches) of the timed frames, ms per frame."""
    import torch
    total = 0.0
    ms = np.zeros(4)
    calls = np.zeros(4, np.int64)
    for frames, depth in scenes:
        v = make_volume(H, cfg, storage, knobs=PLAIN_KNOBS)
        for k in range(untimed):
            if v.integrate_depth(depth[k], frames[k]):
                raise RuntimeError("integrate failed")
        v.synchronize()
        torch.cuda.synchronize()
        if timing:
            v.read_timing()
            v.set_timing(True)
        t0 = time.perf_counter()
        for k in range(untimed, untimed + steps):
            if v.integrate_depth(depth[k], frames[k]):
                raise RuntimeError("integrate failed")
        v.synchronize()
        total += time.perf_counter() - t0
        if timing:
            m, c = v.read_timing()
            ms += m
            calls += c
        v.close()
    return total, ms / np.maximum(calls, 1)


def run_plain_graph(H, cfg, storage, scenes, untimed, steps):
    """One scene at a time, each frame's three kernels replayed from a
    one-scene graph with the one-scene launch shape."""
    import torch
    total = 0.0
    for frames, depth in scenes:
        v = make_volume(H, cfg, storage, knobs=PLAIN_KNOBS)
        b = H.Batch([v])
        apply_knobs(b, PLAIN_KNOBS)
        n = untimed + steps
        if b.set_schedule([[depth[k]] for k in range(n)], [[frames[k]] for k in range(n)]):
            raise RuntimeError("set_schedule failed")
        if b.run(untimed, use_graph=True):
            raise RuntimeError("batch run failed")
        b.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if b.run(steps, use_graph=True):
            raise RuntimeError("batch run failed")
        b.synchronize()
        total += time.perf_counter() - t0
        b.close()
        v.close()
    return total


def make_batch(H, cfg, storage, scenes, n=None):
    vols = [make_volume(H, cfg, storage) for _ in scenes]
    b = H.Batch(vols)
    apply_knobs(b, KNOBS)
    n = n or scenes[0][1].shape[0]
    depths = [[sc[1][k] for sc in scenes] for k in range(n)]
    frames = [[sc[0][k] for sc in scenes] for k in range(n)]
    if b.set_schedule(depths, frames):
        raise RuntimeError("set_schedule failed")
    return b, vols


def close_all(b, vols):
    b.close()
    for v in vols:
        v.close()


def run_batched(H, cfg, storage, scenes, untimed, steps, graph=True):
    """All scenes per step in one launch per stage, replayed from a hipGraph."""
    import torch
    b, vols = make_batch(H, cfg, storage, scenes)
    if b.run(untimed, use_graph=graph):
        raise RuntimeError("batch run failed")
    b.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if b.run(steps, use_graph=graph):
        raise RuntimeError("batch run failed")
    b.synchronize()
    dt = time.perf_counter() - t0
    blocks = sum(v.block_count for v in vols)
    close_all(b, vols)
    return dt, blocks


def run_batched_instrumented(H, cfg, storage, scenes, untimed, steps, bytes_per_voxel,
                             with_all=False):
    """Per-stage HIP-event times of back-to-back eager batched steps, and the
    per-step touched-block counts (the same frames give the same touched sets,
    so they are collected in a second, synchronised pass that is not timed);
    with_all: also the counts of every step, the untimed ones included."""
    b, vols = make_batch(H, cfg, storage, scenes)
    b.run(untimed)
    b.synchronize()
    b.read_timing()
    b.set_timing(True)
    b.run(steps)
    ms, calls = b.read_timing()
    close_all(b, vols)
    b, vols = make_batch(H, cfg, storage, scenes)
    touched = []
    for _ in range(untimed + steps):
        b.run(1)
        touched.append(sum(v.touched_count for v in vols))  # waits for the batch stream
    W, Hh = cfg["width"], cfg["height"]
    tb_all = np.array(touched, np.float64)
    tb = tb_all[untimed:]
    alg_bytes = tb * 512 * 2 * bytes_per_voxel + len(scenes) * W * Hh * 4 + tb * 12
    close_all(b, vols)
    if with_all:
        return ms, calls, tb, alg_bytes, tb_all
    return ms, calls, tb, alg_bytes


def percentile_from_hist(hist, q):
    """Probe depth (steps) at quantile q of a histogram whose bin i counts
    lookups of i + 1 steps (the last bin: that many or more)."""
    c = np.cumsum(hist.astype(np.float64))
    if c[-1] == 0:
        return 0
    return int(np.searchsorted(c, q * c[-1]) + 1)


def alloc_statistics(H, cfg, scenes, untimed, steps, alloc_ms):
    """The allocate kernel at the headline workload: distinct-key lookups and
    probe steps per launch (probe-depth pass), compare-and-swap attempts per
    launch (counting pass), both over the timed steps of the same frames,
    against an uncontended 64-bit CAS ceiling over 2^22 scattered slots
    measured now. alloc_ms: the normal kernel's average duration per launch."""
    b, vols = make_batch(H, cfg, H.STORE_F16, scenes)
    if b.set_alloc_mode(2):
        raise RuntimeError("alloc mode failed")
    b.run(untimed)
    b.synchronize()
    before = [v.probe_stats() for v in vols]
    b.run(steps)
    b.synchronize()
    after = [v.probe_stats() for v in vols]
    close_all(b, vols)
    hist = sum(a[0] - p[0] for a, p in zip(after, before))
    lane = sum(a[1] - p[1] for a, p in zip(after, before))
    wave = sum(a[2] - p[2] for a, p in zip(after, before))
    look = sum(a[3] - p[3] for a, p in zip(after, before))
    b, vols = make_batch(H, cfg, H.STORE_F16, scenes)
    if b.set_alloc_mode(1):
        raise RuntimeError("alloc mode failed")
    b.run(untimed)
    b.synchronize()
    c0 = sum(H.cas_attempts(v) for v in vols)
    b.run(steps)
    b.synchronize()
    cas = sum(H.cas_attempts(v) for v in vols) - c0
    close_all(b, vols)
    ceil_ms = H.cas_ceiling(22, 5)
    ceiling = (1 << 22) / (ceil_ms * 1e-3)
    t = alloc_ms * 1e-3
    pixels = len(scenes) * cfg["width"] * cfg["height"]
    lookups = look / steps
    return {"bound": "valu (band walk and key handling; the table atomics run far below "
                     "their ceiling)",
            "pixels_per_launch": pixels,
            "lookups_per_launch": round(lookups, 1),
            "cas_attempts_per_launch": round(cas / steps, 1),
            "avg_launch_ms": round(alloc_ms, 5),
            "lookups_per_s": round(lookups / t, 1),
            "cas_per_s": round(cas / steps / t, 1),
            "ceiling_cas_per_s": round(ceiling, 1),
            "ceiling": "2^22 uncontended 64-bit CAS on scattered slots, one lane each",
            "frac": round(lookups / t / ceiling, 4),
            "mean_probe_steps": round(lane / max(look, 1), 3),
            "probe_p50": percentile_from_hist(hist, 0.5),
            "probe_p99": percentile_from_hist(hist, 0.99),
            "probe_max_bin": int(np.nonzero(hist)[0].max() + 1) if hist.sum() else 0,
            "wave_over_lane_steps": round(wave / max(lane, 1), 3)}


def measured_ceilings(touched_blocks):
    """Two ceilings measured on this GPU: a plain device-to-device copy of
    1 GiB (bytes read + written per second) and the update's own memory
    pattern with no arithmetic: `touched_blocks` random 2 KiB blocks of a
    2 GiB pool read and written back, the next block's load issued before the
    current block's store (what-irregularity-costs_amd/rmw_ceiling, its
    "unordered" rmw-next line)."""
    import subprocess
    import torch
    src = torch.empty(1 << 28, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    src.fill_(1.0)
    for _ in range(2):
        dst.copy_(src)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    reps = 10
    for _ in range(reps):
        dst.copy_(src)
    e1.record()
    e1.synchronize()
    copy_gbs = 2 * src.numel() * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del src, dst
    torch.cuda.empty_cache()
    out = {"stream_copy_gbs": round(copy_gbs, 1),
           "stream_copy": "torch copy_ of 1 GiB on the device, 10 repetitions, read + write"}
    exe = os.path.join(ROOT, "what-irregularity-costs_amd", "rmw_ceiling")
    if not os.path.exists(exe):
        out["pattern_error"] = f"{exe} is missing: build it (what-irregularity-costs_amd/build.py)"
        log(out["pattern_error"])
        return out
    r = subprocess.run([exe, str(int(touched_blocks)), str(1 << 20), "16384"],
                       capture_output=True, text=True, timeout=120)
    line = next((l for l in r.stdout.splitlines() if l.startswith("unordered")), None)
    if r.returncode or line is None:
        out["pattern_error"] = f"rmw_ceiling failed ({r.returncode}): {r.stderr.strip()[-200:]}"
        log(out["pattern_error"])
        return out
    us = float(line.split("rmw-next")[1].split("us")[0])
    out["pattern_ceiling_us"] = us
    out["pattern"] = (f"{int(touched_blocks)} random 2 KiB blocks read and written back, next "
                      "block's load before the current store, no arithmetic (rmw_ceiling)")
    return out


CHAIN_VERTS = 1 << 25  # per-scene mesh buffer (vertices); noisy 1 cm fields make dense soups


def run_chain(H, cfg, storage, scenes, untimed, steps, batched):
    """The full per-frame chain allocate -> assign -> update -> extract (mesh of
    every scene into a device buffer after every frame): batched in one graph
    per step, or one scene at a time through the single-volume calls.
    Returns (seconds, vertex counts of the last meshes; -1 = overflow)."""
    import torch
    n = len(scenes)
    bufs = [torch.empty(CHAIN_VERTS * 6, dtype=torch.float32, device="cuda") for _ in range(n)]
    counts = torch.zeros(n, dtype=torch.int32, device="cuda")
    if batched:
        b, vols = make_batch(H, cfg, storage, scenes)
        if b.set_extract(bufs, [CHAIN_VERTS] * n, counts, 0.5, 0.0, every=1):
            raise RuntimeError("set_extract failed")
        if b.run(untimed, use_graph=True):
            raise RuntimeError("batch run failed")
        b.synchronize()
        t0 = time.perf_counter()
        if b.run(steps, use_graph=True):
            raise RuntimeError("batch run failed")
        b.synchronize()
        dt = time.perf_counter() - t0
        b.close()
    else:
        dt = 0.0
        vols = []
        for s, (frames, depth) in enumerate(scenes):
            v = make_volume(H, cfg, storage, knobs=PLAIN_KNOBS)
            vols.append(v)
            for k in range(untimed):
                if v.integrate_depth(depth[k], frames[k]):
                    raise RuntimeError("integrate failed")
            v.synchronize()
            t0 = time.perf_counter()
            for k in range(untimed, untimed + steps):
                if v.integrate_depth(depth[k], frames[k]) or v.extract_device(
                        bufs[s], None, CHAIN_VERTS, 0.5, 0.0, counts[s:s + 1]):
                    raise RuntimeError("integrate/extract failed")
            v.synchronize()
            dt += time.perf_counter() - t0
    out = counts.cpu().tolist()
    for v in vols:
        v.close()
    del bufs
    torch.cuda.empty_cache()
    return dt, out


def run_chain_instrumented(H, cfg, storage, scenes, untimed, steps):
    """The batched chain through the eager path with per-stage HIP-event times
    (allocate, assign, update, extract = count + scan + emit of every scene),
    ms per step over `steps` steps after `untimed` ones, and the vertices of
    the last step's meshes (the meshes grow a little per step: keep `steps`
    small for a per-step byte figure) and the blocks allocated."""
    import torch
    n = len(scenes)
    bufs = [torch.empty(CHAIN_VERTS * 6, dtype=torch.float32, device="cuda") for _ in range(n)]
    counts = torch.zeros(n, dtype=torch.int32, device="cuda")
    b, vols = make_batch(H, cfg, storage, scenes)
    if b.set_extract(bufs, [CHAIN_VERTS] * n, counts, 0.5, 0.0, every=1):
        raise RuntimeError("set_extract failed")
    if b.run(untimed):
        raise RuntimeError("batch run failed")
    b.synchronize()
    b.read_timing()
    b.set_timing(True)
    if b.run(steps):
        raise RuntimeError("batch run failed")
    b.synchronize()
    ms, calls = b.read_timing()
    out = counts.cpu().tolist()
    blocks = sum(v.block_count for v in vols)  # waits for the batch
    close_all(b, vols)
    del bufs
    torch.cuda.empty_cache()
    return ms / np.maximum(calls, 1), out, blocks


def run_batched_workload(H, S, seed=4, n_scenes=24, scale=0.15, tune_frames=12):
    """The batched workload of uneven scenes. Each scene renders tune_frames
    frames before its scheduled ones; the autotuner picks the launch shape on
    those, and the timed run integrates the scheduled frames from empty
    volumes (the first step, which also captures the graph, untimed). A
    scene idles (NULL depth) after its last frame."""
    import torch
    import autotune
    cfgs = S.batched_workload(seed, n_scenes, scale)
    n = len(cfgs)
    rend = []
    for c in cfgs:
        fr, d = render_frames(H, S, c, c["seed"], tune_frames + c["frames"])
        rend.append((fr, d))
    tdep = [[rend[s][1][k] for s in range(n)] for k in range(tune_frames)]
    tfr = [[rend[s][0][k] for s in range(n)] for k in range(tune_frames)]

    def vols_for(hk, small_pool):
        return [H.HipVolume(c["voxel"], min(c["table"] // 2, 1 << 17) if small_pool
                            else c["table"] // 2, c["trunc"], 32.0, storage=H.STORE_F16,
                            table_slots=c["table"], hash_kind=hk) for c in cfgs]

    best, tlog = autotune.tune(H, lambda hk: vols_for(hk, True), tdep, tfr, warmup=4,
                               steps=tune_frames - 4)
    steps = max(c["frames"] for c in cfgs)
    dep = [[rend[s][1][tune_frames + k] if k < cfgs[s]["frames"] else None for s in range(n)]
           for k in range(steps)]
    frs = [[rend[s][0][tune_frames + k] if k < cfgs[s]["frames"] else None for s in range(n)]
           for k in range(steps)]
    timed_frames = sum(max(0, c["frames"] - 1) for c in cfgs)
    pixels = sum(max(0, c["frames"] - 1) * c["width"] * c["height"] for c in cfgs)

    def batch():
        vols = vols_for(best["hash_kind"], False)
        b = H.Batch(vols)
        apply_knobs(b, best)
        if b.set_schedule(dep, frs):
            raise RuntimeError("set_schedule failed")
        return b, vols

    b, vols = batch()
    b.run(1, use_graph=True)
    b.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if b.run(steps - 1, use_graph=True):
        raise RuntimeError("batch run failed")
    b.synchronize()
    dt = time.perf_counter() - t0
    blocks = [v.block_count for v in vols]
    drops = [v.drop_count for v in vols]
    close_all(b, vols)
    b, vols = batch()
    b.run(1)
    b.synchronize()
    b.read_timing()
    b.set_timing(True)
    b.run(steps - 1)
    ms, calls = b.read_timing()
    close_all(b, vols)
    # The same schedule with every scene's mesh extracted inside the graph
    # after every step (allocate -> assign -> update -> count/scan/emit for all
    # scenes), each scene into its own device buffer.
    bufs = [torch.empty(CHAIN_VERTS * 6, dtype=torch.float32, device="cuda") for _ in range(n)]
    counts = torch.zeros(n, dtype=torch.int32, device="cuda")
    b, vols = batch()
    if b.set_extract(bufs, [CHAIN_VERTS] * n, counts, 0.5, 0.0, every=1):
        raise RuntimeError("set_extract failed")
    b.run(1, use_graph=True)
    b.synchronize()
    t0 = time.perf_counter()
    if b.run(steps - 1, use_graph=True):
        raise RuntimeError("batch run failed")
    b.synchronize()
    dt_x = time.perf_counter() - t0
    mesh_verts = counts.cpu().tolist()
    close_all(b, vols)
    if min(mesh_verts) < 0:
        raise RuntimeError(f"a mesh overflowed its buffer: {mesh_verts}")
    del bufs
    del rend
    torch.cuda.empty_cache()
    sizes = {}
    for c in cfgs:
        key = f"{c['width']}x{c['height']}"
        sizes[key] = sizes.get(key, 0) + 1
    return {"workload": f"{n} scenes, sizes {sizes}, tables 2^"
                        f"{min(c['table'].bit_length() - 1 for c in cfgs)}..2^"
                        f"{max(c['table'].bit_length() - 1 for c in cfgs)}, frames U{{50..400}} x "
                        f"{scale} ({min(c['frames'] for c in cfgs)}..{steps}), 1 cm, fp16, graph",
            "steps_timed": steps - 1, "frames_timed": timed_frames,
            "frames_per_s": round(timed_frames / dt, 1),
            "megapixels_per_s": round(pixels / dt / 1e6, 1),
            "ms_per_step": round(dt * 1e3 / (steps - 1), 4),
            "stage_ms_per_step": {s: round(float(ms[i] / max(calls[i], 1)), 5)
                                  for i, s in enumerate(H.STAGES[:3])},
            "blocks_allocated": int(sum(blocks)), "scenes_with_drops": int(sum(d > 0 for d in drops)),
            "with_extraction": {
                "stages": "allocate + assign + update + extract (device mesh of every scene "
                          "after every step), one graph per step",
                "frames_per_s": round(timed_frames / dt_x, 1),
                "ms_per_step": round(dt_x * 1e3 / (steps - 1), 4),
                "last_mesh_vertices": int(sum(mesh_verts)),
                "largest_mesh_vertices": int(max(mesh_verts))},
            "autotune": {"chosen": best, "trials": len(tlog["trials"]),
                         "seconds": tlog["seconds"],
                         "tuned_on": f"{tune_frames} frames per scene before the timed ones"}}


PRECISION_MESH_VERTS = 96 << 20


def precision(H, cfg, frames, depth, n):
    """FP16 and FP32 states against the FP64 oracle kernel after n frames:
    |dtsdf| over every observed voxel, and the extracted meshes (min weight
    0.5, iso 0) against the FP64 mesh through the mesh-comparison gate
    (Chamfer mean = symmetric mean surface distance, Hausdorff, p99)."""
    import mesh_metrics as MM
    vols = {}
    for name, st in (("f64", H.STORE_F64), ("f32", H.STORE_F32), ("f16", H.STORE_F16)):
        v = make_volume(H, cfg, st, knobs={})
        for k in range(n):
            v.integrate_depth(depth[k], frames[k])
        v.synchronize()
        vols[name] = v
    meshes = {name: v.extract(0.5, 0.0, cap=PRECISION_MESH_VERTS) for name, v in vols.items()}
    keys = vols["f64"].keys()
    t64, w64 = vols["f64"].read_voxels_f64()
    vols["f64"].close()
    out = {"frames": n, "blocks": int(len(keys))}
    for name in ("f32", "f16"):
        k, (t, w) = vols[name].keys(), vols[name].read_voxels()
        vols[name].close()
        assert np.array_equal(k, keys), "key sets differ between storages"
        obs = w64 > 0
        assert np.array_equal(w.astype(np.float64)[obs], w64[obs])
        err = np.abs(t.astype(np.float64) - t64)[obs]
        vox = cfg["trunc"] / cfg["voxel"]
        out[name] = {"max_trunc": float(err.max()), "rms_trunc": float(np.sqrt((err ** 2).mean())),
                     "max_voxels": float(err.max() * vox),
                     "rms_voxels": float(np.sqrt((err ** 2).mean()) * vox),
                     "observed_voxels": int(obs.sum()),
                     "saturated_voxels": int((w64 == 32.0).sum())}
        if meshes[name] is not None and meshes["f64"] is not None:
            ag = MM.compare(meshes[name], meshes["f64"], 200000, cfg["voxel"], 0xC0FFEE)
            out[name]["mesh_vs_fp64"] = {
                "chamfer_mean_voxels": ag.mean_surface / cfg["voxel"],
                "hausdorff_p99_voxels": ag.hausdorff_p99 / cfg["voxel"],
                "hausdorff_voxels": ag.hausdorff / cfg["voxel"],
                "vertices": int(len(meshes[name])), "vertices_fp64": int(len(meshes["f64"])),
                "samples_per_side": 200000}
    return out


def _cpu_frames(H, B, vol, frames, depth, budget_s, max_frames):
    """Allocate + update per frame on the reference CPU arm, nearest-rank
    percentiles, then one extraction of the volume."""
    import torch
    W, Hh = frames[0].width, frames[0].height
    buf = torch.empty(W * Hh * 3, dtype=torch.float32, device="cuda")
    spent = 0.0
    n = 0
    times = []
    t_start = time.perf_counter()
    for k in range(min(max_frames, len(frames))):
        H.backproject(depth[k], frames[k], buf)
        p = buf.cpu().numpy().reshape(-1, 3)
        pts = np.ascontiguousarray(p[~np.isnan(p[:, 0])].ravel())
        cam = tuple(frames[k].t_wc[:])
        t0 = time.perf_counter()
        vol.allocate(pts, cam)
        vol.update(pts, cam)
        dt = time.perf_counter() - t0
        if k > 0:  # the first frame is warmup
            times.append(dt)
            spent += dt
            n += 1
        if spent > budget_s or time.perf_counter() - t_start > 3 * budget_s:
            break
    t0 = time.perf_counter()
    mesh = vol.extract(0.5, 0.0, cap=64 << 20)
    ext = time.perf_counter() - t0
    t = np.sort(np.array(times))
    pct = lambda q: float(t[min(len(t) - 1, int(np.ceil(q * len(t))) - 1)])  # nearest rank
    return n, spent, pct, ext, (len(mesh) if mesh is not None else -1)


def cpu_baseline(H, S, cfg, frames, depth, budget_s=12.0, max_frames=400):
    """The reference CPU arm (all threads) on the same frames' points, then on
    the small workload's 20 frames. Reported beside the GPU, not a target."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import bindings as B
    if not B.ref_available():
        msg = ("oracle/_ref/libosn_ref.so is missing (make -C oracle ref, where the reference "
               "checkout exists): no CPU baseline was timed")
        log("ERROR: " + msg)
        return {"error": msg}
    vol = B.RefVolume(cfg["voxel"], cfg["table"] // 2, cfg["trunc"], 32.0)
    threads = vol.threads
    n, spent, pct, ext, verts = _cpu_frames(H, B, vol, frames, depth, budget_s, max_frames)
    del vol
    small = dict(S.SMALL)
    sf, sd = render_frames(H, S, small, 5, small["frames"])
    svol = B.RefVolume(small["voxel"], small["table"] // 2, small["trunc"], 32.0)
    sn, sspent, spct, sext, sverts = _cpu_frames(H, B, svol, sf, sd, budget_s, small["frames"])
    del svol
    return {"value": n / spent, "unit": "frames/s", "cores": int(threads), "kind": "reference",
            "sample": f"frames 1..{n} of the same trajectory (frame 0 warmup), allocate+update "
                      f"per frame on back-projected points, p50 {pct(0.5) * 1e3:.1f} ms, "
                      f"p95 {pct(0.95) * 1e3:.1f} ms, p99 {pct(0.99) * 1e3:.1f} ms",
            "extract_ms": round(ext * 1e3, 1), "extract_vertices": verts,
            "frames_per_s_with_one_extract_per_frame": round(1.0 / (spent / n + ext), 2),
            "small_workload": {
                "workload": "4x4x3 m room, 8 spheres, 20 frames 640x480, voxel 2 cm, table 2^18",
                "frames_per_s": round(sn / sspent, 1),
                "p50_ms": round(spct(0.5) * 1e3, 2), "p95_ms": round(spct(0.95) * 1e3, 2),
                "extract_ms": round(sext * 1e3, 1), "extract_vertices": sverts,
                "frames": sn}}


def tune_batched(H, cfg, scenes, preroll):
    """Launch shape for the batched step, on pre-roll frames only."""
    import autotune
    n_t = min(preroll, 30)
    depths = [[sc[1][k] for sc in scenes] for k in range(n_t)]
    frs = [[sc[0][k] for sc in scenes] for k in range(n_t)]
    w = min(10, n_t // 3)
    best, tlog = autotune.tune(
        H, lambda hk: [make_volume(H, cfg, H.STORE_F16, hash_kind=hk, pool=1 << 17, knobs={})
                       for _ in scenes],
        depths, frs, warmup=w, steps=n_t - w,
        log_path=os.path.join(ROOT, "gpurun_out", "autotune.json")
        if os.path.isdir(os.path.join(ROOT, "gpurun_out")) else None)
    return best, {"chosen": best, "trials": len(tlog["trials"]), "seconds": tlog["seconds"],
                  "tuned_on": f"pre-roll frames 0..{n_t - 1}"}


def tune_plain(H, cfg, scene, preroll):
    """Launch shape for one scene at a time (FP32 storage), on its pre-roll
    frames, over grids sized for a single frame."""
    import autotune
    n_t = min(preroll, 30)
    frames, depth = scene
    w = min(10, n_t // 3)
    best, tlog = autotune.tune(
        H, lambda hk: [make_volume(H, cfg, H.STORE_F32, hash_kind=hk, pool=1 << 17, knobs={})],
        [[depth[k]] for k in range(n_t)], [[frames[k]] for k in range(n_t)], warmup=w,
        steps=n_t - w, update_grids=autotune.PLAIN_UPDATE_GRIDS)
    return best, {"chosen": best, "trials": len(tlog["trials"]), "seconds": tlog["seconds"],
                  "tuned_on": f"one scene, pre-roll frames 0..{n_t - 1}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--preroll", type=int, default=40,
                    help="frames each scene integrates before the warmup (untimed); the "
                         "autotuner uses these frames only")
    ap.add_argument("--scenes", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--precision-frames", type=int, default=500)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-workload", action="store_true", help="skip the uneven-scene leg")
    ap.add_argument("--no-tune", action="store_true", help="library default launch shape")
    ap.add_argument("--knobs", default=None,
                    help="fixed launch shape as JSON (alloc_grid, lds_set, hash_kind, "
                         "update_threads, update_grid); skips the autotuner")
    ap.add_argument("--profile-only", action="store_true",
                    help="only the batched fp16 run and one 256 MiB device copy (the copy "
                         "calibrates HBM byte counters), for rocprofv3 passes")
    ap.add_argument("--profile-chain", action="store_true",
                    help="only the batched fp16 chain with extraction every frame (graph, "
                         "then the instrumented eager pass) and the calibration copy")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus above 1 is not supported"}))
        sys.exit(2)

    import torch
    import osn_tsdf_hip as H
    import scenes as S
    if not torch.cuda.is_available():
        sys.exit("bench.py needs a GPU")
    H.lib()
    cfg = dict(S.PRECISE)
    P = args.preroll
    untimed = P + args.warmup
    n = untimed + args.steps
    log(f"rendering {args.scenes} scenes x {n} frames {cfg['width']}x{cfg['height']} "
        f"({P} pre-roll, {args.warmup} warmup, {args.steps} timed)")
    scenes = [render_frames(H, S, cfg, args.seed + i, n) for i in range(args.scenes)]
    frames_timed = args.steps * args.scenes

    tune_log = plain_tune_log = None
    if args.knobs:
        KNOBS.update(json.loads(args.knobs))
        PLAIN_KNOBS.update(KNOBS)
        tune_log = {"chosen": dict(KNOBS), "trials": 0, "seconds": 0.0}
    elif not args.no_tune and P >= 6:
        log("autotuning the batched launch shape on the pre-roll frames")
        best, tune_log = tune_batched(H, cfg, scenes, P)
        KNOBS.update(best)
        log(f"chosen {best}")
        if not (args.profile_only or args.profile_chain):
            log("autotuning the one-scene launch shape")
            pbest, plain_tune_log = tune_plain(H, cfg, scenes[0], P)
            PLAIN_KNOBS.update(pbest)
            log(f"chosen {pbest}")

    if args.profile_only:
        src = torch.empty(1 << 26, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        dst.copy_(src)  # 256 MiB read + 256 MiB written
        torch.cuda.synchronize()
        dt, _ = run_batched(H, cfg, H.STORE_F16, scenes, untimed, args.steps,
                            graph=not args.no_graph)
        log(f"batched fp16 {frames_timed / dt:.0f} frames/s")
        # The same frames through the instrumented eager pass: per-stage
        # kernel times from the launches' own events, and the touched-block
        # counts behind the update's algorithmic bytes.
        ms, calls, tb, alg, tb_all = run_batched_instrumented(
            H, cfg, H.STORE_F16, scenes, untimed, args.steps, 4, with_all=True)
        per = ms / np.maximum(calls, 1)
        W, Hh = cfg["width"], cfg["height"]
        alg_all = tb_all * 512 * 2 * 4 + len(scenes) * W * Hh * 4 + tb_all * 12
        log(f"stage ms per launch: allocate {per[0]:.5f} assign {per[1]:.5f} update {per[2]:.5f}; "
            f"touched blocks per step {tb.mean():.0f}; update algorithmic bytes "
            f"{alg.mean() / 1e6:.1f} MB -> {alg.mean() / (per[2] * 1e-3) / 1e9:.0f} GB/s")
        log(f"every update launch of this run (3 passes over steps 0..{len(tb_all) - 1}): "
            f"touched blocks per step {tb_all.mean():.0f}, algorithmic bytes "
            f"{alg_all.mean() / 1e6:.1f} MB; timed steps {untimed}..{n - 1}: "
            f"{tb.mean():.0f} blocks, {alg.mean() / 1e6:.1f} MB")
        # The allocate's table statistics over the same timed steps (its
        # counting and probe variants, separate kernels in a trace).
        log("allocate statistics " + json.dumps(
            alloc_statistics(H, cfg, scenes, untimed, args.steps, per[0])))
        return

    if args.profile_chain:
        src = torch.empty(1 << 26, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        dst.copy_(src)  # 256 MiB read + 256 MiB written
        torch.cuda.synchronize()
        dt, verts = run_chain(H, cfg, H.STORE_F16, scenes, untimed, args.steps, True)
        log(f"chain batched fp16 {frames_timed / dt:.0f} frames/s")
        per, verts, blocks = run_chain_instrumented(H, cfg, H.STORE_F16, scenes, n - 3, 3)
        out_b = sum(verts) * 24
        log(f"stage ms per step: allocate {per[0]:.5f} assign {per[1]:.5f} update {per[2]:.5f} "
            f"extract {per[3]:.5f} (last 3 steps); last step {sum(verts)} vertices, "
            f"{out_b / 1e6:.1f} MB of mesh, {blocks} blocks allocated -> "
            f"{out_b / (per[3] * 1e-3) / 1e9:.0f} GB/s over the extract stage")
        return

    log("batched fp16 (timed)")
    dt_b16, blocks = run_batched(H, cfg, H.STORE_F16, scenes, untimed, args.steps,
                                 graph=not args.no_graph)
    log("plain fp32 / fp16, eager and one-scene graph (timed)")
    dt_p32, _ = run_plain(H, cfg, H.STORE_F32, scenes, untimed, args.steps)
    dt_p32g = run_plain_graph(H, cfg, H.STORE_F32, scenes, untimed, args.steps)
    dt_p16, _ = run_plain(H, cfg, H.STORE_F16, scenes, untimed, args.steps)
    dt_p16g = run_plain_graph(H, cfg, H.STORE_F16, scenes, untimed, args.steps)
    log("batched fp32 (timed)")
    dt_b32, _ = run_batched(H, cfg, H.STORE_F32, scenes, untimed, args.steps,
                            graph=not args.no_graph)
    fps = {k: frames_timed / v for k, v in
           (("batched_fp16", dt_b16), ("plain_fp32", dt_p32), ("plain_fp32_graph", dt_p32g),
            ("batched_fp32", dt_b32), ("plain_fp16", dt_p16), ("plain_fp16_graph", dt_p16g))}
    log(" ".join(f"{k} {v:.0f}" for k, v in fps.items()))
    best_p32 = max(fps["plain_fp32"], fps["plain_fp32_graph"])
    best_p16 = max(fps["plain_fp16"], fps["plain_fp16_graph"])
    _, st_p32 = run_plain(H, cfg, H.STORE_F32, scenes, untimed, args.steps, timing=True)
    _, st_p16 = run_plain(H, cfg, H.STORE_F16, scenes, untimed, args.steps, timing=True)

    log("chain with extraction every frame (timed)")
    chain_steps = min(args.steps, 50)
    dt_cb, verts_b = run_chain(H, cfg, H.STORE_F16, scenes, untimed, chain_steps, True)
    dt_cp, verts_p = run_chain(H, cfg, H.STORE_F32, scenes, untimed, chain_steps, False)
    chain = {
        "stages": "allocate + assign + update + extract (device mesh) per frame",
        "steps": chain_steps,
        "batched_fp16_frames_per_s": round(chain_steps * args.scenes / dt_cb, 1),
        "plain_fp32_frames_per_s": round(chain_steps * args.scenes / dt_cp, 1),
        "last_mesh_vertices_fp16": verts_b,
        "last_mesh_vertices_fp32": verts_p,
    }
    chain["ratio"] = round(chain["batched_fp16_frames_per_s"] / chain["plain_fp32_frames_per_s"], 3)
    per, verts_i, blocks_i = run_chain_instrumented(H, cfg, H.STORE_F16, scenes,
                                                    untimed + chain_steps - 3, 3)
    chain["batched_fp16_stage_ms_per_step"] = {
        "allocate": round(float(per[0]), 5), "assign": round(float(per[1]), 5),
        "update": round(float(per[2]), 5), "extract": round(float(per[3]), 5)}
    # extract's algorithmic bytes: every allocated block's voxels read once
    # (FP16 + u16: 4 B) and 24 B per vertex written; count and emit in fact
    # read the blocks twice (emit only the non-empty ones), with neighbour faces
    mesh_b = sum(verts_i) * 24
    alg_b = blocks_i * 512 * 4 + mesh_b
    chain["extract_mesh_bytes_per_step"] = mesh_b
    chain["extract_roofline"] = {
        "bound": "hbm", "blocks": blocks_i, "vertices": int(sum(verts_i)),
        "algorithmic_bytes": alg_b,
        "achieved": round(alg_b / (float(per[3]) * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(alg_b / (float(per[3]) * 1e-3) / 8e12, 4),
        "stage": "count + scan + emit of all scenes, events around the launches"}
    log(f"chain batched fp16 {chain['batched_fp16_frames_per_s']:.0f} plain fp32 "
        f"{chain['plain_fp32_frames_per_s']:.0f}")

    log("instrumented batched runs")
    ms16, calls16, tb16, bytes16 = run_batched_instrumented(H, cfg, H.STORE_F16, scenes,
                                                            untimed, args.steps, 4)
    ms32, calls32, tb32, bytes32 = run_batched_instrumented(H, cfg, H.STORE_F32, scenes,
                                                            untimed, args.steps, 8)
    upd_ms16 = ms16[2] / max(calls16[2], 1)
    upd_ms32 = ms32[2] / max(calls32[2], 1)
    achieved16 = bytes16.sum() / (ms16[2] * 1e-3) / 1e9
    achieved32 = bytes32.sum() / (ms32[2] * 1e-3) / 1e9
    log("allocate statistics")
    alloc = alloc_statistics(H, cfg, scenes, untimed, args.steps,
                             float(ms16[0] / max(calls16[0], 1)))
    log("measured ceilings")
    ceil = measured_ceilings(tb16.mean())
    if "pattern_ceiling_us" in ceil:
        ceil["update_over_pattern_ceiling"] = round(float(upd_ms16) * 1e3 / ceil["pattern_ceiling_us"], 3)
    ceil["frac_of_stream_copy"] = round(float(achieved16) / ceil["stream_copy_gbs"], 4)

    workload = None
    if not args.no_workload:
        log("batched workload of uneven scenes")
        workload = run_batched_workload(H, S)

    prec = None
    if not args.no_precision:
        log(f"precision against the FP64 oracle kernel ({args.precision_frames} frames)")
        pf, pd = render_frames(H, S, cfg, args.seed, args.precision_frames)
        prec = precision(H, cfg, pf, pd, args.precision_frames)
        del pf, pd
        torch.cuda.empty_cache()

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(H, S, cfg, scenes[0][0], scenes[0][1])

    out = {
        "metric": "frames/s 640x480@1cm: batched fp16-TSDF vs plain fp32, max err <=0.01 trunc",
        "value": round(fps["batched_fp16"], 1),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt_b16 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded analytic scenes, GPU-rendered depth with noise and dropouts)",
        "config": {"workload": "precise: 6x6x3 m rooms with 20 spheres + 10 boxes, 640x480 depth, "
                               "voxel 1 cm, trunc 4 cm, table 2^22, weight cap 32",
                   "scenes_per_step": args.scenes, "storage": "fp16 tsdf/trunc + u16 weight",
                   "graph": not args.no_graph, "preroll_frames": P,
                   "timed_frames": f"{untimed}..{n - 1} of each scene",
                   "blocks_allocated": int(blocks)},
        "plain_fp32_frames_per_s": round(best_p32, 1),
        "batched_fp16_over_plain_fp32": round(fps["batched_fp16"] / best_p32, 3),
        "batched_fp32_frames_per_s": round(fps["batched_fp32"], 1),
        "plain_fp16_frames_per_s": round(best_p16, 1),
        "attribution": {
            "plain_fp32_eager": round(fps["plain_fp32"], 1),
            "plain_fp32_graph": round(fps["plain_fp32_graph"], 1),
            "plain_fp16_eager": round(fps["plain_fp16"], 1),
            "plain_fp16_graph": round(fps["plain_fp16_graph"], 1),
            "batched_fp32": round(fps["batched_fp32"], 1),
            "batched_fp16": round(fps["batched_fp16"], 1),
            "batching_fp32": round(fps["batched_fp32"] / best_p32, 3),
            "batching_fp16": round(fps["batched_fp16"] / best_p16, 3),
            "fp16_storage_batched": round(fps["batched_fp16"] / fps["batched_fp32"], 3),
            "fp16_storage_plain": round(best_p16 / best_p32, 3),
            "fp16_storage_update_kernel": round(float(upd_ms32 / upd_ms16), 3)},
        "stage_ms_per_step_fp16": {s: round(float(ms16[i] / max(calls16[i], 1)), 5)
                                   for i, s in enumerate(H.STAGES[:3])},
        "stage_ms_per_step_fp32": {s: round(float(ms32[i] / max(calls32[i], 1)), 5)
                                   for i, s in enumerate(H.STAGES[:3])},
        "stage_ms_per_frame_plain_fp32": {s: round(float(st_p32[i]), 5)
                                          for i, s in enumerate(H.STAGES[:3])},
        "stage_ms_per_frame_plain_fp16": {s: round(float(st_p16[i]), 5)
                                          for i, s in enumerate(H.STAGES[:3])},
        "touched_blocks_per_step": round(float(tb16.mean()), 1),
        "update_fp32_gbs": round(float(achieved32), 1),
        "precision_vs_fp64": prec,
        "chain_with_extraction": chain,
        "batched_workload": workload,
        "roofline": {"bound": "hbm", "achieved": round(float(achieved16), 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(float(achieved16 / HBM_PEAK_GBS), 4),
                     "traffic": UPDATE_TRAFFIC["hbm_bytes_per_launch"],
                     "traffic_measured_on": UPDATE_TRAFFIC,
                     "kernel": "k_proj_update_batch<fp16>",
                     "avg_launch_ms": round(float(upd_ms16), 5),
                     "bytes_per_launch": round(float(bytes16.mean()), 0),
                     "measured_ceilings": ceil},
        "alloc_roofline": alloc,
        "cpu_baseline": cpu,
        "autotune": tune_log,
        "autotune_plain": plain_tune_log,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
