"""Benchmark: frames/s of depth-frame TSDF integration at 640x480, 1 cm voxels.

Workload ("precise" scene): a seeded 6x6x3 m room with 20 spheres and 10 boxes,
640x480 planar depth frames rendered on the GPU with N(0, 0.0012 z^2) noise and
1% dropouts, fx=fy=320, cx=320, cy=240; voxel 1 cm, truncation 4 cm, table
2^22 slots, pool 2^21 blocks, weight cap 32. A step is one frame: allocate
(FP32 back-projection + band walk + insert), index assignment, projective
update. All depth frames are resident in HBM before the timed region.

Reported:
  value          frames/s of the FP16-storage pipeline (FP32 arithmetic, FP16
                 normalised TSDF + u16 weight), warmup frames untimed;
  fp32_*         the same frames through plain FP32 storage, for the ratio;
  precision      max / RMS |dtsdf| against the FP64 oracle kernel on the same
                 frames, in truncation units and in voxels;
  roofline       the update kernel: algorithmic bytes (touched blocks x 512
                 voxels x (read + write) x bytes/voxel + depth image + block
                 list) / its average duration from HIP events that its own
                 launches record (start and end of the dispatch, on the batch
                 stream), against 8 TB/s;
  chain_*        the same frames with a device-resident mesh extraction of
                 every scene after every frame (allocate -> assign -> update
                 -> extract), batched fp16 against plain fp32;
  cpu_baseline   the reference CPU arm (compiled from the reference sources,
                 oracle/_ref) integrating the same frames' back-projected points
                 with all host threads, on a bounded sample of frames.
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
# WRITE x 1.0). Those launches are steps 0..29 of the same 16 scenes, so their
# algorithmic bytes are given beside them; counters cannot run inside this
# timed benchmark.
UPDATE_TRAFFIC = {"hbm_bytes_per_launch": 395.4e6, "read": 238.6e6, "write": 156.8e6,
                  "algorithmic_bytes_per_launch": 432.9e6,
                  "launches": "k_proj_update_batch<fp16>, 16 scenes, steps 0..29, "
                              "100.6k touched blocks per launch",
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


KNOBS = {}  # launch shape chosen by the autotuner (empty = library defaults)


def make_volume(H, cfg, storage, hash_kind=None, pool=None):
    """Table cfg["table"] slots; pool table/2 blocks unless given (the tuner's
    short trials use a smaller pool over the same table)."""
    hk = KNOBS.get("hash_kind", 0) if hash_kind is None else hash_kind
    return H.HipVolume(cfg["voxel"], pool or cfg["table"] // 2, cfg["trunc"], 32.0,
                       storage=storage, table_slots=cfg["table"],
                       hash_kind=hk, lds_set=KNOBS.get("lds_set", -1),
                       alloc_grid=KNOBS.get("alloc_grid", 0),
                       update_threads=KNOBS.get("update_threads", 0),
                       update_grid=KNOBS.get("update_grid", 0))


def run_plain(H, cfg, storage, scenes, warmup, steps):
    """One scene at a time, one frame per call: 3 launches per frame."""
    import torch
    total = 0.0
    blocks = 0
    for frames, depth in scenes:
        v = make_volume(H, cfg, storage)
        for k in range(warmup):
            if v.integrate_depth(depth[k], frames[k]):
                raise RuntimeError("integrate failed")
        v.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(warmup, warmup + steps):
            if v.integrate_depth(depth[k], frames[k]):
                raise RuntimeError("integrate failed")
        v.synchronize()
        total += time.perf_counter() - t0
        blocks += v.block_count
        v.close()
    return total, blocks


def make_batch(H, cfg, storage, scenes):
    vols = [make_volume(H, cfg, storage) for _ in scenes]
    b = H.Batch(vols)
    if KNOBS and b.tune(KNOBS["alloc_grid"], KNOBS["update_grid"], KNOBS["update_threads"],
                        KNOBS["lds_set"]):
        raise RuntimeError("batch tune failed")
    n = scenes[0][1].shape[0]
    depths = [[sc[1][k] for sc in scenes] for k in range(n)]
    frames = [[sc[0][k] for sc in scenes] for k in range(n)]
    if b.set_schedule(depths, frames):
        raise RuntimeError("set_schedule failed")
    return b, vols


def run_batched(H, cfg, storage, scenes, warmup, steps, graph=True):
    """All scenes per step in one launch per stage, replayed from a hipGraph."""
    import torch
    b, vols = make_batch(H, cfg, storage, scenes)
    if b.run(warmup, use_graph=graph):
        raise RuntimeError("batch run failed")
    b.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if b.run(steps, use_graph=graph):
        raise RuntimeError("batch run failed")
    b.synchronize()
    dt = time.perf_counter() - t0
    blocks = sum(v.block_count for v in vols)
    b.close()
    for v in vols:
        v.close()
    return dt, blocks


def run_batched_instrumented(H, cfg, storage, scenes, warmup, steps, bytes_per_voxel,
                             with_all=False):
    """Per-stage HIP-event times of back-to-back eager batched steps, and the
    per-step touched-block counts (the same frames give the same touched sets,
    so they are collected in a second, synchronised pass that is not timed);
    with_all: also the counts of every step, warmup included."""
    b, vols = make_batch(H, cfg, storage, scenes)
    b.run(warmup)
    b.synchronize()
    b.read_timing()
    b.set_timing(True)
    b.run(steps)
    ms, calls = b.read_timing()
    b.close()
    for v in vols:
        v.close()
    b, vols = make_batch(H, cfg, storage, scenes)
    touched = []
    for _ in range(warmup + steps):
        b.run(1)
        touched.append(sum(v.touched_count for v in vols))  # waits for the batch stream
    W, Hh = cfg["width"], cfg["height"]
    tb_all = np.array(touched, np.float64)
    tb = tb_all[warmup:]
    alg_bytes = tb * 512 * 2 * bytes_per_voxel + len(scenes) * W * Hh * 4 + tb * 12
    b.close()
    for v in vols:
        v.close()
    if with_all:
        return ms, calls, tb, alg_bytes, tb_all
    return ms, calls, tb, alg_bytes


CHAIN_VERTS = 1 << 25  # per-scene mesh buffer (vertices); noisy 1 cm fields make dense soups


def run_chain(H, cfg, storage, scenes, warmup, steps, batched):
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
        if b.run(warmup, use_graph=True):
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
            v = make_volume(H, cfg, storage)
            vols.append(v)
            for k in range(warmup):
                if v.integrate_depth(depth[k], frames[k]):
                    raise RuntimeError("integrate failed")
            v.synchronize()
            t0 = time.perf_counter()
            for k in range(warmup, warmup + steps):
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


def run_chain_instrumented(H, cfg, storage, scenes, warmup, steps):
    """The batched chain through the eager path with per-stage HIP-event times
    (allocate, assign, update, extract = count + scan + emit of every scene),
    ms per step over `steps` steps after `warmup` untimed ones, and the
    vertices of the last step's meshes (the meshes grow a little per step:
    keep `steps` small for a per-step byte figure) and the blocks allocated."""
    import torch
    n = len(scenes)
    bufs = [torch.empty(CHAIN_VERTS * 6, dtype=torch.float32, device="cuda") for _ in range(n)]
    counts = torch.zeros(n, dtype=torch.int32, device="cuda")
    b, vols = make_batch(H, cfg, storage, scenes)
    if b.set_extract(bufs, [CHAIN_VERTS] * n, counts, 0.5, 0.0, every=1):
        raise RuntimeError("set_extract failed")
    if b.run(warmup):
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
    b.close()
    for v in vols:
        v.close()
    del bufs
    torch.cuda.empty_cache()
    return ms / np.maximum(calls, 1), out, blocks


PRECISION_MESH_VERTS = 40 << 20


def precision(H, cfg, frames, depth, n):
    """FP16 and FP32 states against the FP64 oracle kernel after n frames:
    |dtsdf| over every observed voxel, and the extracted meshes (min weight
    0.5, iso 0) against the FP64 mesh through the mesh-comparison gate
    (Chamfer mean = symmetric mean surface distance, Hausdorff, p99)."""
    import mesh_metrics as MM
    vols = {}
    for name, st in (("f64", H.STORE_F64), ("f32", H.STORE_F32), ("f16", H.STORE_F16)):
        v = make_volume(H, cfg, st)
        for k in range(n):
            v.integrate_depth(depth[k], frames[k])
        v.synchronize()
        vols[name] = v
    meshes = {name: v.extract(0.5, 0.0, cap=PRECISION_MESH_VERTS) for name, v in vols.items()}
    keys = vols["f64"].keys()
    t64, w64 = vols["f64"].read_voxels_f64()
    vols["f64"].close()
    out = {}
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
                     "observed_voxels": int(obs.sum())}
        if meshes[name] is not None and meshes["f64"] is not None:
            ag = MM.compare(meshes[name], meshes["f64"], 200000, cfg["voxel"], 0xC0FFEE)
            out[name]["mesh_vs_fp64"] = {
                "chamfer_mean_voxels": ag.mean_surface / cfg["voxel"],
                "hausdorff_p99_voxels": ag.hausdorff_p99 / cfg["voxel"],
                "hausdorff_voxels": ag.hausdorff / cfg["voxel"],
                "vertices": int(len(meshes[name])), "vertices_fp64": int(len(meshes["f64"])),
                "samples_per_side": 200000}
    return out


def cpu_baseline(H, cfg, frames, depth, budget_s=12.0, max_frames=400):
    """Reference CPU arm (all threads) on the same frames' points."""
    import torch
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import bindings as B
    kind = "reference" if B.ref_available() else "port"
    W, Hh = cfg["width"], cfg["height"]
    buf = torch.empty(W * Hh * 3, dtype=torch.float32, device="cuda")
    if kind == "reference":
        vol = B.RefVolume(cfg["voxel"], cfg["table"] // 2, cfg["trunc"], 32.0)
        threads = vol.threads
    else:
        vol = B.Volume(cfg["voxel"], cfg["table"] // 2, cfg["trunc"], 32.0)
        threads = 1
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
    t = np.sort(np.array(times))
    pct = lambda q: float(t[min(len(t) - 1, int(len(t) * q))])  # nearest rank
    return {"value": n / spent, "unit": "frames/s", "cores": int(threads), "kind": kind,
            "sample": f"frames 1..{n} of the same trajectory (frame 0 warmup), allocate+update "
                      f"per frame on back-projected points, p50 {pct(0.5) * 1e3:.1f} ms, "
                      f"p95 {pct(0.95) * 1e3:.1f} ms"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--scenes", type=int, default=16)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-precision", action="store_true")
    ap.add_argument("--no-graph", action="store_true")
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
    n = args.warmup + args.steps
    log(f"rendering {args.scenes} scenes x {n} frames {cfg['width']}x{cfg['height']}")
    scenes = [render_frames(H, S, cfg, args.seed + i, n) for i in range(args.scenes)]
    frames_timed = args.steps * args.scenes

    tune_log = None
    if args.knobs:
        KNOBS.update(json.loads(args.knobs))
        tune_log = {"chosen": dict(KNOBS), "trials": 0, "seconds": 0.0}
    elif not args.no_tune:
        import autotune
        log("autotuning the launch shape on the first frames")
        n_t = min(n, 30)
        depths = [[sc[1][k] for sc in scenes] for k in range(n_t)]
        frs = [[sc[0][k] for sc in scenes] for k in range(n_t)]
        best, tlog = autotune.tune(
            H, lambda hk: [make_volume(H, cfg, H.STORE_F16, hash_kind=hk, pool=1 << 17)
                           for _ in scenes],
            depths, frs, warmup=min(10, n_t // 3), steps=n_t - min(10, n_t // 3),
            log_path=os.path.join(ROOT, "gpurun_out", "autotune.json")
            if os.path.isdir(os.path.join(ROOT, "gpurun_out")) else None)
        KNOBS.update(best)
        tune_log = {"chosen": best, "trials": len(tlog["trials"]), "seconds": tlog["seconds"]}
        log(f"chosen {best}")

    if args.profile_only:
        src = torch.empty(1 << 26, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        dst.copy_(src)  # 256 MiB read + 256 MiB written
        torch.cuda.synchronize()
        dt, _ = run_batched(H, cfg, H.STORE_F16, scenes, args.warmup, args.steps,
                            graph=not args.no_graph)
        log(f"batched fp16 {frames_timed / dt:.0f} frames/s")
        # The same frames through the instrumented eager pass: per-stage
        # kernel times from the launches' own events, and the touched-block
        # counts behind the update's algorithmic bytes.
        ms, calls, tb, alg, tb_all = run_batched_instrumented(
            H, cfg, H.STORE_F16, scenes, args.warmup, args.steps, 4, with_all=True)
        per = ms / np.maximum(calls, 1)
        W, Hh = cfg["width"], cfg["height"]
        alg_all = tb_all * 512 * 2 * 4 + len(scenes) * W * Hh * 4 + tb_all * 12
        log(f"stage ms per launch: allocate {per[0]:.5f} assign {per[1]:.5f} update {per[2]:.5f}; "
            f"touched blocks per step {tb.mean():.0f}; update algorithmic bytes "
            f"{alg.mean() / 1e6:.1f} MB -> {alg.mean() / (per[2] * 1e-3) / 1e9:.0f} GB/s")
        log(f"every update launch of this run (3 passes over steps 0..{len(tb_all) - 1}): "
            f"touched blocks per step {tb_all.mean():.0f}, algorithmic bytes "
            f"{alg_all.mean() / 1e6:.1f} MB")
        return

    if args.profile_chain:
        src = torch.empty(1 << 26, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        src.fill_(1.0)
        dst.copy_(src)  # 256 MiB read + 256 MiB written
        torch.cuda.synchronize()
        dt, verts = run_chain(H, cfg, H.STORE_F16, scenes, args.warmup, args.steps, True)
        log(f"chain batched fp16 {frames_timed / dt:.0f} frames/s")
        per, verts, _ = run_chain_instrumented(H, cfg, H.STORE_F16, scenes,
                                               args.warmup + args.steps - 3, 3)
        out_b = sum(verts) * 24
        log(f"stage ms per step: allocate {per[0]:.5f} assign {per[1]:.5f} update {per[2]:.5f} "
            f"extract {per[3]:.5f} (last 3 steps); last step {sum(verts)} vertices, "
            f"{out_b / 1e6:.1f} MB of mesh -> {out_b / (per[3] * 1e-3) / 1e9:.0f} GB/s over "
            f"the extract stage")
        return

    log("batched fp16 (timed)")
    dt_b16, blocks = run_batched(H, cfg, H.STORE_F16, scenes, args.warmup, args.steps,
                                 graph=not args.no_graph)
    log("plain fp32 (timed)")
    dt_p32, _ = run_plain(H, cfg, H.STORE_F32, scenes, args.warmup, args.steps)
    log("batched fp32 / plain fp16 (timed)")
    dt_b32, _ = run_batched(H, cfg, H.STORE_F32, scenes, args.warmup, args.steps,
                            graph=not args.no_graph)
    dt_p16, _ = run_plain(H, cfg, H.STORE_F16, scenes, args.warmup, args.steps)
    fps = {k: frames_timed / v for k, v in
           (("batched_fp16", dt_b16), ("plain_fp32", dt_p32), ("batched_fp32", dt_b32),
            ("plain_fp16", dt_p16))}
    log(" ".join(f"{k} {v:.0f}" for k, v in fps.items()))

    log("chain with extraction every frame (timed)")
    chain_steps = min(args.steps, 50)
    dt_cb, verts_b = run_chain(H, cfg, H.STORE_F16, scenes, args.warmup, chain_steps, True)
    dt_cp, verts_p = run_chain(H, cfg, H.STORE_F32, scenes, args.warmup, chain_steps, False)
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
                                                    args.warmup + chain_steps - 3, 3)
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
        "bound": "hbm", "blocks": blocks_i, "algorithmic_bytes": alg_b,
        "achieved": round(alg_b / (float(per[3]) * 1e-3) / 1e9, 1), "peak": 8000.0,
        "unit": "GB/s", "frac": round(alg_b / (float(per[3]) * 1e-3) / 8e12, 4),
        "stage": "count + scan + emit of all 16 scenes, events around the launches"}
    log(f"chain batched fp16 {chain['batched_fp16_frames_per_s']:.0f} plain fp32 "
        f"{chain['plain_fp32_frames_per_s']:.0f}")

    log("instrumented batched runs")
    ms16, calls16, tb16, bytes16 = run_batched_instrumented(H, cfg, H.STORE_F16, scenes,
                                                            args.warmup, args.steps, 4)
    ms32, calls32, tb32, bytes32 = run_batched_instrumented(H, cfg, H.STORE_F32, scenes,
                                                            args.warmup, args.steps, 8)
    upd_ms16 = ms16[2] / max(calls16[2], 1)
    achieved16 = bytes16.sum() / (ms16[2] * 1e-3) / 1e9
    achieved32 = bytes32.sum() / (ms32[2] * 1e-3) / 1e9

    prec = None
    if not args.no_precision:
        log("precision against the FP64 oracle kernel")
        prec = precision(H, cfg, scenes[0][0], scenes[0][1], n)

    cpu = None
    if not args.no_cpu:
        log("cpu baseline")
        cpu = cpu_baseline(H, cfg, scenes[0][0], scenes[0][1])

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
                   "graph": not args.no_graph, "frames_per_scene": n,
                   "blocks_allocated": int(blocks)},
        "plain_fp32_frames_per_s": round(fps["plain_fp32"], 1),
        "batched_fp16_over_plain_fp32": round(fps["batched_fp16"] / fps["plain_fp32"], 3),
        "batched_fp32_frames_per_s": round(fps["batched_fp32"], 1),
        "plain_fp16_frames_per_s": round(fps["plain_fp16"], 1),
        "stage_ms_per_step_fp16": {s: round(float(ms16[i] / max(calls16[i], 1)), 5)
                                   for i, s in enumerate(H.STAGES[:3])},
        "stage_ms_per_step_fp32": {s: round(float(ms32[i] / max(calls32[i], 1)), 5)
                                   for i, s in enumerate(H.STAGES[:3])},
        "touched_blocks_per_step": round(float(tb16.mean()), 1),
        "update_fp32_gbs": round(float(achieved32), 1),
        "precision_vs_fp64": prec,
        "chain_with_extraction": chain,
        "roofline": {"bound": "hbm", "achieved": round(float(achieved16), 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(float(achieved16 / HBM_PEAK_GBS), 4),
                     "traffic": UPDATE_TRAFFIC["hbm_bytes_per_launch"],
                     "traffic_measured_on": UPDATE_TRAFFIC,
                     "kernel": "k_proj_update_batch<fp16>",
                     "avg_launch_ms": round(float(upd_ms16), 5),
                     "bytes_per_launch": round(float(bytes16.mean()), 0)},
        "cpu_baseline": cpu,
        "autotune": tune_log,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
