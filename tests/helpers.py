"""Shared helpers for the parity tests."""
from __future__ import annotations

import numpy as np


def radius_errors(mesh: np.ndarray, R: float):
    r = np.sqrt((mesh[:, :3].astype(np.float64) ** 2).sum(1))
    e = np.abs(r - R)
    return float(e.mean()), float(e.max())


def sorted_triangles(mesh: np.ndarray) -> np.ndarray:
    """Triangles (9 position floats each) in a canonical order, for comparing
    two soups emitted in different orders."""
    tri = mesh[:, :3].reshape(-1, 9)
    view = np.ascontiguousarray(tri).view(np.dtype((np.void, tri.dtype.itemsize * 9))).ravel()
    return tri[np.argsort(view, kind="stable")]


def nearest_distances(a: np.ndarray, b: np.ndarray, cell: float) -> np.ndarray:
    """For each point of a, the distance to the nearest point of b (grid hash)."""
    from collections import defaultdict
    inv = 1.0 / cell
    kb = np.floor(b * inv).astype(np.int64)
    grid = defaultdict(list)
    for i, k in enumerate(map(tuple, kb)):
        grid[k].append(i)
    ka = np.floor(a * inv).astype(np.int64)
    out = np.empty(len(a))
    offs = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)]
    for i, (k, q) in enumerate(zip(map(tuple, ka), a)):
        best = np.inf
        for dx, dy, dz in offs:
            idx = grid.get((k[0] + dx, k[1] + dy, k[2] + dz))
            if idx:
                d = np.sqrt(((b[idx] - q) ** 2).sum(1)).min()
                best = min(best, d)
        out[i] = best
    return out


def voxel_maps(keys: np.ndarray, tsdf: np.ndarray, weight: np.ndarray):
    """key -> (tsdf[512], weight[512])."""
    return {int(k): (tsdf[i], weight[i]) for i, k in enumerate(keys)}


def scene_sdf(scene, p: np.ndarray) -> np.ndarray:
    """Unsigned distance from points p (n,3) to the nearest surface of a
    scenes.Scene (room walls seen from inside, spheres, boxes), FP64."""
    p = np.asarray(p, np.float64)
    lo, hi = np.asarray(scene.room_lo), np.asarray(scene.room_hi)
    d = np.minimum(p - lo, hi - p).min(1)
    d = np.abs(d)
    for s in scene.spheres:
        d = np.minimum(d, np.abs(np.linalg.norm(p - s[:3], axis=1) - s[3]))
    for b in scene.boxes:
        c = (b[:3] + b[3:]) / 2
        h = (b[3:] - b[:3]) / 2
        q = np.abs(p - c) - h
        outside = np.linalg.norm(np.maximum(q, 0.0), axis=1)
        inside = np.minimum(q.max(1), 0.0)
        d = np.minimum(d, np.abs(outside + inside))
    return d


def render_depth_np(scene, R, t, W, H, fx, fy, cx, cy) -> np.ndarray:
    """Closed-form FP64 ray casting, the checker for the device renderer."""
    u, v = np.meshgrid(np.arange(W, dtype=np.float64), np.arange(H, dtype=np.float64))
    dc = np.stack([(u - cx) / fx, (v - cy) / fy, np.ones_like(u)], -1).reshape(-1, 3)
    d = dc @ np.asarray(R, np.float64).T
    o = np.asarray(t, np.float64)
    lo, hi = np.asarray(scene.room_lo), np.asarray(scene.room_hi)
    with np.errstate(divide="ignore", invalid="ignore"):
        tw = np.where(d > 0, (hi - o) / d, np.where(d < 0, (lo - o) / d, np.inf))
    best = tw.min(1)
    for s in scene.spheres:
        oc = o - s[:3]
        a = (d * d).sum(1)
        b = (d * oc).sum(1)
        c = (oc * oc).sum() - s[3] ** 2
        disc = b * b - a * c
        ok = disc >= 0
        sq = np.sqrt(np.where(ok, disc, 0))
        t0 = (-b - sq) / a
        t1 = (-b + sq) / a
        tt = np.where(t0 > 1e-4, t0, t1)
        best = np.where(ok & (tt > 1e-4), np.minimum(best, tt), best)
    for bx in scene.boxes:
        with np.errstate(divide="ignore", invalid="ignore"):
            a = (bx[:3] - o) / d
            b = (bx[3:] - o) / d
        t0 = np.minimum(a, b).max(1)
        t1 = np.maximum(a, b).min(1)
        ok = t0 <= t1
        tt = np.where(t0 > 1e-4, t0, t1)
        best = np.where(ok & (tt > 1e-4), np.minimum(best, tt), best)
    best = np.where(np.isfinite(best), best, 0.0)
    return best.reshape(H, W)
