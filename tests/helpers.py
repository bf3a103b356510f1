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
