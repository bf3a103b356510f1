"""Generate the golden fixtures from the compiled reference CPU arm.

Runs the reference CPU arm (oracle/_ref/libosn_ref.so, compiled from
cpp/src/{volume,marching_tets,c_api}.cpp by oracle/Makefile) single-threaded,
so its float sums are accumulated in point order and are reproducible, on a
few small seeded inputs, and writes inputs and outputs to reference_arm.npz:

  <case>_points           the input points (float32, xyz interleaved)
  <case>_params           voxel, pool, trunc, weight_cap, cam xyz, radius
  <case>_blocks           block count, drop count
  <case>_mesh_n           vertex count of extract(min_weight=0.5, iso=0)
  <case>_mesh_sha256      sha256 of the vertex buffer (posnor float32)
  <case>_mesh_pos_sorted  a sorted subsample of vertex coordinates

The reference's C ABI exposes only counts and the mesh, so those are the
outputs recorded. Single-threaded, the mesh bytes depend on the allocated key
set, the pool (arrival) order and every weighted sum the extractor reads, so
an equal sha256 pins all three.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import bindings as B  # noqa: E402


def cases():
    rng = np.random.default_rng(20261203)
    yield "sphere2cm", B.sphere_points(0.4, 100, 200), (0.02, 1 << 12, -1.0, 32.0, 0, 0, 0, 0.0)
    yield "sphere1cm_nocap", B.sphere_points(0.25, 120, 240), (0.01, 1 << 12, -1.0, 0.0, 0, 0, 0, 0.0)
    cloud = rng.uniform(-0.35, 0.35, size=(6000, 3)).astype(np.float32).ravel()
    yield "cloud_origin", cloud, (0.01, 1 << 14, -1.0, 0.0, 0.013, -0.021, 0.007, 0.0)
    plane = B.plane_points(-0.2, 0.3, 120)
    yield "plane_neg_radius", plane, (0.01, 1 << 13, 0.03, 8.0, 0.05, 0.02, 0.3, 0.58)


def main():
    if not B.ref_available():
        B.build()
    if not B.ref_available():
        sys.exit("reference CPU arm not built (reference sources absent)")
    out = {}
    for name, pts, (voxel, pool, trunc, wcap, cx, cy, cz, radius) in cases():
        ref = B.RefVolume(voxel, pool, trunc, wcap, threads=1)
        assert ref.integrate(pts, (cx, cy, cz), radius) == 0
        mesh = ref.extract(0.5, 0.0)
        out[f"{name}_points"] = pts
        out[f"{name}_params"] = np.array([voxel, pool, trunc, wcap, cx, cy, cz, radius], np.float64)
        out[f"{name}_blocks"] = np.array([ref.block_count, ref.drop_count], np.int64)
        out[f"{name}_mesh_n"] = np.array([len(mesh)], np.int64)
        out[f"{name}_mesh_sha256"] = np.frombuffer(
            hashlib.sha256(np.ascontiguousarray(mesh).tobytes()).digest(), np.uint8)
        # vertex positions, sorted, so a key-set mismatch is also visible here
        out[f"{name}_mesh_pos_sorted"] = np.sort(mesh[:, :3].astype(np.float32).view(np.uint32)
                                                 .view(np.float32).ravel())[:: max(1, len(mesh) // 2000)]
        print(name, ref.block_count, ref.drop_count, len(mesh))
    np.savez_compressed(os.path.join(HERE, "reference_arm.npz"), **out)


if __name__ == "__main__":
    main()
