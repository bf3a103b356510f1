"""Pins for the CPU oracle (oracle/tsdf_oracle.c), on the CPU.

Each check is chosen so that a plausible mistake in the oracle fails it:
  * key codec / floor division / hash: closed forms and brute force over small
    cubes and the extreme coordinates (types.hpp:52-69,150-165);
  * the allocate/update/extract path: the reference's published known answers
    (block and vertex counts per synthetic cell, results/sweep_*/sweep.csv and
    docs/geometry/manifest.json) and its analytic gates
    (cpp/tests/test_analytic.cpp:64-203);
  * the FP32 back-projection and projective decision: the pixel convention of
    crates/tartanair/src/lib.rs:323-326 and agreement with an FP64 evaluation.
"""
import math

import numpy as np
import pytest

import bindings as B


def test_pack_coord_closed_form_and_roundtrip():
    L = B.lib()
    rng = np.random.default_rng(1)
    cases = [(0, 0, 0), (-1, -1, -1), ((1 << 20) - 1, -(1 << 20), 5), (-(1 << 20), (1 << 20) - 1, -7)]
    cases += [tuple(int(x) for x in rng.integers(-(1 << 20), 1 << 20, 3)) for _ in range(500)]
    cases += [(x, y, z) for x in range(-4, 5) for y in range(-4, 5) for z in range(-4, 5)]
    out = np.zeros(3, np.int32)
    keys = set()
    for x, y, z in cases:
        k = L.orc_pack_coord(x, y, z)
        assert k == ((x + (1 << 20)) << 42) | ((y + (1 << 20)) << 21) | (z + (1 << 20))
        assert 0 <= k < (1 << 63)
        L.orc_unpack_key(k, out.ctypes.data_as(B.i32p))
        assert tuple(out) == (x, y, z)
        keys.add(k)
    assert len(keys) == len(set(cases))  # injective


def test_floor_div_brute_force():
    L = B.lib()
    for a in range(-200, 201):
        for b in range(1, 10):
            assert L.orc_floor_div(a, b) == math.floor(a / b)


def test_hash_closed_form():
    L = B.lib()
    rng = np.random.default_rng(2)
    for _ in range(300):
        x, y, z = (int(v) for v in rng.integers(-(1 << 20), 1 << 20, 3))
        mask = (1 << int(rng.integers(1, 25))) - 1
        h = (((x * 73856093) & 0xFFFFFFFF) ^ ((y * 19349663) & 0xFFFFFFFF) ^
             ((z * 83492791) & 0xFFFFFFFF)) & mask
        assert L.orc_hash_block(x, y, z, mask) == h


@pytest.mark.parametrize("R,a,b,pool,blocks,verts", [
    (0.5, 400, 800, 1 << 14, 1160, 846288),   # base
    (0.5, 100, 200, 1 << 14, 1104, 7632),     # pts-20k
    (0.5, 200, 400, 1 << 14, 1160, 648336),   # pts-80k
    (0.25, 400, 800, 1 << 14, 280, 212544),   # r-0.25
    (2.0, 400, 800, 1 << 17, 18160, 124722),  # r-2.0
])
def test_published_sphere_cells(R, a, b, pool, blocks, verts):
    v = B.Volume(0.01, pool, -1.0, 32.0)
    v.integrate(B.sphere_points(R, a, b), (0, 0, 0))
    assert v.block_count == blocks
    assert v.drop_count == 0
    assert v.count_mesh(0.5, 0.0) == verts


def test_published_plane_cell():
    """plane-320k: plane(0.5, 1.0, 566) seen from the origin."""
    v = B.Volume(0.01, 1 << 14, -1.0, 32.0)
    v.integrate(B.plane_points(0.5, 1.0, 566), (0, 0, 0))
    assert v.block_count == 1060
    assert v.count_mesh(0.5, 0.0) == 950424


def test_analytic_sphere_and_plane_gates():
    v = B.Volume(0.01, 1 << 14, -1.0, 32.0)
    v.integrate(B.sphere_points(0.5, 400, 800), (0, 0, 0))
    m = v.extract(0.5, 0.0)
    r = np.sqrt((m[:, :3].astype(np.float64) ** 2).sum(1))
    assert np.abs(r - 0.5).mean() < 0.0025 and np.abs(r - 0.5).max() < 0.01
    # normals are unit length and point along the radius (outward or inward)
    n = m[:, 3:6].astype(np.float64)
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-5)
    p = v = B.Volume(0.01, 1 << 14, -1.0, 32.0)
    p.integrate(B.plane_points(0.30, 0.25, 500), (0, 0, 0))
    m = p.extract(0.5, 0.0)
    dz = np.abs(m[:, 2].astype(np.float64) - 0.30)
    assert len(m) > 1000 and dz.mean() < 0.0025 and dz.max() < 0.01


def test_pool_exhaustion_and_bad_arguments():
    v = B.Volume(0.005, 16, -1.0, 32.0)
    v.integrate(B.sphere_points(0.5, 200, 400), (0, 0, 0))
    assert v.drop_count > 0 and v.block_count <= 16
    L = B.lib()
    assert L.orc_allocate_blocks(v.h, None, 10, 0, 0, 0, 0) == 1
    assert L.orc_create(0.0, 10, -1.0, 1.0) is None
    assert L.orc_create(0.01, 0, -1.0, 1.0) is None
    # extract with too small a buffer reports overflow
    v2 = B.Volume(0.02, 1 << 12, -1.0, 32.0)
    v2.integrate(B.sphere_points(0.4, 100, 200), (0, 0, 0))
    n = v2.count_mesh()
    assert n > 0 and v2.extract(cap=n - 1) is None and len(v2.extract(cap=n)) == n


def test_truncation_default_and_steps():
    """trunc <= 0 selects 4 voxels (types.hpp:97-106); band = 2*4+1 samples
    per point, so a single point seen head-on yields 9 visits."""
    v = B.Volume(0.01, 64, -1.0, 0.0)
    assert abs(v.trunc - 0.04) < 1e-9
    pt = np.array([0.505, 0.005, 0.005], np.float32)
    assert v.count_visits(pt, (0, 0, 0)) == 9


def test_backprojection_convention_and_fp64_agreement():
    rng = np.random.default_rng(3)
    W, H = 64, 48
    fx, fy, cx, cy = 32.0, 32.0, 32.0, 24.0
    depth = rng.uniform(0.5, 4.0, size=(H, W)).astype(np.float32)
    depth[5, 7] = 0.0
    depth[9, 9] = np.nan
    ang = 0.3
    R = np.array([[math.cos(ang), -math.sin(ang), 0], [math.sin(ang), math.cos(ang), 0], [0, 0, 1]])
    t = np.array([0.1, -0.2, 1.3])
    pts = B.backproject_f32(depth, fx, fy, cx, cy, R, t)
    valid = np.isfinite(depth) & (depth > 0)
    assert len(pts) == valid.sum() * 3
    vv, uu = np.nonzero(valid)
    z = depth[valid].astype(np.float64)
    pc = np.stack([(uu - cx) * z / fx, (vv - cy) * z / fy, z], 1)
    pw = pc @ R.T + t
    err = np.abs(pts.reshape(-1, 3) - pw)
    assert err.max() < 2e-6 * (np.abs(pw).max() + 1)
    # the optical-axis pixel (cx, cy) lands on the camera's z axis
    k = np.nonzero((uu == 32) & (vv == 24))[0][0]
    assert np.allclose(pts.reshape(-1, 3)[k], R[:, 2] * depth[24, 32] + t, atol=1e-5)


def _rot(ax_deg, ay_deg):
    a, b = np.radians(ax_deg), np.radians(ay_deg)
    rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    return (rx @ ry).astype(np.float32)


# With the identity pose the depth is constant along a run of 4 voxels in x;
# the rotated poses make it vary, so the run's batch-inverted reciprocals are
# told apart (a swapped or misindexed reciprocal moves the pixel).
@pytest.mark.parametrize("ax,ay", [(0.0, 0.0), (10.0, 35.0), (-25.0, -50.0)])
def test_projective_decision_matches_fp64_projection(ax, ay):
    L = B.lib()
    rng = np.random.default_rng(4)
    W, H = 80, 60
    fx, fy, cx, cy = 40.0, 40.0, 40.0, 30.0
    depth = rng.uniform(0.8, 3.0, size=(H, W)).astype(np.float32)
    Rcw = _rot(ax, ay)
    c = np.array([0.0, 0.0, -1.0], np.float32)
    voxel, trunc = 0.02, 0.08
    x = np.zeros(1, np.float32)
    agree = 0
    total = 0
    for _ in range(6000):
        b = rng.integers(-6, 6, 3)
        l = rng.integers(0, 8, 3)
        ok = L.orc_proj_observe_f32(Rcw.ctypes.data_as(B.f32p), c.ctypes.data_as(B.f32p), fx, fy,
                                    cx, cy, W, H, depth.ctypes.data_as(B.f32p), voxel, trunc,
                                    int(b[0]), int(b[1]), int(b[2]), int(l[0]), int(l[1]),
                                    int(l[2]), x.ctypes.data_as(B.f32p))
        p = Rcw.astype(np.float64) @ ((b * 8 + l + 0.5) * voxel - c.astype(np.float64))
        if p[2] <= 1e-3:
            assert not ok
            continue
        u = fx * p[0] / p[2] + cx + 0.5
        v = fy * p[1] / p[2] + cy + 0.5
        inside = 0 <= u < W and 0 <= v < H
        near_edge = min(abs(u - round(u)), abs(v - round(v))) < 1e-4
        if not inside or near_edge:
            continue
        d = float(depth[int(math.floor(v)), int(math.floor(u))])
        sdf = d - p[2]
        if abs(sdf + trunc) < 1e-5:
            continue
        total += 1
        if sdf < -trunc:
            assert not ok
        else:
            assert ok
            assert abs(x[0] - min(sdf / trunc, 1.0)) < 1e-5
        agree += 1
    assert total > 1000 and agree == total
