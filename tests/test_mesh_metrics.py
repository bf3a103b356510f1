"""The mesh-comparison gate against the checks the reference runs on its own
metric (crates/mesh-metrics/src/bin/verify_metrics.rs:60-131) and against
closed forms: distances of points to triangles computed independently, the
area weighting of the sampler, and determinism. CPU only (host code)."""
import math

import numpy as np
import pytest

import mesh_metrics as M


def sphere(radius, n_lat, n_lon, offset=(0.0, 0.0, 0.0)):
    """Latitude/longitude sphere soup, as verify_metrics.rs:21-50 builds it."""
    out = []

    def p(lat, lon):
        th = math.pi * lat / n_lat
        ph = 2.0 * math.pi * lon / n_lon
        return (offset[0] + radius * math.sin(th) * math.cos(ph),
                offset[1] + radius * math.sin(th) * math.sin(ph),
                offset[2] + radius * math.cos(th))

    for lat in range(n_lat):
        for lon in range(n_lon):
            a, b, c, d = p(lat, lon), p(lat + 1, lon), p(lat + 1, lon + 1), p(lat, lon + 1)
            out += [a, b, c, a, c, d]
    return np.array(out, np.float32)


@pytest.fixture(scope="module")
def spheres():
    return sphere(1.00, 64, 128), sphere(1.05, 64, 128), sphere(1.00, 64, 128, (0.012, 0, 0))


def test_concentric_spheres_measure_their_gap(spheres):
    a, b, _ = spheres
    ag = M.compare(a, b, 60000, 0.02, 0xC0FFEE)
    assert abs(ag.mean_surface - 0.05) < 1e-3  # verify_metrics.rs:73-80
    # the chordal tessellation keeps every distance within its sagitta of 0.05
    assert 0.045 < ag.a_to_b.mean < 0.051 and 0.045 < ag.b_to_a.mean < 0.051


def test_identical_meshes_agree_exactly(spheres):
    a, _, _ = spheres
    ag = M.compare(a, a, 60000, 0.02, 0xC0FFEE)
    # point-to-triangle: a sample on a surface is at distance 0 from it
    assert ag.mean_surface < 1e-9 and ag.hausdorff < 1e-6  # reference bound: 2e-3


def test_gate_passes_and_fails_as_the_reference(spheres):
    a, b, c = spheres
    tol = M.Tolerance.for_voxel(0.01)
    assert (tol.mean_surface_m, tol.hausdorff_p99_m, tol.vertex_count_rel) == (0.001, 0.01, 0.02)
    assert M.gate(a, a, tol, 60000, 0.02).passed
    g = M.gate(a, b, tol, 60000, 0.02)
    assert not g.passed and g.failures & M.FAIL_MEAN and g.failures & M.FAIL_P99
    g = M.gate(a, c, tol, 60000, 0.02)
    assert not g.passed and g.failures & M.FAIL_MEAN


def _closest_on_triangle_np(p, a, b, c):
    """Independent closed form: minimise |p - (a + s*ab + t*ac)|^2 over the
    triangle by the unconstrained solution when inside, else the best of the
    three edge segments."""
    ab, ac = b - a, c - a
    G = np.array([[ab @ ab, ab @ ac], [ab @ ac, ac @ ac]])
    rhs = np.array([ab @ (p - a), ac @ (p - a)])
    s, t = np.linalg.solve(G, rhs)
    if s >= 0 and t >= 0 and s + t <= 1:
        q = a + s * ab + t * ac
        return float((p - q) @ (p - q))
    best = np.inf
    for u, v in ((a, b), (b, c), (c, a)):
        d = v - u
        k = np.clip(((p - u) @ d) / (d @ d), 0.0, 1.0)
        q = u + k * d
        best = min(best, float((p - q) @ (p - q)))
    return best


def test_point_triangle_distance_closed_form():
    rng = np.random.default_rng(3)
    tris = rng.normal(size=(60, 3, 3)).astype(np.float32)
    soup = tris.reshape(-1, 3)
    q = rng.normal(scale=2.0, size=(400, 3))
    got = M.nearest_sq(soup, q, cell=0.25, threads=2)
    for i in range(len(q)):
        want = min(_closest_on_triangle_np(q[i], *t.astype(np.float64)) for t in tris)
        assert got[i] == pytest.approx(want, rel=1e-9, abs=1e-12)


def test_sampler_is_area_weighted_on_surface_and_deterministic():
    # two triangles with areas 1 and 3 in the z = 0 plane
    t1 = [(0, 0, 0), (1, 0, 0), (0, 2, 0)]
    t2 = [(5, 0, 0), (8, 0, 0), (5, 2, 0)]
    soup = np.array(t1 + t2, np.float32)
    s = M.sample_surface(soup, 40000, 7)
    assert s.shape == (40000, 3) and np.all(s[:, 2] == 0.0)
    frac2 = np.mean(s[:, 0] >= 5.0)
    assert abs(frac2 - 0.75) < 0.01
    inside1 = s[s[:, 0] < 5.0]
    assert np.all(inside1[:, 0] / 1.0 + inside1[:, 1] / 2.0 <= 1.0 + 1e-12)
    assert np.array_equal(s, M.sample_surface(soup, 40000, 7))
    assert not np.array_equal(s, M.sample_surface(soup, 40000, 8))
    # a zero-area soup stands in its vertices
    flat = np.zeros((6, 3), np.float32)
    assert M.sample_surface(flat, 10, 1).shape == (6, 3)


def test_position_normal_records_are_read_by_stride(spheres):
    a, b, _ = spheres
    a6 = np.concatenate([a, np.ones_like(a)], axis=1)  # 6 floats per vertex
    ag3 = M.compare(a, b, 5000, 0.02, 11)
    ag6 = M.compare(a6, b, 5000, 0.02, 11)
    assert ag3.mean_surface == ag6.mean_surface and ag3.hausdorff == ag6.hausdorff
