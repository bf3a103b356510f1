"""The HIP arm through the reference-shaped C ABI, against the oracle.

Reference gates restated (cpp/tests/test_analytic.cpp:64-203,
cpp/tests/test_cuda_analytic.cu:185-283) plus the stricter parity this build
claims: the allocated key set bit-exact, weights exact and tsdf sums within
1e-5 of sum|terms| (unit weights, so sum|terms| <= weight) against the
sequential oracle at weight_cap = 0, and the extracted mesh byte-identical to
the oracle extractor run over the same voxel state in the same pool order.
"""
import numpy as np
import pytest

import bindings as B
from helpers import radius_errors, sorted_triangles

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def H():
    import osn_tsdf_hip
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    return osn_tsdf_hip


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_sphere_keys_and_sums_bit_parity(H):
    """Sphere R=0.5, 400x800, voxel 1 cm, pool 2^14, cap disabled."""
    pts = B.sphere_points(0.5, 400, 800)
    gv = H.HipVolume(0.01, 1 << 14, -1.0, 0.0)
    assert gv.integrate(dev(pts), (0, 0, 0)) == 0
    ov = B.Volume(0.01, 1 << 14, -1.0, 0.0)
    ov.integrate(pts, (0, 0, 0))
    print("blocks gpu", gv.block_count, "oracle", ov.block_count)
    assert gv.block_count == ov.block_count == 1160
    assert gv.drop_count == 0
    gk = gv.keys()
    assert np.array_equal(np.sort(gk), np.sort(ov.keys()))
    # pool order is the sorted key order
    assert np.all(np.diff(gk) > 0)
    t, w = gv.read_voxels()
    for i, k in enumerate(gk):
        ot, ow, orr, og, ob = ov.block_voxels(int(k))
        assert np.array_equal(w[i], ow), f"weights differ in block {k}"
        tol = 1e-5 * np.maximum(ow, 1.0)
        assert np.all(np.abs(t[i] - ot) <= tol), f"tsdf sums differ in block {k}"


# Weight caps: the reference skips a visit once the voxel's weight has reached
# the cap (cpp/src/volume.cpp:283-289, types.hpp:99-101); its sequential CPU
# arm therefore ends every voxel at min(cap, visits) for an integer cap, and
# ceil(cap) for a fractional one. The device takes a call's visits of a voxel
# against the weight the voxel has (all at once, or per workgroup where they
# are added directly), so it ends at the same counts whatever the order of its
# workgroups. Three integrates of the sphere: most surface voxels pass
# cap 8 within one call, many pass 32.5 by the third.
@pytest.mark.parametrize("cap", [8.0, 32.5])
def test_scatter_weight_cap_counts_exact(H, cap):
    pts = B.sphere_points(0.5, 400, 800)
    d = dev(pts)
    gv = H.HipVolume(0.01, 1 << 14, -1.0, cap)
    oc = B.Volume(0.01, 1 << 14, -1.0, cap)   # sequential, capped
    ou = B.Volume(0.01, 1 << 14, -1.0, 0.0)   # uncapped: visit counts and sums
    for _ in range(3):
        assert gv.integrate(d, (0, 0, 0)) == 0
        oc.integrate(pts, (0, 0, 0))
        ou.integrate(pts, (0, 0, 0))
    gk = gv.keys()
    assert np.array_equal(np.sort(gk), np.sort(oc.keys()))
    t, w = gv.read_voxels()
    wmax = float(np.ceil(cap))
    n_sat = n_free = 0
    dev_means = []
    for i, k in enumerate(gk):
        cw = oc.block_voxels(int(k))[1]
        ut, uw = ou.block_voxels(int(k))[:2]
        assert np.array_equal(w[i], cw), f"capped weights differ in block {k}"
        free = uw <= wmax  # never capped: the full sums
        assert np.all(np.abs(t[i][free] - ut[free]) <= 1e-5 * np.maximum(uw[free], 1.0))
        sat = uw > wmax
        assert np.all(w[i][sat] == wmax)
        n_sat += int(sat.sum())
        n_free += int((free & (uw > 0)).sum())
        if sat.any():
            dev_means.append(np.abs(t[i][sat] / wmax))
    m = np.concatenate(dev_means)
    print(f"cap {cap}: {n_sat} saturated voxels, {n_free} below the cap; |mean| max {m.max():.3f}")
    assert n_sat > 10000 and n_free > 1000
    assert m.max() <= 1.0 + 1e-6  # a convex weighting of observations in [-1, 1]


# Dense clouds: the 1.28M-point sphere gives most blocks more than 64
# workgroup segments in one call, so the visit pass adds part of their visits
# directly; 2.2M points take two chunks of the two-pass update in one call.
# Weights must still equal the sequential oracle's, capped or not, after two
# calls (the second takes the visit pass's smaller LDS set: the first call
# saw many points per touched block). Near the
# poles a voxel collects thousands of visits, where two FP32 sums of the same
# n terms in different orders may differ by up to 2 (n - 1) 2^-24 sum|terms|
# (sum|terms| <= n = the weight), so the tolerance grows to that bound.
def sum_tol(w):
    w = np.maximum(w.astype(np.float64), 1.0)
    return np.maximum(1e-5 * w, 2.0 * 2.0 ** -24 * w * w)


@pytest.mark.parametrize("a,b,cap", [(800, 1600, 0.0), (1100, 2000, 0.0), (1100, 2000, 8.0)])
def test_scatter_dense_clouds_exact(H, a, b, cap):
    pts = B.sphere_points(0.5, a, b)
    gv = H.HipVolume(0.01, 1 << 14, -1.0, cap)
    ov = B.Volume(0.01, 1 << 14, -1.0, cap)
    ou = B.Volume(0.01, 1 << 14, -1.0, 0.0) if cap else ov
    d = dev(pts)
    for _ in range(2):
        assert gv.integrate(d, (0, 0, 0)) == 0
        ov.integrate(pts, (0, 0, 0))
        if cap:
            ou.integrate(pts, (0, 0, 0))
    gk = gv.keys()
    assert np.array_equal(np.sort(gk), np.sort(ov.keys()))
    t, w = gv.read_voxels()
    n_vis = 0
    for i, k in enumerate(gk):
        ow = ov.block_voxels(int(k))[1]
        ut, uw = ou.block_voxels(int(k))[:2]
        assert np.array_equal(w[i], ow), f"weights differ in block {k}"
        free = uw <= (np.ceil(cap) if cap else np.inf)
        err = np.abs(t[i][free].astype(np.float64) - ut[free])
        assert np.all(err <= sum_tol(uw[free])), f"tsdf sums differ in block {k}"
        n_vis += int(uw.sum())
    n_pts = len(pts) // 3
    print(f"{n_pts} points x 2 calls, {n_vis} visits, {len(gk)} blocks")
    assert n_vis > 16 * n_pts


# A wider band: trunc 8 voxels gives 17 samples a point, 4,352 visits per
# workgroup against the visit pass's 3,072-slot LDS set, so crowded sets add
# the rest directly; weights still exact against the oracle.
@pytest.mark.parametrize("cap", [0.0, 8.0])
def test_scatter_wide_band_exact(H, cap):
    pts = B.sphere_points(0.5, 400, 800)
    gv = H.HipVolume(0.01, 1 << 15, 0.08, cap)
    assert gv.integrate(dev(pts), (0, 0, 0)) == 0
    ov = B.Volume(0.01, 1 << 15, 0.08, cap)
    ov.integrate(pts, (0, 0, 0))
    ou = B.Volume(0.01, 1 << 15, 0.08, 0.0) if cap else ov
    if cap:
        ou.integrate(pts, (0, 0, 0))
    gk = gv.keys()
    assert np.array_equal(np.sort(gk), np.sort(ov.keys()))
    t, w = gv.read_voxels()
    for i, k in enumerate(gk):
        ow = ov.block_voxels(int(k))[1]
        ut, uw = ou.block_voxels(int(k))[:2]
        assert np.array_equal(w[i], ow), f"weights differ in block {k}"
        free = uw <= (np.ceil(cap) if cap else np.inf)
        err = np.abs(t[i][free].astype(np.float64) - ut[free])
        assert np.all(err <= sum_tol(uw[free])), f"tsdf sums differ in block {k}"
    print(f"{len(gk)} blocks, trunc 8 voxels, cap {cap}")


# Calls of up to 2^17 points take one sample per lane; the sums and weights
# are those of the band layout. Two calls, capped and not, a 4 cm and an 8 cm
# band.
@pytest.mark.parametrize("trunc,cap", [(-1.0, 0.0), (-1.0, 8.0), (0.08, 0.0)])
def test_scatter_small_calls_exact(H, trunc, cap):
    pts = B.sphere_points(0.5, 200, 400)  # 80,000 points
    gv = H.HipVolume(0.01, 1 << 15, trunc, cap)
    ov = B.Volume(0.01, 1 << 15, trunc, cap)
    ou = B.Volume(0.01, 1 << 15, trunc, 0.0) if cap else ov
    d = dev(pts)
    for _ in range(2):
        assert gv.integrate(d, (0, 0, 0)) == 0
        ov.integrate(pts, (0, 0, 0))
        if cap:
            ou.integrate(pts, (0, 0, 0))
    gk = gv.keys()
    assert np.array_equal(np.sort(gk), np.sort(ov.keys()))
    t, w = gv.read_voxels()
    n_vis = 0
    for i, k in enumerate(gk):
        ow = ov.block_voxels(int(k))[1]
        ut, uw = ou.block_voxels(int(k))[:2]
        assert np.array_equal(w[i], ow), f"weights differ in block {k}"
        free = uw <= (np.ceil(cap) if cap else np.inf)
        err = np.abs(t[i][free].astype(np.float64) - ut[free])
        assert np.all(err <= sum_tol(uw[free])), f"tsdf sums differ in block {k}"
        n_vis += int(uw.sum())
    assert n_vis > 2 * 8 * (len(pts) // 3)


def test_sphere_analytic_gate_and_known_answer(H):
    """cpp/tests/test_cuda_analytic.cu:185-227 with the default cap of 32;
    1,160 blocks and 846,288 vertices are the published counts
    (docs/geometry/manifest.json, cell `base`)."""
    pts = B.sphere_points(0.5, 400, 800)
    gv = H.HipVolume(0.01, 1 << 14, -1.0, 32.0)
    assert gv.integrate(dev(pts), (0, 0, 0)) == 0
    mesh = gv.extract(0.5, 0.0)
    assert mesh is not None
    nv = len(mesh)
    print("vertices", nv)
    assert gv.block_count == 1160 and gv.drop_count == 0
    assert nv > 1000
    assert abs(nv - 846288) / 846288 < 0.01
    mean, mx = radius_errors(mesh, 0.5)
    assert mean < 0.01 * 0.25 and mx < 0.01


def test_extract_bit_identical_to_oracle_extractor(H):
    """Same voxel state, same pool order -> byte-identical triangle soup."""
    pts = B.sphere_points(0.4, 200, 400)
    gv = H.HipVolume(0.02, 1 << 14, -1.0, 0.0)
    assert gv.integrate(dev(pts), (0, 0, 0)) == 0
    keys = gv.keys()
    t, w = gv.read_voxels()
    ov = B.Volume(0.02, 1 << 14, -1.0, 0.0)
    for i, k in enumerate(keys):
        assert ov.load_block(int(k), t[i], w[i], 0) == i
    for min_w, iso in ((0.5, 0.0), (2.0, 0.1)):
        g, grgb = gv.extract(min_w, iso, rgb=True)
        o, orgb = ov.extract(min_w, iso, rgb=True)
        assert len(g) == len(o) > 0
        assert np.array_equal(g, o)
        assert np.array_equal(grgb, orgb)


def test_host_extract_in_chunks_equals_device_extract(H):
    """The host-buffer extract copies through two pinned 4 MiB halves on a
    thread pool: the base sphere's 846 k vertices (20.3 MB, five chunks, an
    uneven last one) and their colours must equal the device extraction byte
    for byte, into an unaligned host buffer too."""
    pts = B.sphere_points(0.5, 400, 800)
    gv = H.HipVolume(0.01, 1 << 14, -1.0, 0.0)
    assert gv.integrate(dev(pts), (0, 0, 0)) == 0
    host, hrgb = gv.extract(0.5, 0.0, rgb=True)
    n = len(host)
    assert n * 24 > 4 * (4 << 20)
    posnor = torch.empty(n * 6, dtype=torch.float32, device="cuda")
    col = torch.empty(n * 3, dtype=torch.uint8, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert gv.extract_device(posnor, col, n, 0.5, 0.0, cnt) == 0
    gv.synchronize()
    assert int(cnt.item()) == n
    assert np.array_equal(posnor.cpu().numpy().reshape(n, 6), host)
    assert np.array_equal(col.cpu().numpy().reshape(n, 3), hrgb)
    # a caller's buffer that starts 4 bytes into an allocation
    raw = np.zeros(n * 6 + 1, np.float32)
    got = gv._L.osn_tsdf_hip_extract_mesh(gv.h, raw[1:].ctypes.data, None, n, 0.5, 0.0)
    assert got == n
    assert np.array_equal(raw[1:].reshape(n, 6), host) and raw[0] == 0.0


def test_extract_deterministic_and_overflow(H):
    pts = B.sphere_points(0.4, 200, 400)
    meshes = []
    for _ in range(2):
        gv = H.HipVolume(0.02, 1 << 14, -1.0, 32.0)
        gv.integrate(dev(pts), (0, 0, 0))
        meshes.append(gv.extract(0.5, 0.0))
    # weight-capped sums are order dependent in the last ulp, so compare the
    # geometry as a set; the uncapped case is byte-compared above
    assert len(meshes[0]) == len(meshes[1])
    n = len(meshes[0])
    assert gv.extract(0.5, 0.0, cap=n - 1) is None
    again = gv.extract(0.5, 0.0, cap=n)
    assert again is not None and len(again) == n


def test_plane_gate(H):
    """cpp/tests/test_analytic.cpp:108-141: plane z=0.30 seen from below."""
    pts = B.plane_points(0.30, 0.25, 500)
    gv = H.HipVolume(0.01, 1 << 14, -1.0, 32.0)
    gv.integrate(dev(pts), (0, 0, 0))
    mesh = gv.extract(0.5, 0.0)
    dz = np.abs(mesh[:, 2].astype(np.float64) - 0.30)
    ov = B.Volume(0.01, 1 << 14, -1.0, 32.0)
    ov.integrate(pts, (0, 0, 0))
    assert gv.block_count == ov.block_count
    assert len(mesh) > 1000 and dz.mean() < 0.0025 and dz.max() < 0.01


@pytest.mark.parametrize("name,R,a,b,pool,blocks,verts", [
    ("pts-20k", 0.5, 100, 200, 1 << 14, 1104, 7632),
    ("pts-80k", 0.5, 200, 400, 1 << 14, 1160, 648336),
    ("r-0.25", 0.25, 400, 800, 1 << 14, 280, 212544),
    ("r-1.0", 1.0, 400, 800, 1 << 16, 4752, 2571120),
    ("r-2.0", 2.0, 400, 800, 1 << 17, 18160, 124722),
])
def test_published_sweep_cells(H, name, R, a, b, pool, blocks, verts):
    """Block counts of the reference sweep (results/sweep_*/sweep.csv) exactly;
    vertex counts (docs/geometry/manifest.json) within the reference's 1%
    cross-arm tolerance (cpp/tests/test_cuda_analytic.cu:247-249)."""
    pts = B.sphere_points(R, a, b)
    gv = H.HipVolume(0.01, pool, -1.0, 32.0)
    gv.integrate(dev(pts), (0, 0, 0))
    assert gv.block_count == blocks
    nv = len(gv.extract(0.5, 0.0, cap=4 << 20))
    print(name, gv.block_count, nv)
    assert abs(nv - verts) <= 0.01 * verts


def test_pool_exhaustion_reported(H):
    """cpp/tests/test_analytic.cpp:186-203."""
    pts = B.sphere_points(0.5, 200, 400)
    gv = H.HipVolume(0.005, 16, -1.0, 32.0)
    assert gv.integrate(dev(pts), (0, 0, 0)) == 0
    assert gv.drop_count > 0
    assert gv.block_count == 16
    ov = B.Volume(0.005, 1 << 16, -1.0, 32.0)
    ov.allocate(pts, (0, 0, 0))
    all_keys = np.sort(ov.keys())
    # the 32-slot table itself fills first-come, so only membership is fixed
    assert np.isin(gv.keys(), all_keys).all()
    # with a table large enough for every key, the pool admits exactly the
    # smallest keys of the call, whatever the arrival order
    big = H.HipVolume(0.005, 64, -1.0, 32.0, storage=H.STORE_SUMS, table_slots=1 << 16)
    assert big.allocate(dev(pts), (0, 0, 0)) == 0
    assert big.block_count == 64
    assert big.drop_count == len(all_keys) - 64
    assert np.array_equal(big.keys(), all_keys[:64])


def test_negative_coordinates_and_origin_crossing(H):
    """Points straddling the origin on every axis: floor division must not
    mirror blocks (types.hpp:157-165)."""
    rng = np.random.default_rng(7)
    pts = rng.uniform(-0.3, 0.3, size=(20000, 3)).astype(np.float32)
    cam = (0.013, -0.021, 0.007)
    gv = H.HipVolume(0.01, 1 << 15, -1.0, 0.0)
    gv.integrate(dev(pts.ravel()), cam, 0.0)
    ov = B.Volume(0.01, 1 << 15, -1.0, 0.0)
    ov.integrate(pts.ravel(), cam)
    assert np.array_equal(np.sort(gv.keys()), np.sort(ov.keys()))
    t, w = gv.read_voxels()
    for i, k in enumerate(gv.keys()):
        ot, ow = ov.block_voxels(int(k))[:2]
        assert np.array_equal(w[i], ow)
        assert np.all(np.abs(t[i] - ot) <= 1e-5 * np.maximum(ow, 1.0))


def test_radius_cull_matches(H):
    pts = B.sphere_points(0.5, 100, 200)
    pts[: len(pts) // 2] *= 3.0  # half the points beyond the radius
    gv = H.HipVolume(0.01, 1 << 15, -1.0, 0.0)
    gv.integrate(dev(pts), (0, 0, 0), 1.0)
    ov = B.Volume(0.01, 1 << 15, -1.0, 0.0)
    ov.integrate(pts, (0, 0, 0), 1.0)
    assert np.array_equal(np.sort(gv.keys()), np.sort(ov.keys()))


def test_error_codes_and_variants(H):
    pts = dev(B.sphere_points(0.5, 100, 200))
    gv = H.HipVolume(0.01, 1 << 14, -1.0, 32.0)
    assert gv._L.osn_tsdf_hip_allocate_blocks(None, pts.data_ptr(), 10, 0, 0, 0, 0) == 1
    assert gv._L.osn_tsdf_hip_allocate_blocks(gv.h, None, 10, 0, 0, 0, 0) == 1
    assert gv._L.osn_tsdf_hip_allocate_blocks(gv.h, pts.data_ptr(), 0, 0, 0, 0, 0) == 1
    assert gv._L.osn_tsdf_hip_update_voxels(gv.h, pts.data_ptr(), -1, 0, 0, 0, 0) == 1
    assert gv._L.osn_tsdf_hip_extract_mesh(gv.h, None, None, 10, 0.5, 0.0) == -1
    assert H.lib().osn_tsdf_hip_create(0.0, 16, -1.0, 32.0) is None
    assert H.lib().osn_tsdf_hip_create(0.01, 0, -1.0, 32.0) is None
    # no-CAS probe leaves the table empty
    assert gv.allocate_no_cas(pts, (0, 0, 0)) == 0
    assert gv.block_count == 0
    # counting CAS: attempts land in drop_count, at least one per block
    assert gv.allocate_counting_cas(pts, (0, 0, 0)) == 0
    nb = gv.block_count
    assert nb > 0 and gv.drop_count >= nb
    # probe depth on a fresh volume: every distinct key probed at least once
    pv = H.HipVolume(0.01, 1 << 14, -1.0, 32.0)
    rc, lane, wave, hist = pv.allocate_probe_depth(pts, (0, 0, 0))
    assert rc == 0 and lane >= pv.block_count and wave >= lane and hist.sum() > 0
    # device view describes the same memory
    rc, view = gv.device_view()
    assert rc == 0 and view.pool_capacity == 1 << 14 and view.block_dim == 8
    assert view.hash_mask == (1 << 15) - 1 and view.scratch_base == 1 << 15
    # reset returns the volume to empty
    assert gv.reset() == 0 and gv.block_count == 0 and gv.drop_count == 0
