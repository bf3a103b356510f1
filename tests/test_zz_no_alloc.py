"""The extraction entry point that allocates nothing
(osn_tsdf_hip_extract_mesh_device_noalloc) against the allocating one
(osn_tsdf_hip_extract_mesh_device), through the C ABI.

Its output must be byte-identical (vertex count, every float, every colour
byte) on two calls in a row, the second over a smaller volume, so whatever
the first call left in the shared scratch cannot leak into the second; an
allocating extraction of the first volume afterwards must still give its
first result (the shared scratch must not touch the volume's own neighbour
cache); an output buffer one vertex too small gives -1 on both; a volume
whose pool exceeds the scratch is refused with 4 before anything runs.
"""
import numpy as np
import pytest

import bindings as B

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def H():
    import osn_tsdf_hip
    assert torch.cuda.is_available()
    return osn_tsdf_hip


def precise_volume(H, n_frames):
    """FP16 storage, 640x480 noisy frames of the precise scene, table 2^20."""
    import scenes as S
    cfg = dict(S.PRECISE, frames=n_frames)
    scene, poses = S.build(cfg, 3)
    fx, fy, cx, cy = S.intrinsics(640, 480)
    sc = scene.to_c()
    v = H.HipVolume(cfg["voxel"], 1 << 19, cfg["trunc"], 32.0, storage=H.STORE_F16)
    for k, (R, t) in enumerate(poses):
        fr = H.make_frame(640, 480, fx, fy, cx, cy, R, t)
        d = torch.empty(640 * 480, dtype=torch.float32, device="cuda")
        assert S.render_depth(sc, fr, cfg["noise_k"], cfg["dropout"], 3, k, d) == 0
        assert v.integrate_depth(d, fr) == 0
    v.synchronize()
    return v


def sphere_volume(H):
    """The reference-shaped volume (weighted sums, colour planes) of a sphere."""
    v = H.HipVolume(0.01, 1 << 14, -1.0, 32.0)
    pts = torch.from_numpy(B.sphere_points(0.5, 200, 400)).cuda()
    assert v.integrate(pts, (0.0, 0.0, 0.0)) == 0
    v.synchronize()
    return v


def extract(v, cap, noalloc, rgb):
    posnor = torch.full((cap * 6 + 6,), -7.0, dtype=torch.float32, device="cuda")
    col = torch.zeros(cap * 3 + 3, dtype=torch.uint8, device="cuda") if rgb else None
    cnt = torch.full((1,), -9, dtype=torch.int32, device="cuda")
    f = v.extract_device_noalloc if noalloc else v.extract_device
    assert f(posnor, col, cap, 0.5, 0.0, cnt) == 0
    v.synchronize()
    n = int(cnt.item())
    out = posnor.cpu().numpy()
    return n, out, (col.cpu().numpy() if rgb else None)


def same(a, b):
    return a[0] == b[0] and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)) and \
        (a[2] is None or np.array_equal(a[2], b[2]))


def test_noalloc_equals_allocating_twice_then_smaller(H):
    big = precise_volume(H, 8)
    cap = 16 << 20
    ref1 = extract(big, cap, False, False)
    got1 = extract(big, cap, True, False)
    assert ref1[0] > 100000 and same(got1, ref1)
    # the second call: a smaller volume, another storage, colour output
    small = sphere_volume(H)
    assert small.block_count < big.block_count
    cap2 = 4 << 20
    got2 = extract(small, cap2, True, True)
    ref2 = extract(small, cap2, False, True)
    assert ref2[0] > 1000 and same(got2, ref2)
    # the volume's own neighbour cache is untouched by the shared scratch
    assert same(extract(big, cap, False, False), ref1)
    # and the no-allocation call after it still agrees
    assert same(extract(big, cap, True, False), ref1)
    print("vertices", ref1[0], ref2[0], "blocks", big.block_count, small.block_count)


def test_noalloc_overflow_and_refusal(H):
    v = sphere_volume(H)
    n = extract(v, 4 << 20, False, False)[0]
    over = extract(v, n - 1, True, False)
    assert over[0] == -1 and np.all(over[1] == -7.0)  # nothing written
    assert extract(v, n - 1, False, False)[0] == -1
    big_pool = H.HipVolume(0.01, (1 << 21) + 64, 0.04, 32.0, storage=H.STORE_F16,
                           table_slots=1 << 23)
    buf = torch.zeros(60, dtype=torch.float32, device="cuda")
    cnt = torch.full((1,), -9, dtype=torch.int32, device="cuda")
    assert big_pool.extract_device_noalloc(buf, None, 10, 0.5, 0.0, cnt) == 4
    torch.cuda.synchronize()
    assert int(cnt.item()) == -9
    assert big_pool.extract_device_noalloc(None, None, 10, 0.5, 0.0, cnt) == 1
