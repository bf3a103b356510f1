"""Large-volume configuration: a seeded 12 x 12 x 4 m room with 60 mixed
primitives, 1280x960 depth, voxel 5 mm, truncation 2 cm, table 2^25 slots,
through the batched FP16 pipeline, checked against the FP64 oracle kernel and
the CPU oracle on the first frames.

With the configuration's noise model (sigma = 0.0012 z^2, up to four voxels
at the room's far walls) the meshes of even a few frames exceed 67M vertices,
so the mesh gate runs on a noise-free frame of the same scene and trajectory,
cut to its central 320x240 pixels (the whole 90 degree view of the room is tens
of square metres of 5 mm surface: ~70M vertices of triangle soup).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from test_gpu_projective import frames_for, host_points  # noqa: E402

POOL = 1 << 20  # blocks; the frames used here allocate a few hundred thousand


@pytest.fixture(scope="module")
def H():
    import osn_tsdf_hip
    return osn_tsdf_hip


@pytest.fixture(scope="module")
def S():
    import scenes
    return scenes


def _volume(H, cfg, storage):
    return H.HipVolume(cfg["voxel"], POOL, cfg["trunc"], 32.0, storage=storage,
                       table_slots=cfg["table"], hash_kind=1)


def _batched(H, cfg, fr, extract=None):
    v = _volume(H, cfg, H.STORE_F16)
    b = H.Batch([v])
    assert b.set_schedule([[d] for f, d, R, t in fr], [[f] for f, d, R, t in fr]) == 0
    if extract is not None:
        assert b.set_extract([extract[0]], [extract[1]], extract[2], 0.5, 0.0,
                             every=len(fr)) == 0
    assert b.run(len(fr), use_graph=True) == 0
    b.synchronize()
    b.close()
    return v


def test_large_key_set_and_fp16_error(H, S):
    import bindings as B
    cfg = S.LARGE
    _, fr = frames_for(H, S, cfg, 21, 4)
    g16 = _batched(H, cfg, fr)
    g64 = _volume(H, cfg, H.STORE_F64)
    ov = B.Volume(cfg["voxel"], POOL, cfg["trunc"], 0.0)
    for f, d, R, t in fr:
        assert g64.integrate_depth(d, f) == 0
        ov.allocate(host_points(H, f, d), tuple(f.t_wc[:]))
    g64.synchronize()
    assert g16.drop_count == 0 and g64.drop_count == 0
    keys = g16.keys()
    assert np.array_equal(keys, g64.keys())
    assert np.array_equal(np.sort(keys), np.sort(ov.keys()))
    t16, w16 = g16.read_voxels()
    t64, w64 = g64.read_voxels_f64()
    obs = w64 > 0
    assert obs.sum() > 10_000_000
    assert np.array_equal(w16.astype(np.float64)[obs], w64[obs])
    err = np.abs(t16.astype(np.float64)[obs] - t64[obs])
    assert err.max() <= 0.01  # tsdf is stored in truncation units


def test_large_mesh_gate_noise_free(H, S):
    import mesh_metrics as M
    cfg = S.LARGE
    _, fr = frames_for(H, S, cfg, 21, 1, noise=False)
    # keep a central 320x240 window of the 1280x960 frame (depth 0 = invalid
    # elsewhere): the whole view is ~70M vertices of triangle soup
    W, Hh = cfg["width"], cfg["height"]
    img = fr[0][1].view(Hh, W)
    keep = torch.zeros_like(img, dtype=torch.bool)
    keep[Hh // 2 - 120:Hh // 2 + 120, W // 2 - 160:W // 2 + 160] = True
    img[~keep] = 0.0
    cap = 1 << 25
    buf = torch.empty(cap * 6, dtype=torch.float32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    g16 = _batched(H, cfg, fr, extract=(buf, cap, cnt))
    n16 = int(cnt.item())
    assert n16 > 0, "batched extraction overflowed"
    mesh16 = buf[: n16 * 6].view(-1, 6).cpu().numpy()
    del buf
    g64 = _volume(H, cfg, H.STORE_F64)
    for f, d, R, t in fr:
        assert g64.integrate_depth(d, f) == 0
    mesh64 = g64.extract(0.5, 0.0, cap=cap)
    assert mesh64 is not None and len(mesh64) > 0
    assert g16.block_count == g64.block_count
    tol = M.Tolerance.for_voxel(cfg["voxel"])
    g = M.gate(mesh16, mesh64, tol, 200000, cfg["voxel"])
    assert g.passed, (g.failures, g.agreement, g.vertex_count_rel)
