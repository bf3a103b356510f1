"""Long trajectories on the GPU.

* Precise workload, the whole 500-frame trajectory (640x480, 1 cm voxels,
  table 2^22, noise and dropouts, weight cap 32): FP16 and FP32 storage
  against the FP64 oracle kernel after most observed voxels have saturated
  at the cap and frozen: weights exact, max |dtsdf| within 0.01 (FP16) and
  1e-4 (FP32) truncation units, key sets and pool order identical.
* Large workload, 100 noisy 1280x960 frames (5 mm voxels, trunc 2 cm, table
  2^25; OSN_LARGE_FRAMES=1000 runs the whole trajectory, pool 2^22), through
  the batched FP16 pipeline with the mesh extracted inside the
  captured step into a device buffer of 2^30 vertices (the soup is about 316 M
  vertices: the noise at the far walls is several voxels), against
  the FP64 oracle kernel over the same frames: every observed voxel within
  0.01 truncation units with exact weights, the soups' vertex counts within
  the mesh gate's tolerance, and the mesh gate on four seeded 0.8 m crops of
  the two soups (the host gate cannot hold soups of this size).
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

from test_gpu_projective import frames_for  # noqa: E402


@pytest.fixture(scope="module")
def H():
    import osn_tsdf_hip
    assert torch.cuda.is_available()
    return osn_tsdf_hip


@pytest.fixture(scope="module")
def S():
    import scenes
    return scenes


def compare_chunked(va, vb, n_blocks, f64_b, chunk=1 << 16):
    """Max |t_a - t_b| over voxels observed in b, and whether every weight
    agrees, block range by block range (va FP16/FP32, vb FP64 or FP32)."""
    err = 0.0
    observed = 0
    saturated = 0
    for first in range(0, n_blocks, chunk):
        n = min(chunk, n_blocks - first)
        ta, wa = va.read_voxels(first, n)
        tb, wb = vb.read_voxels_f64(first, n) if f64_b else vb.read_voxels(first, n)
        obs = wb > 0
        if not np.array_equal(wa.astype(np.float64)[obs], wb.astype(np.float64)[obs]):
            return None
        if obs.any():
            err = max(err, float(np.abs(ta.astype(np.float64) - tb)[obs].max()))
        observed += int(obs.sum())
        saturated += int((wb == 32).sum())
        print(f"  compared blocks {first}..{first + n} of {n_blocks}", flush=True)
    return err, observed, saturated


def test_precise_full_trajectory(H, S):
    cfg = dict(S.PRECISE)
    _, fr = frames_for(H, S, cfg, 1, cfg["frames"])
    vols = {}
    for name, st in (("f64", H.STORE_F64), ("f32", H.STORE_F32), ("f16", H.STORE_F16)):
        v = H.HipVolume(cfg["voxel"], 1 << 20, cfg["trunc"], 32.0, storage=st,
                        table_slots=cfg["table"])
        for f, d, R, t in fr:
            assert v.integrate_depth(d, f) == 0
        v.synchronize()
        assert v.drop_count == 0
        vols[name] = v
    keys = vols["f64"].keys()
    n = len(keys)
    for name, tol in (("f32", 1e-4), ("f16", 0.01)):
        assert np.array_equal(vols[name].keys(), keys)
        r = compare_chunked(vols[name], vols["f64"], n, True)
        assert r is not None, f"{name} weights differ"
        err, observed, saturated = r
        print(f"{name}: {n} blocks, {observed} observed voxels, {saturated} frozen at the cap, "
              f"max |dt| {err:.3e} trunc")
        assert observed > 10_000_000 and saturated > observed // 2
        assert err <= tol


CAP = 1 << 30  # vertices per soup buffer (24 GiB of position + normal records)


def crops(m64, n_crops=4, half=0.4, seed=5):
    """Crop boxes centred on seeded triangles of the FP64 soup (device)."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    tri = m64.view(-1, 3, 6)
    idx = torch.randint(0, tri.shape[0], (n_crops,), generator=g)
    return [tri[int(i), :, :3].mean(0) for i in idx], half


def crop(m, centre, half):
    tri = m.view(-1, 3, 6)
    cen = tri[:, :, :3].mean(1)
    keep = ((cen - centre).abs() < half).all(1)
    return tri[keep].reshape(-1, 6).cpu().numpy()


def test_large_100_noisy_frames_batched_with_soup(H, S):
    import mesh_metrics as M
    cfg = dict(S.LARGE)
    n_frames = int(os.environ.get("OSN_LARGE_FRAMES", "100"))
    t0 = time.time()
    _, fr = frames_for(H, S, cfg, 21, n_frames)
    print(f"{n_frames} frames rendered in {time.time() - t0:.1f} s", flush=True)
    pool = 1 << 21 if n_frames <= 100 else 1 << 22
    # The whole trajectory observes every wall from afar, whose noise (sigma
    # 0.0012 z^2, several voxels) leaves more than 2^31 vertices of surface
    # at min weight 0.5: there the soups keep voxels seen 4 times or more.
    cap = CAP if n_frames <= 100 else (1 << 31) - 64
    min_w = 0.5 if n_frames <= 100 else 4.0
    v16 = H.HipVolume(cfg["voxel"], pool, cfg["trunc"], 32.0, storage=H.STORE_F16,
                      table_slots=cfg["table"], hash_kind=1)
    buf16 = torch.empty(cap * 6, dtype=torch.float32, device="cuda")
    counts = torch.zeros(1, dtype=torch.int32, device="cuda")
    b = H.Batch([v16])
    assert b.set_schedule([[d] for f, d, R, t in fr], [[f] for f, d, R, t in fr]) == 0
    assert b.set_extract([buf16], [cap], counts, min_w, 0.0, every=n_frames) == 0
    t0 = time.time()
    assert b.run(n_frames, use_graph=True) == 0
    b.synchronize()
    b.close()
    n16 = int(counts.item())
    print(f"batched FP16 with the soup extracted: {time.time() - t0:.1f} s; {v16.block_count} "
          f"blocks, {v16.drop_count} drops, soup {n16}", flush=True)
    assert n16 > 0, "the soup overflowed its buffer"
    v64 = H.HipVolume(cfg["voxel"], pool, cfg["trunc"], 32.0, storage=H.STORE_F64,
                      table_slots=cfg["table"], hash_kind=1)
    t0 = time.time()
    for f, d, R, t in fr:
        assert v64.integrate_depth(d, f) == 0
    v64.synchronize()
    print(f"FP64 oracle kernel: {time.time() - t0:.1f} s", flush=True)
    buf64 = torch.empty(cap * 6, dtype=torch.float32, device="cuda")
    c64 = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert v64.extract_device(buf64, None, cap, min_w, 0.0, c64) == 0
    v64.synchronize()
    n64 = int(c64.item())
    assert n64 > 0
    assert v16.drop_count == 0 and v64.drop_count == 0
    keys = v16.keys()
    assert np.array_equal(keys, v64.keys())
    r = compare_chunked(v16, v64, len(keys), True)
    assert r is not None, "weights differ"
    err, observed, _ = r
    tol = M.Tolerance.for_voxel(cfg["voxel"])
    rel = abs(n16 - n64) / n64
    print(f"{len(keys)} blocks, {observed} observed voxels, max |dt| {err:.3e} trunc; "
          f"soups {n16} / {n64} vertices (rel {rel:.2e})")
    assert err <= 0.01
    assert rel <= tol.vertex_count_rel
    m16 = buf16[: n16 * 6]
    m64 = buf64[: n64 * 6]
    centres, half = crops(m64)
    for c in centres:
        a, bb = crop(m16, c, half), crop(m64, c, half)
        g = M.gate(a, bb, tol, 200000, cfg["voxel"])
        print(f"crop at {[round(float(x), 2) for x in c]}: {len(a)} / {len(bb)} vertices, "
              f"mean {g.agreement.mean_surface / cfg['voxel']:.2e} vox, p99 "
              f"{g.agreement.hausdorff_p99 / cfg['voxel']:.2e} vox")
        assert len(bb) > 100000
        assert g.passed, (g.failures, g.vertex_count_rel)
