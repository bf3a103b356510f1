"""Depth-frame integration on the GPU: back-projection, allocation, and the
projective update in FP32, FP16 and FP64 storage, against the oracle.

Pins (the projective update has no reference counterpart):
  * device back-projection == the oracle's FP32 definition, bit for bit;
  * allocated key set per frame == the reference-semantics oracle run on the
    same back-projected points, bit-exact (allocation is defined on points);
  * touched set per frame == the oracle's distinct visited blocks;
  * FP64 device oracle kernel == CPU FP64 oracle, bit for bit;
  * FP64, FP32 and FP16 running averages against their closed form (the mean
    of a voxel's first min(n, cap) observations) on a constant-depth wall;
  * FP32 / FP16 storage against the FP64 oracle: weights exact, max |dtsdf|
    within 1e-4 (FP32) and 0.01 (FP16), in units of the truncation distance
    (FP32 camera depth at 3-6 m carries ~5e-7 m of rounding, ~1.2e-5 trunc
    units per observation);
  * extracted surface within a voxel of the analytic scene;
  * the device renderer against closed-form FP64 ray casting.
"""
import numpy as np
import pytest

import bindings as B
from helpers import render_depth_np, scene_sdf

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def H():
    import osn_tsdf_hip
    assert torch.cuda.is_available()
    return osn_tsdf_hip


@pytest.fixture(scope="module")
def S():
    import scenes
    return scenes


def frames_for(H, S, cfg, seed, n, W=None, Hh=None, noise=True):
    scene, poses = S.build(dict(cfg, frames=n), seed)
    W = W or cfg["width"]
    Hh = Hh or cfg["height"]
    fx, fy, cx, cy = S.intrinsics(W, Hh)
    sc = scene.to_c()
    out = []
    for k, (R, t) in enumerate(poses):
        fr = H.make_frame(W, Hh, fx, fy, cx, cy, R, t)
        d = torch.empty(Hh * W, dtype=torch.float32, device="cuda")
        nk = cfg["noise_k"] if noise else 0.0
        dr = cfg["dropout"] if noise else 0.0
        assert S.render_depth(sc, fr, nk, dr, seed, k, d) == 0
        out.append((fr, d, R, t))
    torch.cuda.synchronize()
    return scene, out


def host_points(H, fr, d):
    W, Hh = fr.width, fr.height
    buf = torch.empty(W * Hh * 3, dtype=torch.float32, device="cuda")
    assert H.backproject(d, fr, buf) == 0
    p = buf.cpu().numpy().reshape(-1, 3)
    return p[~np.isnan(p[:, 0])].ravel()


def test_renderer_matches_closed_form(H, S):
    scene, fr = frames_for(H, S, S.PRECISE, 11, 3, W=160, Hh=120, noise=False)
    fx, fy, cx, cy = S.intrinsics(160, 120)
    for f, d, R, t in fr:
        ref = render_depth_np(scene, R, t, 160, 120, fx, fy, cx, cy)
        got = d.cpu().numpy().reshape(120, 80 * 2)
        rel = np.abs(got - ref) / np.maximum(ref, 1e-3)
        assert np.percentile(rel, 99.9) < 1e-4 and (got > 0).mean() > 0.99


def test_backprojection_bit_exact(H, S):
    _, fr = frames_for(H, S, S.PRECISE, 3, 2)
    for f, d, R, t in fr:
        got = host_points(H, f, d)
        depth = d.cpu().numpy().reshape(f.height, f.width)
        want = B.backproject_f32(depth, f.fx, f.fy, f.cx, f.cy, np.array(f.R_wc[:]),
                                 np.array(f.t_wc[:]))
        assert got.shape == want.shape and np.array_equal(got, want)


# trunc 8 cm = 4 voxels: the walk takes its bisection shortcut on nearly every
# ray. trunc 7 cm = 3.5 voxels: the band still has 4 samples a side, but two
# of them can be culled, so every wave walks sample by sample.
@pytest.mark.parametrize("trunc", [0.08, 0.07])
def test_depth_allocation_key_sets_bit_exact(H, S, trunc):
    """Small config: 20 frames 640x480, 2 cm voxels, table 2^18."""
    cfg = dict(S.SMALL, trunc=trunc)
    _, fr = frames_for(H, S, cfg, 5, cfg["frames"])
    gv = H.HipVolume(cfg["voxel"], cfg["table"] // 2, cfg["trunc"], 0.0, storage=H.STORE_F32)
    ov = B.Volume(cfg["voxel"], cfg["table"] // 2, cfg["trunc"], 0.0)
    for k, (f, d, R, t) in enumerate(fr):
        assert gv.allocate_depth(d, f) == 0
        pts = host_points(H, f, d)
        cam = tuple(f.t_wc[:])
        ov.allocate(pts, cam)
        touched_o = ov.touched_keys(pts, cam)
        assert gv.touched_count == len(touched_o), f"frame {k}"
        if k % 5 == 4 or k == 0:
            assert gv.block_count == ov.block_count
            assert np.array_equal(np.sort(gv.keys()), np.sort(ov.keys())), f"frame {k}"
    assert gv.drop_count == 0


def test_scatter_update_on_small_frames_within_1e5(H, S):
    """The reference-semantics scatter arm on the small workload's 20 frames
    (640x480, 2 cm voxels, table 2^18), back-projected on the device: key set
    bit-exact, weights exact and tsdf sums within 1e-5 of sum|terms| (unit
    weights) against the oracle, at weight_cap 0 (cpp/src/volume.cpp:200-299)."""
    cfg = dict(S.SMALL)
    _, fr = frames_for(H, S, cfg, 5, cfg["frames"])
    pool = cfg["table"] // 2
    gv = H.HipVolume(cfg["voxel"], pool, cfg["trunc"], 0.0)
    ov = B.Volume(cfg["voxel"], pool, cfg["trunc"], 0.0)
    for k, (f, d, R, t) in enumerate(fr):
        W, Hh = f.width, f.height
        buf = torch.empty(W * Hh * 3, dtype=torch.float32, device="cuda")
        assert H.backproject(d, f, buf) == 0
        p3 = buf.view(-1, 3)
        dpts = p3[~torch.isnan(p3[:, 0])].contiguous().view(-1)
        cam = tuple(f.t_wc[:])
        assert gv.integrate(dpts, cam) == 0
        ov.integrate(dpts.cpu().numpy(), cam)
    gv.synchronize()
    assert gv.drop_count == 0 and gv.block_count == ov.block_count
    gk = gv.keys()
    assert np.array_equal(np.sort(gk), np.sort(ov.keys()))
    t, w = gv.read_voxels()
    worst = 0.0
    for i, k in enumerate(gk):
        ot, ow = ov.block_voxels(int(k))[:2]
        assert np.array_equal(w[i], ow), f"weights differ in block {k}"
        e = np.abs(t[i] - ot) / np.maximum(ow, 1.0)
        worst = max(worst, float(e.max()))
    print(f"{len(gk)} blocks, max |dsum| / weight {worst:.2e}")
    assert worst <= 1e-5


def _run(H, storage, cfg, fr, wcap):
    v = H.HipVolume(cfg["voxel"], cfg["table"] // 2, cfg["trunc"], wcap, storage=storage)
    for f, d, R, t in fr:
        assert v.integrate_depth(d, f) == 0
    v.synchronize()
    return v


def _by_key(v):
    keys = v.keys()
    t, w = v.read_voxels()
    return keys, t, w


def _other_depth(d, k):
    """A different image of the same size: depth scaled by 1.004, every 7th
    pixel (offset k) invalid."""
    e = d * 1.004
    e[k % 7::7] = 0.0
    return e


# 157x117: neither side a multiple of the depth tiles (8x4), so the tiled
# copy's padding and its last partial tiles are exercised. separate=True
# allocates on one image and updates from another through update_depth, which
# must tile its own image rather than read the copy allocation left.
@pytest.mark.parametrize("W,Hh,separate", [(160, 120, False), (157, 117, False),
                                           (157, 117, True)])
def test_fp64_kernel_equals_cpu_fp64_oracle(H, S, W, Hh, separate):
    cfg = dict(S.PRECISE, table=1 << 18)
    _, fr = frames_for(H, S, cfg, 9, 4, W=W, Hh=Hh)
    if separate:
        gv = H.HipVolume(cfg["voxel"], cfg["table"] // 2, cfg["trunc"], 32.0, storage=H.STORE_F64)
        # the volume's stream reads the images after the calls return: keep
        # them alive (and written) until it has finished
        others = [_other_depth(d, k) for k, (f, d, R, t) in enumerate(fr)]
        torch.cuda.synchronize()
        for (f, d, R, t), e in zip(fr, others):
            assert gv.allocate_depth(d, f) == 0
            assert gv.update_depth(e, f) == 0
        gv.synchronize()
    else:
        gv = _run(H, H.STORE_F64, cfg, fr, 32.0)
    ov = B.Volume(cfg["voxel"], cfg["table"] // 2, cfg["trunc"], 0.0)
    po = B.ProjOracle(cfg["voxel"], cfg["trunc"], 32, cfg["table"] // 2)
    for k, (f, d, R, t) in enumerate(fr):
        pts = host_points(H, f, d)
        cam = tuple(f.t_wc[:])
        ov.allocate(pts, cam)
        touched = ov.touched_keys(pts, cam)
        Rwc64 = np.array(f.R_wc64[:]).reshape(3, 3)
        Rwc32 = np.array(f.R_wc[:], np.float32).reshape(3, 3)
        du = _other_depth(d, k) if separate else d
        po.integrate(touched, du.cpu().numpy().reshape(f.height, f.width), f.fx, f.fy, f.cx, f.cy,
                     Rwc64.T, np.array(f.t_wc64[:]), Rwc32.T, np.array(f.t_wc[:], np.float32))
    keys = gv.keys()
    t64, w64 = gv.read_voxels_f64()
    n_obs = 0
    for i, k in enumerate(keys):
        ot, ow = po.block(int(k))
        assert np.array_equal(w64[i], ow.astype(np.float64)), f"weights differ in {k}"
        assert np.array_equal(t64[i], ot), f"tsdf differs in {k}"
        n_obs += int((ow > 0).sum())
    assert n_obs > 10000


# Weight cap 32: voxels seen in more than 32 of the 40 frames saturate, and the
# FP16 kernel takes its averaging reciprocals from the table (caps below 64);
# cap 64 runs the FP16 kernel without the table.
@pytest.mark.parametrize("wcap", [32.0, 64.0])
def test_fp32_and_fp16_against_fp64(H, S, wcap):
    cfg = dict(S.PRECISE, table=1 << 20)
    _, fr = frames_for(H, S, cfg, 21, 40)
    ref = _run(H, H.STORE_F64, cfg, fr, wcap)
    k64 = ref.keys()
    t64, w64 = ref.read_voxels_f64()
    if wcap == 32.0:
        assert (w64 == 32.0).sum() > 1000  # the cap is reached
    for storage, tol in ((H.STORE_F32, 1e-4), (H.STORE_F16, 0.01)):
        v = _run(H, storage, cfg, fr, wcap)
        k, t, w = _by_key(v)
        assert np.array_equal(k, k64)  # same sorted-key pool order
        assert np.array_equal(w.astype(np.float64), w64)
        obs = w64 > 0
        err = np.abs(t.astype(np.float64) - t64)[obs]
        print(f"storage {storage}: max {err.max():.3e} rms {np.sqrt((err ** 2).mean()):.3e} "
              f"over {obs.sum()} voxels")
        assert err.max() <= tol


def _wall_expectation(keys, touched_per_frame, depths, voxel, trunc, w_max, W, Hh, f):
    """Closed form of the running average with the cap's freeze, per voxel of
    every block in keys: t = mean of the first min(m, w_max) observations,
    w = min(m, w_max), where frame k observes a voxel of a block it touched
    when the voxel projects inside the image and d_k - z >= -trunc (identity
    pose, constant depth d_k). Returns (t, w, checkable)."""
    n = len(keys)
    t = np.zeros((n, 512))
    w = np.zeros((n, 512), np.int64)
    ok = np.ones((n, 512), bool)
    lz, ly, lx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
    lx, ly, lz = lx.reshape(512), ly.reshape(512), lz.reshape(512)
    vs, tr = float(np.float32(voxel)), float(np.float32(trunc))  # the volume's FP32 values
    for i, k in enumerate(keys):
        b = [((int(k) >> s) & ((1 << 21) - 1)) - (1 << 20) for s in (42, 21, 0)]
        x = ((b[0] * 8 + lx) + 0.5) * vs
        y = ((b[1] * 8 + ly) + 0.5) * vs
        z = ((b[2] * 8 + lz) + 0.5) * vs
        u = f * x / z + W / 2 + 0.5
        v = f * y / z + Hh / 2 + 0.5
        inside = (u >= 0) & (u < W) & (v >= 0) & (v < Hh)
        edge = (np.minimum(np.abs(u), np.abs(u - W)) < 1e-3) | \
               (np.minimum(np.abs(v), np.abs(v - Hh)) < 1e-3)
        ok[i] &= ~edge
        s_ = np.zeros(512)
        c_ = np.zeros(512, np.int64)
        for d, touched in zip(depths, touched_per_frame):
            if int(k) not in touched:
                continue
            sdf = float(d) - z
            ok[i] &= ~(np.abs(sdf + tr) < 1e-5)
            take = inside & (sdf >= -tr) & (c_ < w_max)
            s_[take] += np.clip(sdf / tr, -1.0, 1.0)[take]
            c_[take] += 1
        t[i] = np.where(c_ > 0, s_ / np.maximum(c_, 1), 0.0)
        w[i] = c_
    return t, w, ok


# The running average against its closed form, with no oracle arithmetic: a
# camera at the origin looking along +z sees a wall at a different constant
# depth in each of 12 frames, so every observation x = clamp((d - z) / trunc)
# is known. With cap 5 most voxels in front of the walls saturate and must
# freeze (the reference's cap, types.hpp:99-101); with no cap every
# observation counts. Touched blocks come from the oracle's allocation.
@pytest.mark.parametrize("wcap", [5.0, 0.0])
@pytest.mark.parametrize("storage,tol", [("STORE_F64", 1e-12), ("STORE_F32", 2e-5),
                                         ("STORE_F16", 5e-3)])
def test_running_average_closed_form_on_wall(H, storage, tol, wcap):
    W, Hh, f = 64, 48, 32.0
    voxel, trunc = 0.02, 0.08
    depths = np.random.default_rng(7).uniform(0.95, 1.35, 12).astype(np.float32)
    fr = H.make_frame(W, Hh, f, f, W / 2, Hh / 2, np.eye(3), np.zeros(3))
    v = H.HipVolume(voxel, 1 << 14, trunc, wcap, storage=getattr(H, storage))
    ov = B.Volume(voxel, 1 << 14, trunc, 0.0)
    touched = []
    imgs = [torch.full((Hh * W,), float(d), dtype=torch.float32, device="cuda") for d in depths]
    for d, img in zip(depths, imgs):
        assert v.integrate_depth(img, fr) == 0
        pts = B.backproject_f32(np.full((Hh, W), d, np.float32), f, f, W / 2, Hh / 2,
                                np.eye(3), np.zeros(3))
        ov.allocate(pts, (0.0, 0.0, 0.0))
        touched.append(set(int(k) for k in ov.touched_keys(pts, (0.0, 0.0, 0.0))))
    v.synchronize()
    keys = v.keys()
    assert np.array_equal(np.sort(keys), np.sort(ov.keys()))
    if storage == "STORE_F64":
        t, w = v.read_voxels_f64()
    else:
        t, w = v.read_voxels()
    w_max = 5 if wcap > 0 else 65535
    te, we, ok = _wall_expectation(keys, touched, depths, voxel, trunc, w_max, W, Hh, f)
    assert ok.sum() > 20000
    assert np.array_equal(w.astype(np.int64)[ok], we[ok])
    err = np.abs(t.astype(np.float64) - te)[ok]
    print(f"{storage} cap {wcap}: max |t - closed form| {err.max():.3e} over {ok.sum()} voxels, "
          f"{int((we[ok] == w_max).sum())} saturated")
    assert err.max() <= tol
    if wcap > 0:
        assert (we[ok] == w_max).sum() > 5000


def test_projective_mesh_on_analytic_surface_and_deterministic(H, S):
    cfg = dict(S.SMALL)
    scene, fr = frames_for(H, S, cfg, 2, 20, noise=False)
    meshes = []
    for _ in range(2):
        v = _run(H, H.STORE_F16, cfg, fr, 32.0)
        meshes.append(v.extract(1.0, 0.0))
    assert np.array_equal(meshes[0], meshes[1])
    m = meshes[0]
    assert len(m) > 10000
    dist = scene_sdf(scene, m[:, :3])
    print("surface distance mean", dist.mean(), "p99", np.percentile(dist, 99))
    assert dist.mean() < cfg["voxel"] / 4
    assert np.percentile(dist, 99) < cfg["voxel"]


def test_device_extract_matches_host_extract(H, S):
    cfg = dict(S.SMALL)
    _, fr = frames_for(H, S, cfg, 4, 6, noise=False)
    v = _run(H, H.STORE_F32, cfg, fr, 32.0)
    host = v.extract(1.0, 0.0)
    n = len(host)
    posnor = torch.empty(n * 6 + 60, dtype=torch.float32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert v.extract_device(posnor, None, n + 10, 1.0, 0.0, cnt) == 0
    v.synchronize()
    assert int(cnt.item()) == n
    assert np.array_equal(posnor[: n * 6].cpu().numpy().reshape(n, 6), host)
    assert v.extract_device(posnor, None, n - 1, 1.0, 0.0, cnt) == 0
    v.synchronize()
    assert int(cnt.item()) == -1
    # the host-buffer extract of the same state, loaded into the oracle extractor
    keys = v.keys()
    t, w = v.read_voxels()
    ov = B.Volume(cfg["voxel"], len(keys) + 8, cfg["trunc"], 0.0)
    for i, k in enumerate(keys):
        ov.load_block(int(k), t[i], w[i], 1)
    assert np.array_equal(ov.extract(1.0, 0.0), host)


@pytest.mark.parametrize("storage", ["STORE_F16", "STORE_F32"])
def test_dense_blocks_extract_bit_identical(H, storage):
    """Independent random depths per pixel, about one pixel per voxel: nearly
    every cube of the band crosses, so blocks hold more triangles than one
    listing pass takes (768). Device extraction into host buffers, and into a
    device buffer one float past an 8-byte boundary, against the oracle
    extractor over the same voxel state (marching_tets.cpp:84-243)."""
    W, Hh = 128, 96
    fr = H.make_frame(W, Hh, 128.0, 128.0, 64.0, 48.0, np.eye(3), np.zeros(3))
    rng = np.random.default_rng(17)
    depth = torch.from_numpy(rng.uniform(0.95, 1.05, W * Hh).astype(np.float32)).cuda()
    v = H.HipVolume(0.01, 1 << 13, 0.04, 32.0, storage=getattr(H, storage))
    assert v.integrate_depth(depth, fr) == 0
    host = v.extract(0.5, 0.0, cap=4 << 20)
    assert host is not None
    n = len(host)
    # the densest blocks exceed one listing pass
    cell = np.floor(host[:, :3].astype(np.float64) / 0.01 - 0.5).astype(np.int64)
    _, per_block = np.unique(np.floor_divide(cell, 8), axis=0, return_counts=True)
    print("vertices", n, "blocks", v.block_count, "max triangles per block", per_block.max() // 3)
    assert per_block.max() > 3 * 768
    keys = v.keys()
    t, w = v.read_voxels()
    ov = B.Volume(0.01, len(keys) + 8, 0.04, 0.0)
    for i, k in enumerate(keys):
        ov.load_block(int(k), t[i], w[i], 1)
    assert np.array_equal(ov.extract(0.5, 0.0, cap=4 << 20), host)
    buf = torch.empty(n * 6 + 8, dtype=torch.float32, device="cuda")
    cnt = torch.zeros(1, dtype=torch.int32, device="cuda")
    assert v.extract_device(buf[1:], None, n, 0.5, 0.0, cnt) == 0
    v.synchronize()
    assert int(cnt.item()) == n
    assert np.array_equal(buf[1:1 + n * 6].cpu().numpy().reshape(n, 6), host)


def test_extract_after_every_frame_and_reset(H, S):
    """Extraction after every frame reuses the neighbour indices the earlier
    calls found; blocks allocated later next to old ones and a reset must not
    leave stale indices. The last mesh of each round against the oracle
    extractor over the same voxel state."""
    cfg = dict(S.SMALL)
    _, fr = frames_for(H, S, cfg, 5, 6)
    v = H.HipVolume(cfg["voxel"], 1 << 14, cfg["trunc"], 32.0, storage=H.STORE_F16)
    for _ in range(2):
        for f, d, _, _ in fr:
            assert v.integrate_depth(d, f) == 0
            assert v.extract(0.5, 0.0, cap=4 << 20) is not None
        host = v.extract(0.5, 0.0, cap=4 << 20)
        keys = v.keys()
        t, w = v.read_voxels()
        ov = B.Volume(cfg["voxel"], len(keys) + 8, cfg["trunc"], 0.0)
        for i, k in enumerate(keys):
            ov.load_block(int(k), t[i], w[i], 1)
        assert len(host) > 0 and np.array_equal(ov.extract(0.5, 0.0, cap=4 << 20), host)
        assert v.reset() == 0


@pytest.mark.parametrize("storage", ["f16", "f32"])
def test_extraction_sequence_matches_oracle(H, S, storage):
    """Extractions of one volume reuse the neighbour indices earlier calls
    found (kept in the volume's scratch). Meshes through a sequence where
    anything kept from an earlier call could go stale -- an update after an
    extraction, a repeated update_depth of one allocation, blocks allocated
    next to old ones, other min_weight and iso and back with no update in
    between, a reset -- equal the oracle extractor over the same voxel
    state."""
    st = H.STORE_F16 if storage == "f16" else H.STORE_F32
    cfg = dict(S.SMALL)
    _, fr = frames_for(H, S, cfg, 7, 5)
    v = H.HipVolume(cfg["voxel"], 1 << 14, cfg["trunc"], 32.0, storage=st)

    def check(min_w, iso, what):
        host = v.extract(min_w, iso, cap=4 << 20)
        keys = v.keys()
        t, w = v.read_voxels()
        ov = B.Volume(cfg["voxel"], len(keys) + 8, cfg["trunc"], 0.0)
        for i, k in enumerate(keys):
            ov.load_block(int(k), t[i], w[i], 1)
        ref = ov.extract(min_w, iso, cap=4 << 20)
        assert host is not None and len(host) == len(ref) > 0, what
        assert np.array_equal(ref, host), what

    for rnd in range(2):
        f, d, _, _ = fr[0]
        assert v.integrate_depth(d, f) == 0
        check(0.5, 0.0, "first frame")
        check(0.5, 0.0, "no change")
        for i, (f, d, _, _) in enumerate(fr[1:3]):
            assert v.integrate_depth(d, f) == 0
            check(0.5, 0.0, f"frame {i + 1}")
        # one allocation, two updates, a count between them
        f, d, _, _ = fr[3]
        assert v.allocate_depth(d, f) == 0 and v.update_depth(d, f) == 0
        check(0.5, 0.0, "allocate + update")
        assert v.update_depth(d, f) == 0
        check(0.5, 0.0, "the same frame's update again")
        # other parameters (no update between), then back
        check(0.5, 0.1, "iso 0.1")
        check(2.5, 0.1, "min weight 2.5")
        check(0.5, 0.0, "back to the first parameters")
        f, d, _, _ = fr[4]
        assert v.integrate_depth(d, f) == 0
        check(0.5, 0.0, f"round {rnd} last frame")
        assert v.reset() == 0


def test_storage_errors(H, S):
    v = H.HipVolume(0.02, 1024, -1.0, 32.0, storage=H.STORE_F16)
    pts = torch.zeros(30, dtype=torch.float32, device="cuda")
    assert v.update(pts, (0, 0, 0)) == 3          # scatter update needs sum storage
    assert v.device_view()[0] == 3
    r = H.HipVolume(0.02, 1024, -1.0, 32.0)
    fr = H.make_frame(8, 8, 4, 4, 4, 4, np.eye(3), np.zeros(3))
    d = torch.ones(64, dtype=torch.float32, device="cuda")
    assert r.update_depth(d, fr) == 3             # projective update needs mean storage
    assert r.allocate_depth(None, fr) == 1


def test_degenerate_and_ragged_frames_fp64_bit_exact(H, S):
    """Edge cases of the depth path, bit-exact against the CPU FP64 oracle on
    one volume, frame after frame: a ragged 31x5 image (neither side a
    multiple of the 8x4 depth tiles or of the 16x16 walk tiles), the same
    image with every pixel invalid (0, negative, NaN, +inf: nothing is
    allocated or updated), with a single valid pixel, a 1x1 image, a 4099x1
    row (one pixel tall, wider than any tile row), and valid pixels mixed
    with NaN, +inf and negative ones."""
    cfg = dict(S.PRECISE, table=1 << 16)
    _, base = frames_for(H, S, cfg, 31, 3, W=31, Hh=5)
    seq = []
    f0, d0, _, _ = base[0]
    seq.append((f0, d0))
    bad = d0.clone()
    bad[0::4] = 0.0
    bad[1::4] = -1.0
    bad[2::4] = float("nan")
    bad[3::4] = float("inf")
    seq.append((base[1][0], bad))
    one = torch.zeros_like(d0)
    one[77] = base[2][1][77]
    seq.append((base[2][0], one))
    for W, Hh in ((1, 1), (4099, 1)):
        _, fr = frames_for(H, S, cfg, 33, 1, W=W, Hh=Hh, noise=False)
        seq.append((fr[0][0], fr[0][1]))
    # valid pixels with invalid ones among them: voxels of the blocks the
    # valid ones touch project onto NaN (skipped), +inf (observed as free
    # space, x = 1) and negative (skipped) depths
    mixed = base[2][1].clone()
    mixed[0::5] = float("nan")
    mixed[1::7] = float("inf")
    mixed[2::11] = -2.0
    seq.append((base[2][0], mixed))
    gv = H.HipVolume(cfg["voxel"], cfg["table"] // 2, cfg["trunc"], 32.0, storage=H.STORE_F64,
                     table_slots=cfg["table"])
    for f, d in seq:
        assert gv.integrate_depth(d, f) == 0
    gv.synchronize()
    ov = B.Volume(cfg["voxel"], cfg["table"] // 2, cfg["trunc"], 0.0)
    po = B.ProjOracle(cfg["voxel"], cfg["trunc"], 32, cfg["table"] // 2)
    counts = []
    for f, d in seq:
        pts = host_points(H, f, d)
        cam = tuple(f.t_wc[:])
        before = ov.block_count
        if pts.size:
            ov.allocate(pts, cam)
            touched = ov.touched_keys(pts, cam)
        else:
            touched = np.zeros(0, np.int64)
        counts.append((pts.size // 3, ov.block_count - before, touched.size))
        Rwc64 = np.array(f.R_wc64[:]).reshape(3, 3)
        Rwc32 = np.array(f.R_wc[:], np.float32).reshape(3, 3)
        po.integrate(touched, d.cpu().numpy().reshape(f.height, f.width), f.fx, f.fy, f.cx, f.cy,
                     Rwc64.T, np.array(f.t_wc64[:]), Rwc32.T, np.array(f.t_wc[:], np.float32))
    print("points / new blocks / touched per frame:", counts)
    assert counts[1] == (0, 0, 0)  # every pixel invalid
    assert counts[2][0] == 1 and counts[2][2] >= 1  # one pixel: its band's blocks
    assert counts[3][0] == 1 and counts[4][0] > 1000 and counts[5][2] > 0
    keys = gv.keys()
    assert np.array_equal(np.sort(keys), np.sort(ov.keys()))
    assert gv.drop_count == 0
    t64, w64 = gv.read_voxels_f64()
    for i, k in enumerate(keys):
        ot, ow = po.block(int(k))
        assert np.array_equal(w64[i], ow.astype(np.float64)), f"weights differ in {k}"
        assert np.array_equal(t64[i], ot), f"tsdf differs in {k}"
