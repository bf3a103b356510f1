"""Batched multi-scene integration against per-scene integration.

Each scene's arithmetic is the same as through the single-volume entry points
and no two scenes share memory, so the batched state must be byte-identical
to integrating every scene on its own, whatever the launch shape, and a
captured-graph replay must match eager launches after every step.
"""
import numpy as np
import pytest

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


SIZES = [(320, 240), (640, 480), (160, 120)]


def make_inputs(H, S, n_steps, seed0=40):
    cfg = S.PRECISE
    depths, frames = [], []
    for k in range(n_steps):
        depths.append([None] * len(SIZES))
        frames.append([None] * len(SIZES))
    for s, (W, Hh) in enumerate(SIZES):
        scene, poses = S.build(dict(cfg, frames=n_steps), seed0 + s)
        fx, fy, cx, cy = S.intrinsics(W, Hh)
        sc = scene.to_c()
        for k, (R, t) in enumerate(poses):
            if s == 2 and k % 3 == 1:
                continue  # scene 2 idles every third step
            fr = H.make_frame(W, Hh, fx, fy, cx, cy, R, t)
            d = torch.empty(W * Hh, dtype=torch.float32, device="cuda")
            assert S.render_depth(sc, fr, cfg["noise_k"], cfg["dropout"], seed0 + s, k, d) == 0
            depths[k][s] = d
            frames[k][s] = fr
    torch.cuda.synchronize()
    return depths, frames


def volumes(H, storage, tables=(1 << 18, 1 << 20, 1 << 16)):
    return [H.HipVolume(0.01, t // 2, 0.04, 32.0, storage=storage) for t in tables]


def state(v):
    k = v.keys()
    t, w = v.read_voxels()
    return k, t.view(np.uint32), w.view(np.uint32)


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


@pytest.mark.parametrize("storage", [2, 1])
def test_batched_equals_per_scene(H, S, storage):
    n_steps = 9
    depths, frames = make_inputs(H, S, n_steps)
    solo = volumes(H, storage)
    for k in range(n_steps):
        for s, v in enumerate(solo):
            if depths[k][s] is not None:
                assert v.integrate_depth(depths[k][s], frames[k][s]) == 0
    batched = volumes(H, storage)
    b = H.Batch(batched)
    assert b.set_schedule(depths, frames) == 0
    assert b.run(4) == 0
    assert b.run(5) == 0
    assert b.run(1) == 1  # past the schedule
    b.synchronize()
    for v1, v2 in zip(solo, batched):
        assert v1.block_count == v2.block_count > 0
        assert same(state(v1), state(v2))
    b.close()
    # after the batch is gone the volumes continue through the single calls
    for s, v in enumerate(batched):
        if depths[0][s] is not None:
            assert v.integrate_depth(depths[0][s], frames[0][s]) == 0
            assert solo[s].integrate_depth(depths[0][s], frames[0][s]) == 0
    for v1, v2 in zip(solo, batched):
        assert same(state(v1), state(v2))


# Twelve small scenes: scene 1 mod 4 idles on odd steps and scene 5 never has
# a frame, so scenes with no touched blocks sit in the middle of the batch's
# work list as well as at its end.
def test_batched_many_scenes_with_idle_ones(H, S):
    n_steps, W, Hh, n_sc = 4, 160, 120, 12
    cfg = S.PRECISE
    fx, fy, cx, cy = S.intrinsics(W, Hh)
    depths = [[None] * n_sc for _ in range(n_steps)]
    frames = [[None] * n_sc for _ in range(n_steps)]
    for s in range(n_sc):
        scene, poses = S.build(dict(cfg, frames=n_steps), 70 + s)
        sc = scene.to_c()
        for k, (R, t) in enumerate(poses):
            if s == 5 or (s % 4 == 1 and k % 2 == 1):
                continue
            fr = H.make_frame(W, Hh, fx, fy, cx, cy, R, t)
            d = torch.empty(W * Hh, dtype=torch.float32, device="cuda")
            assert S.render_depth(sc, fr, cfg["noise_k"], cfg["dropout"], 70 + s, k, d) == 0
            depths[k][s] = d
            frames[k][s] = fr
    torch.cuda.synchronize()
    solo = volumes(H, 2, tables=[1 << 16] * n_sc)
    for k in range(n_steps):
        for s, v in enumerate(solo):
            if depths[k][s] is not None:
                assert v.integrate_depth(depths[k][s], frames[k][s]) == 0
    batched = volumes(H, 2, tables=[1 << 16] * n_sc)
    b = H.Batch(batched)
    assert b.set_schedule(depths, frames) == 0
    assert b.run(n_steps, use_graph=True) == 0
    b.synchronize()
    for s, (v1, v2) in enumerate(zip(solo, batched)):
        assert v1.block_count == v2.block_count
        assert (v1.block_count > 0) == (s != 5)
        assert same(state(v1), state(v2)), f"scene {s}"
    b.close()


def test_graph_replay_matches_eager_every_step(H, S):
    n_steps = 6
    depths, frames = make_inputs(H, S, n_steps, seed0=70)
    eager_v = volumes(H, 2)
    graph_v = volumes(H, 2)
    be = H.Batch(eager_v)
    bg = H.Batch(graph_v)
    assert be.set_schedule(depths, frames) == 0
    assert bg.set_schedule(depths, frames) == 0
    for k in range(n_steps):
        assert be.run(1, use_graph=False) == 0
        assert bg.run(1, use_graph=True) == 0
        be.synchronize()
        bg.synchronize()
        # eager work of the program's own between replays
        _ = torch.arange(1000, device="cuda").sum().item()
        for v1, v2 in zip(eager_v, graph_v):
            assert v1.block_count == v2.block_count
    for v1, v2 in zip(eager_v, graph_v):
        assert same(state(v1), state(v2))


@pytest.mark.parametrize("every", [0, 1, 5])
def test_multi_step_graphs_match_eager(H, S, every):
    """A graph call of 19 steps replays runs of 8 alike steps (all or none
    extracting) from one 8-step graph and the rest from one-step graphs:
    with no extraction (two 8-step graphs, three single steps), extraction
    after every step (the same, each step meshing every scene) and after
    every 5th (no run of 8 is alike: single steps only). States and the last
    meshes equal eager runs of the same schedule."""
    n_steps = 19
    depths, frames = make_inputs(H, S, n_steps, seed0=150)
    caps = [1 << 24] * 3
    runs = []
    for use_graph in (False, True):
        vs = volumes(H, 2)
        b = H.Batch(vs)
        assert b.set_schedule(depths, frames) == 0
        bufs = [torch.empty(c * 6, dtype=torch.float32, device="cuda") for c in caps]
        counts = torch.full((3,), -7, dtype=torch.int32, device="cuda")
        if every:
            assert b.set_extract(bufs, caps, counts, 0.5, 0.0, every=every) == 0
        assert b.run(n_steps, use_graph=use_graph) == 0
        b.synchronize()
        runs.append((vs, bufs, counts.cpu().numpy()))
        b.close()
    (ve, be, ce), (vg, bg, cg) = runs
    for v1, v2 in zip(ve, vg):
        assert v1.block_count == v2.block_count > 0
        assert same(state(v1), state(v2))
    if every:
        assert np.array_equal(ce, cg) and (ce > 0).all()
        for s in range(3):
            assert _same_mesh(be[s], bg[s], ce[s]), f"scene {s}"


def test_mixed_graph_and_eager_runs(H, S):
    """Graph runs read their step index from a device counter that their
    update kernels advance; eager runs in between leave it behind, and a
    graph run after one must set it again: 4 graph, 2 eager, 9 graph (an
    8-step graph and a single step), 1 graph, 1 eager steps equal 17 eager
    steps."""
    n_steps = 17
    depths, frames = make_inputs(H, S, n_steps, seed0=170)
    ref = volumes(H, 2)
    b0 = H.Batch(ref)
    assert b0.set_schedule(depths, frames) == 0
    assert b0.run(n_steps) == 0
    b0.synchronize()
    vs = volumes(H, 2)
    b = H.Batch(vs)
    assert b.set_schedule(depths, frames) == 0
    for k, g in ((4, True), (2, False), (9, True), (1, True), (1, False)):
        assert b.run(k, use_graph=g) == 0
    b.synchronize()
    for v1, v2 in zip(ref, vs):
        assert v1.block_count == v2.block_count > 0
        assert same(state(v1), state(v2))
    b0.close()
    b.close()


def test_member_reads_wait_for_the_batch(H, S):
    """Counters and voxels of a batched volume read straight after an
    asynchronous batch step (no explicit synchronisation) are that step's: the
    reads wait for the batch's own stream, not only the volume's."""
    n_steps = 5
    depths, frames = make_inputs(H, S, n_steps, seed0=110)
    ref = volumes(H, 2)
    b0 = H.Batch(ref)
    assert b0.set_schedule(depths, frames) == 0
    want = []
    for k in range(n_steps):
        assert b0.run(1) == 0
        b0.synchronize()
        want.append([(v.touched_count, v.block_count) for v in ref])
    vs = volumes(H, 2)
    b = H.Batch(vs)
    assert b.set_schedule(depths, frames) == 0
    for k in range(n_steps):
        assert b.run(1) == 0
        assert [(v.touched_count, v.block_count) for v in vs] == want[k], f"step {k}"
    for v1, v2 in zip(ref, vs):
        assert same(state(v1), state(v2))


def test_batch_tuning_keeps_results(H, S):
    n_steps = 4
    depths, frames = make_inputs(H, S, n_steps, seed0=90)
    ref = volumes(H, 2)
    b0 = H.Batch(ref)
    b0.set_schedule(depths, frames)
    b0.run(n_steps)
    b0.synchronize()
    for alloc_grid, upd_grid, threads, lds in ((64, 128, 128, 0), (0, 0, 512, 256)):
        vs = volumes(H, 2)
        b = H.Batch(vs)
        assert b.tune(alloc_grid, upd_grid, threads, lds) == 0
        b.set_schedule(depths, frames)
        assert b.run(n_steps, use_graph=True) == 0
        b.synchronize()
        for v1, v2 in zip(ref, vs):
            assert same(state(v1), state(v2))
        b.close()


def test_batch_rejects_bad_groups(H):
    a = H.HipVolume(0.01, 1024, 0.04, 32.0, storage=2)
    c = H.HipVolume(0.01, 1024, 0.04, 32.0, storage=1)
    r = H.HipVolume(0.01, 1024, 0.04, 32.0)
    with pytest.raises(RuntimeError):
        H.Batch([a, c])       # mixed storage
    with pytest.raises(RuntimeError):
        H.Batch([r])          # weighted-sum storage has no projective update
    with pytest.raises(RuntimeError):
        H.Batch([a, a])       # the same volume twice
    b = H.Batch([a])
    with pytest.raises(RuntimeError):
        H.Batch([a])          # already in a batch
    fr = H.make_frame(8, 8, 4, 4, 4, 4, np.eye(3), np.zeros(3))
    d = torch.ones(64, dtype=torch.float32, device="cuda")
    assert a.integrate_depth(d, fr) == 3  # single calls refused while batched
    b.close()
    assert a.integrate_depth(d, fr) == 0


def _same_mesh(a, b, count):
    """Bitwise equality of the first `count` vertices, on the device (noisy
    depth makes triangle soups of millions of vertices)."""
    n = int(count) * 6
    return bool(torch.equal(a[:n].view(torch.int32), b[:n].view(torch.int32)))


@pytest.mark.parametrize("use_graph", [False, True])
def test_batched_extraction_equals_per_scene(H, S, use_graph):
    """Meshes extracted inside batched steps (every 3rd step, graph-captured
    or eager) are byte-identical to each scene's own device extraction of the
    same state; a buffer too small reports -1."""
    n_steps = 6
    depths, frames = make_inputs(H, S, n_steps, seed0=110)
    solo = volumes(H, 2)
    batched = volumes(H, 2)
    b = H.Batch(batched)
    assert b.set_schedule(depths, frames) == 0
    caps = [1 << 25, 1 << 25, 64]  # the last scene's mesh does not fit
    bufs = [torch.empty(c * 6, dtype=torch.float32, device="cuda") for c in caps]
    bufs[2].fill_(7.0)  # an overflowing extraction writes nothing
    counts = torch.full((3,), -7, dtype=torch.int32, device="cuda")
    assert b.set_extract(bufs, caps, counts, 0.5, 0.0, every=3) == 0
    for k in range(n_steps):
        for s, v in enumerate(solo):
            if depths[k][s] is not None:
                assert v.integrate_depth(depths[k][s], frames[k][s]) == 0
        assert b.run(1, use_graph=use_graph) == 0
        b.synchronize()
        if k % 3 != 2:
            continue
        got = counts.cpu().numpy()
        assert got[2] == -1
        assert bool(torch.all(bufs[2] == 7.0))
        for s in range(2):
            ref = torch.empty(caps[s] * 6, dtype=torch.float32, device="cuda")
            rc = torch.zeros(1, dtype=torch.int32, device="cuda")
            assert solo[s].extract_device(ref, None, caps[s], 0.5, 0.0, rc) == 0
            solo[s].synchronize()
            assert got[s] == int(rc.item()) > 0
            assert _same_mesh(bufs[s], ref, got[s])
            del ref
        counts.fill_(-7)
    b.close()


# The instrumented allocate modes of a batch (probe-depth statistics, counted
# compare-and-swap attempts) change no result, and their counters are
# consistent: every distinct-key lookup lands in one histogram bin, the
# deepest lane of each 64-lane batch bounds the wave steps from above, and
# every inserted block took at least one exchange attempt.
@pytest.mark.parametrize("mode", [1, 2])
def test_instrumented_batch_allocate(H, S, mode):
    import ctypes as C
    n_steps = 5
    depths, frames = make_inputs(H, S, n_steps, seed0=90)
    plain = volumes(H, 2)
    inst = volumes(H, 2)
    bp, bi = H.Batch(plain), H.Batch(inst)
    for b in (bp, bi):
        assert b.set_schedule(depths, frames) == 0
    assert bi.set_alloc_mode(mode) == 0
    assert bp.run(n_steps) == 0 and bi.run(n_steps, use_graph=True) == 0
    bp.synchronize()
    bi.synchronize()
    for v1, v2 in zip(plain, inst):
        assert same(state(v1), state(v2))
        hist, lane, wave, look = v2.probe_stats()
        if mode == 2:
            assert look > v2.block_count and hist.sum() == look
            assert lane >= look and wave >= lane
            assert H.cas_attempts(v2) == 0
        else:
            assert H.cas_attempts(v2) >= v2.block_count > 0
            assert look == 0
    assert bi.set_alloc_mode(3) == 1
