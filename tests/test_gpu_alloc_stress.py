"""Allocator stress on the GPU: 2^22 Zipf-shuffled inserts over 2^18 keys into
tables at load factors 0.5 / 0.7 / 0.9 (non-power-of-two capacities, mixed
hash) and the reference hash at 0.5. The key set must equal the universe (the
closed-form oracle) exactly, nothing may drop, and the pool order is sorted.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def data():
    import stress
    uni, stream = stress.make_stream(seed=11)
    return stress, uni, stream


@pytest.mark.parametrize("label,lf,hk", [("0.5", 0.5, 1), ("0.7", 0.7, 1), ("0.9", 0.9, 1),
                                         ("0.5-ref", 0.5, 0)])
def test_stress_key_set_exact(data, label, lf, hk):
    import osn_tsdf_hip as H
    stress, uni, stream = data
    K = len(uni)
    cap = stress.capacity(K, lf, hk)
    v = H.HipVolume(0.01, K, -1.0, 0.0, storage=H.STORE_F16, table_slots=cap, hash_kind=hk)
    d = torch.from_numpy(stream.ravel()).cuda()
    assert v.insert_blocks(d) == 0
    assert v.block_count == K
    assert v.drop_count == 0
    keys = v.keys()
    assert np.all(np.diff(keys) > 0)  # pool order = sorted keys
    assert np.array_equal(keys, np.sort(stress.pack_keys(uni)))
    # inserting the stream again finds everything and adds nothing
    assert v.insert_blocks(d) == 0
    assert v.block_count == K and v.drop_count == 0


def test_probe_histogram_consistent(data):
    import osn_tsdf_hip as H
    stress, uni, stream = data
    K = len(uni)
    d = torch.from_numpy(stream.ravel()).cuda()
    means = []
    for lf in (0.5, 0.9):
        v = H.HipVolume(0.01, K, -1.0, 0.0, storage=H.STORE_F16,
                        table_slots=stress.capacity(K, lf, 1), hash_kind=1)
        assert H.insert_blocks_mode(v, d, 2) == 0
        hist, lane, wave, lookups = v.probe_stats()
        assert int(hist.sum()) == lookups >= K
        assert wave >= lane >= lookups
        means.append(lane / lookups)
        assert np.array_equal(v.keys(), np.sort(stress.pack_keys(uni)))
    assert means[1] > means[0]


def test_counting_cas_and_ceiling(data):
    import osn_tsdf_hip as H
    stress, uni, stream = data
    K = len(uni)
    v = H.HipVolume(0.01, K, -1.0, 0.0, storage=H.STORE_F16,
                    table_slots=stress.capacity(K, 0.7, 1), hash_kind=1)
    d = torch.from_numpy(stream.ravel()).cuda()
    assert H.insert_blocks_mode(v, d, 1) == 0
    assert H.cas_attempts(v) >= K
    assert H.cas_ceiling(16, reps=2) > 0


def _coords(rng, n, lim=1 << 10):
    return rng.integers(-lim, lim, size=(n, 3), dtype=np.int32)


@pytest.mark.parametrize("sizes", [(300, 5000), (20000, 9000, 1), (70000, 4097)])
def test_consecutive_calls_pool_order(sizes):
    """Each call's new keys get indices bc0, bc0+1, ... in key order, for calls
    whose new-key counts need one tile, bitonic tiles and multi-level merges,
    on both counter sets (consecutive calls)."""
    import stress
    import osn_tsdf_hip as H
    rng = np.random.default_rng(5)
    v = H.HipVolume(0.01, 1 << 17, -1.0, 0.0, storage=H.STORE_F16, table_slots=1 << 18,
                    hash_kind=1)
    seen = set()
    expect = []
    for n in sizes:
        c = _coords(rng, n)
        assert v.insert_blocks(torch.from_numpy(c.ravel()).cuda()) == 0
        k = stress.pack_keys(c)
        new = np.unique(k[~np.isin(k, np.fromiter(seen, np.int64, len(seen)))])
        expect.extend(new.tolist())
        seen.update(new.tolist())
        got = v.keys()
        assert v.block_count == len(expect)
        assert np.array_equal(got, np.array(expect, np.int64))


def test_extreme_key_sorts_last():
    """Block (2^20-1)^3 packs to INT64_MAX, the value the in-LDS sort pads
    with; it must survive the sort as the largest key."""
    import stress
    import osn_tsdf_hip as H
    rng = np.random.default_rng(9)
    c = np.concatenate([_coords(rng, 700), np.array([[(1 << 20) - 1] * 3, [-(1 << 20)] * 3],
                                                     np.int32)])
    v = H.HipVolume(0.01, 4096, -1.0, 0.0, storage=H.STORE_F16, table_slots=1 << 13,
                    hash_kind=1)
    assert v.insert_blocks(torch.from_numpy(c.ravel()).cuda()) == 0
    want = np.unique(stress.pack_keys(c))
    assert want[-1] == np.iinfo(np.int64).max
    assert np.array_equal(v.keys(), want)


@pytest.mark.parametrize("box,offset", [((16, 16, 16), (0, 0, 0)), ((1, 1, 4096), (-7, 3, -2000)),
                                        ((64, 64, 1), (100, -200, 5)), ((32, 32, 64), (-16, -16, -32)),
                                        ((40, 40, 40), (1, 2, 3))])
def test_small_box_calls_pool_order(box, offset):
    """Calls whose new keys fill a small box of block coordinates (ranked
    through a bitmap of at most 2^16 cells: 16^3, a line, a plane, the
    32x32x64 limit; 40^3 exceeds it and is sorted), with more than 256 new
    keys per call: pool order is the per-call key order, over consecutive
    calls into the same volume."""
    import stress
    import osn_tsdf_hip as H
    rng = np.random.default_rng(11)
    v = H.HipVolume(0.01, 1 << 15, -1.0, 0.0, storage=H.STORE_F16, table_slots=1 << 16,
                    hash_kind=1)
    seen = set()
    expect = []
    lo = np.array(offset, np.int32)
    ext = np.array(box, np.int32)
    for n in (700, 3000, 2500):
        c = (lo + (rng.random((n, 3)) * ext).astype(np.int32)).astype(np.int32)
        assert v.insert_blocks(torch.from_numpy(c.ravel()).cuda()) == 0
        k = stress.pack_keys(c)
        new = np.unique(k[~np.isin(k, np.fromiter(seen, np.int64, len(seen)))])
        expect.extend(new.tolist())
        seen.update(new.tolist())
        assert v.block_count == len(expect)
        assert np.array_equal(v.keys(), np.array(expect, np.int64))
    assert len(expect) > 1000
