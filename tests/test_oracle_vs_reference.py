"""The oracle against the reference CPU arm itself.

Two routes, on the CPU:
  * the committed fixtures (tests/golden/reference_arm.npz, written by
    tests/golden/make_golden.py from the compiled reference arm run
    single-threaded): block counts, vertex counts and the sha256 of the mesh
    bytes must match exactly -- the mesh bytes depend on the key set, the pool
    order and every weighted sum the extractor reads;
  * when oracle/_ref/libosn_ref.so is built (it is compiled from the
    reference's own sources by oracle/Makefile), the same comparison on more
    inputs, plus the multi-threaded reference arm's block count.
"""
import hashlib
import os

import numpy as np
import pytest

import bindings as B

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_arm.npz")


def _cases(z):
    return sorted({k.rsplit("_", 1)[0] for k in z.files if k.endswith("_params")})


def _run_oracle(pts, p):
    voxel, pool, trunc, wcap, cx, cy, cz, radius = p
    v = B.Volume(float(voxel), int(pool), float(trunc), float(wcap))
    assert v.integrate(pts, (cx, cy, cz), float(radius)) == 0
    return v


def test_oracle_matches_golden_reference_outputs():
    z = np.load(GOLDEN)
    names = _cases(z)
    assert len(names) >= 4
    for name in names:
        v = _run_oracle(z[f"{name}_points"], z[f"{name}_params"])
        blocks, drops = z[f"{name}_blocks"]
        assert (v.block_count, v.drop_count) == (blocks, drops), name
        n = v.count_mesh(0.5, 0.0)
        assert n == z[f"{name}_mesh_n"][0], name
        if n:
            mesh = v.extract(0.5, 0.0)
            digest = hashlib.sha256(np.ascontiguousarray(mesh).tobytes()).digest()
            assert digest == bytes(z[f"{name}_mesh_sha256"]), f"{name}: mesh bytes differ"


def _ref_or_skip():
    if not B.ref_available():
        try:
            B.build()
        except Exception:
            pass
    if not B.ref_available():
        pytest.skip("reference CPU arm not built here; the golden fixtures cover it")


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_equals_single_threaded_reference(seed):
    _ref_or_skip()
    rng = np.random.default_rng(100 + seed)
    # a noisy shell seen from an off-centre camera, crossing the origin
    n = 40000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = rng.uniform(0.3, 0.6, size=(n, 1))
    pts = (d * r + rng.normal(scale=0.01, size=(n, 3))).astype(np.float32).ravel()
    cam = tuple(rng.uniform(-0.05, 0.05, 3))
    for wcap in (0.0, 4.0):
        ref = B.RefVolume(0.01, 1 << 15, -1.0, wcap, threads=1)
        ref.integrate(pts, cam)
        v = B.Volume(0.01, 1 << 15, -1.0, wcap)
        v.integrate(pts, cam)
        assert v.block_count == ref.block_count
        a = ref.extract(0.5, 0.0)
        b = v.extract(0.5, 0.0)
        assert np.array_equal(a, b)


def test_multithreaded_reference_block_count_matches():
    _ref_or_skip()
    pts = B.sphere_points(0.5, 400, 800)
    ref = B.RefVolume(0.01, 1 << 14, -1.0, 32.0)
    ref.integrate(pts, (0, 0, 0))
    v = B.Volume(0.01, 1 << 14, -1.0, 32.0)
    v.integrate(pts, (0, 0, 0))
    assert ref.block_count == v.block_count == 1160
    assert len(ref.extract(0.5, 0.0)) == v.count_mesh(0.5, 0.0) == 846288
