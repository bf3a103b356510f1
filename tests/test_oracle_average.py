"""Closed-form pin of the oracle's projective running average, on the CPU.

The oracle's projective update (orc_proj_integrate) folds each observation x
into t <- (t * w + x) / (w + 1), w <- w + 1, and leaves a voxel alone once its
weight has reached the cap w_max: the reference's weight cap freezes a voxel
(cpp/include/osn_tsdf/types.hpp:99-101; the skip is cpp/src/volume.cpp:289),
and its read-side mean is sum / w (cpp/src/volume.cpp:322-324). With unit
weights both say the same thing in closed form:

    t = (x_1 + ... + x_m') / m',   w = m',   m' = min(m, w_max)

where x_1..x_m are the voxel's observations in frame order. The scene makes
every x known without the oracle's arithmetic: the camera sits at the origin
looking along +z (identity pose) and frame k's depth image is one constant d_k,
so a voxel centre at depth z sees x_k = clamp((d_k - z) / trunc, -1, 1)
whenever d_k - z >= -trunc.

The check is shown to catch the plausible mistakes: the oracle source is
compiled again with each of four planted errors (-DORC_PLANT=n: the cap only
bounding the divisor, i.e. no freeze; the cap applied before averaging; the
last observation kept alone; freezing one observation early), and every one
fails it.
"""
import os
import subprocess

import numpy as np
import pytest

import bindings as B

VOXEL = np.float32(0.02)
TRUNC = np.float32(0.08)
W, H = 64, 48
FX = FY = 32.0
CX, CY = 32.0, 24.0
N_FRAMES = 12
# blocks (bx, by) in {-1, 0}^2 and bz 5..9: voxel centres within 0.15 m of the
# optical axis at 0.81-1.59 m, all inside the image
BLOCKS = [(bx, by, bz) for bx in (-1, 0) for by in (-1, 0) for bz in range(5, 10)]


def wall_depths(seed=7):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.95, 1.35, N_FRAMES).astype(np.float32)


def run_oracle(L, w_max, depths):
    po = B.ProjOracle(float(VOXEL), float(TRUNC), w_max, len(BLOCKS), L=L)
    keys = np.array([B.pack_coord(*b) for b in BLOCKS], np.int64)
    I64 = np.eye(3)
    I32 = np.eye(3, dtype=np.float32)
    for d in depths:
        img = np.full((H, W), d, np.float32)
        po.integrate(keys, img, FX, FY, CX, CY, I64, np.zeros(3), I32, np.zeros(3, np.float32))
    out = {}
    for b, k in zip(BLOCKS, keys):
        out[b] = po.block(int(k))
    return out


def closed_form(depths, w_max):
    """Expected (t, w) per voxel of every block, and a mask of voxels whose sdf
    gate is too close to call (|d - z + trunc| < 1e-5 in some frame)."""
    vs = float(VOXEL)
    tr = float(TRUNC)
    res = {}
    for b in BLOCKS:
        lz, ly, lx = np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij")
        z = ((b[2] * 8 + lz) + 0.5) * vs  # voxel index order (lz * 8 + ly) * 8 + lx
        z = z.reshape(512)
        obs = [(float(d) - z) for d in depths]
        ambiguous = np.zeros(512, bool)
        sums = np.zeros(512)
        cnt = np.zeros(512, np.int64)
        for s in obs:
            ambiguous |= np.abs(s + tr) < 1e-5
            seen = s >= -tr
            take = seen & (cnt < w_max)  # the first w_max observations count
            x = np.clip(s / tr, -1.0, 1.0)
            sums[take] += x[take]
            cnt[take] += 1
        t = np.where(cnt > 0, sums / np.maximum(cnt, 1), 0.0)
        res[b] = (t, cnt, ambiguous)
    return res


def average_errors(L, w_max, depths):
    """(max |t - closed form|, number of weight mismatches, voxels checked,
    voxels whose observation count exceeds w_max)."""
    got = run_oracle(L, w_max, depths)
    exp = closed_form(depths, w_max)
    err = 0.0
    wbad = 0
    checked = 0
    saturated = 0
    for b in BLOCKS:
        t, w = got[b]
        te, we, amb = exp[b]
        keep = ~amb
        err = max(err, float(np.abs(t - te)[keep].max()))
        wbad += int((w[keep] != we[keep]).sum())
        checked += int(keep.sum())
        saturated += int(((we == w_max) & keep).sum())
    return err, wbad, checked, saturated


@pytest.mark.parametrize("w_max", [1, 5, 65535])
def test_running_average_is_mean_of_first_wmax_observations(w_max):
    depths = wall_depths()
    err, wbad, checked, saturated = average_errors(B.lib(), w_max, depths)
    assert checked > 4000
    assert wbad == 0
    assert err < 1e-12
    if w_max < N_FRAMES:
        assert saturated > 1000  # the cap is reached, so the freeze is exercised


def test_unobserved_and_partially_observed_voxels():
    """Voxels behind every wall position stay at (0, 0); voxels in front of
    every position see all frames, clipped to x = 1 far in front."""
    depths = wall_depths()
    exp = closed_form(depths, 65535)
    got = run_oracle(B.lib(), 65535, depths)
    zero = full = partial = 0
    for b in BLOCKS:
        t, w = got[b]
        _, we, _ = exp[b]
        zero += int(((w == 0) & (t == 0.0)).sum())
        full += int((w == N_FRAMES).sum())
        partial += int(((w > 0) & (w < N_FRAMES)).sum())
    assert zero > 0 and full > 0 and partial > 0
    # a block entirely in front of the nearest wall position (z < 0.95 - 0.08)
    t, w = got[(0, 0, 5)]
    near = np.arange(512) // 64 < 2  # lz 0, 1: z = 0.81, 0.83 m
    assert np.all(w[near] == N_FRAMES) and np.all(t[near] == 1.0)


@pytest.fixture(scope="module")
def planted(tmp_path_factory):
    src = os.path.join(os.path.dirname(B.__file__), "tsdf_oracle.c")
    d = tmp_path_factory.mktemp("plant")
    libs = {}
    for k in (1, 2, 3, 4):
        so = str(d / f"liboracle_plant{k}.so")
        subprocess.run(["gcc", "-std=c11", "-O2", "-ffp-contract=off", "-fPIC", "-shared",
                        f"-DORC_PLANT={k}", "-o", so, src, "-lm"], check=True)
        libs[k] = B.load_variant(so)
    return libs


@pytest.mark.parametrize("plant", [1, 2, 3, 4])
def test_planted_average_mistakes_fail_the_closed_form(planted, plant):
    err, wbad, checked, _ = average_errors(planted[plant], 5, wall_depths())
    assert checked > 4000
    assert err > 1e-6 or wbad > 0, f"planted mistake {plant} passed the closed-form check"
