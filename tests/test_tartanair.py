"""TartanAir ingest (host only): NED pose composition, trajectory files, depth
PNG decoding, and the exporter's world points.

Pins (conventions from crates/tartanair/src/lib.rs:43-58,104-112,170-214 and
tools/tartanair_points.py:21-27,55-100):
  * the identity pose maps a camera point (x, y, z) to body/world (z, x, y);
  * R_wc against an independent numpy restatement of the exporter's
    permutation form (body = cam[(2, 0, 1)], world = R(q) body), for random
    unnormalised quaternions, and against closed-form rotations;
  * depth decoding through PNG files written with each of the five scanline
    filters, bit for bit, special values included;
  * the exporter's FP64 points against the device path's FP32 back-projection
    (the oracle's restatement) of the same frame.
"""
import math
import struct
import zlib

import numpy as np
import pytest

import bindings as B

TA = pytest.importorskip("tartanair")


def _quat_matrix(q):
    x, y, z, w = np.asarray(q, np.float64) / np.linalg.norm(q)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def test_identity_pose_is_the_ned_permutation():
    R = TA.pose_rotation([0, 0, 0, 0, 0, 0, 1])
    assert np.array_equal(R @ np.array([1.0, 2.0, 3.0]), [3.0, 1.0, 2.0])


def test_rotation_matches_the_permutation_form():
    rng = np.random.default_rng(3)
    for _ in range(200):
        q = rng.normal(size=4) * rng.uniform(0.2, 5.0)  # not unit: R(q) normalises
        pose = np.concatenate([rng.normal(size=3), q])
        R = TA.pose_rotation(pose)
        assert np.allclose(R @ R.T, np.eye(3), atol=1e-12) and abs(np.linalg.det(R) - 1) < 1e-12
        cam = rng.normal(size=3)
        body = cam[[2, 0, 1]]  # forward, right, down
        assert np.allclose(R @ cam, _quat_matrix(q) @ body, atol=1e-12)


def test_closed_form_rotation():
    # 90 degrees about NED z (down): forward turns to the right
    s = math.sqrt(0.5)
    R = TA.pose_rotation([0, 0, 0, 0, 0, s, s])
    # camera forward (0, 0, 1) is body forward (1, 0, 0), rotated to body right (0, 1, 0)
    assert np.allclose(R @ [0.0, 0.0, 1.0], [0.0, 1.0, 0.0], atol=1e-15)
    # camera right (1, 0, 0) is body right, rotated to body backward
    assert np.allclose(R @ [1.0, 0.0, 0.0], [-1.0, 0.0, 0.0], atol=1e-15)
    with pytest.raises(TA.TartanAirError):
        TA.pose_rotation([0, 0, 0, 0, 0, 0, 0])


def test_frame_fields():
    pose = [1.25, -2.5, 0.75, 0.1, -0.2, 0.3, 0.9]
    f = TA.make_frame(pose, max_range=8.0)
    R = TA.pose_rotation(pose).ravel()
    assert (f.width, f.height) == (640, 640)
    assert (f.fx, f.fy, f.cx, f.cy) == (320.0, 320.0, 320.0, 320.0)
    assert list(f.R_wc64) == list(R) and list(f.t_wc64) == pose[:3]
    assert list(f.R_wc) == [float(np.float32(x)) for x in R]
    assert list(f.t_wc) == [float(np.float32(x)) for x in pose[:3]]
    assert f.max_range_m == 8.0


def test_load_poses(tmp_path):
    p = tmp_path / "pose.txt"
    p.write_text("1 2 3 0 0 0 1\n\n  4.5\t5 6 0.1 0.2 0.3 0.9  \n7 8 9 x 0 0 0 1\n")
    got = TA.load_poses(str(p))
    assert got.shape == (3, 7)
    assert np.array_equal(got[0], [1, 2, 3, 0, 0, 0, 1])
    assert np.array_equal(got[1], [4.5, 5, 6, 0.1, 0.2, 0.3, 0.9])
    assert np.array_equal(got[2], [7, 8, 9, 0, 0, 0, 1])  # a non-numeric token is skipped
    bad = tmp_path / "bad.txt"
    bad.write_text("1 2 3 0 0 0 1\n1 2 3 0 0 1\n")
    with pytest.raises(TA.TartanAirError) as e:
        TA.load_poses(str(bad))
    assert e.value.rc == 3
    with pytest.raises(TA.TartanAirError) as e:
        TA.load_poses(str(tmp_path / "missing.txt"))
    assert e.value.rc == 2


def _write_png(path, rgba, ftype):
    """8-bit RGBA PNG, every scanline with filter `ftype` (PNG spec 9.2)."""
    h, w, _ = rgba.shape
    img = rgba.astype(np.int32)
    rows = []
    for y in range(h):
        cur = img[y].reshape(-1)
        up = img[y - 1].reshape(-1) if y else np.zeros_like(cur)
        left = np.concatenate([np.zeros(4, np.int32), cur[:-4]])
        upl = np.concatenate([np.zeros(4, np.int32), up[:-4]])
        if ftype == 0:
            f = cur
        elif ftype == 1:
            f = cur - left
        elif ftype == 2:
            f = cur - up
        elif ftype == 3:
            f = cur - (left + up) // 2
        else:
            p = left + up - upl
            pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - upl)
            f = cur - np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, upl))
        rows.append(bytes([ftype]) + (f & 0xFF).astype(np.uint8).tobytes())

    def chunk(kind, body):
        return (struct.pack(">I", len(body)) + kind + body +
                struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF))

    with open(path, "wb") as fh:
        fh.write(b"\x89PNG\r\n\x1a\n")
        fh.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)))
        fh.write(chunk(b"IDAT", zlib.compress(b"".join(rows))))
        fh.write(chunk(b"IEND", b""))


@pytest.mark.parametrize("ftype", [0, 1, 2, 3, 4])
def test_depth_png_round_trip(tmp_path, ftype):
    rng = np.random.default_rng(ftype)
    depth = rng.uniform(0.1, 40.0, size=(9, 13)).astype(np.float32)
    depth[0, 0], depth[0, 1], depth[0, 2], depth[0, 3] = 0.0, np.inf, np.nan, 1e-30
    le = depth.view(np.uint8).reshape(9, 13, 4)  # little-endian bytes b0..b3
    rgba = np.stack([le[..., 2], le[..., 1], le[..., 0], le[..., 3]], -1)  # B,G,R,A = b0..b3
    path = tmp_path / f"d{ftype}.png"
    _write_png(path, rgba, ftype)
    assert np.array_equal(TA.read_png_rgba(str(path)), rgba)
    got = TA.read_depth(str(path))
    assert got.shape == depth.shape and np.array_equal(got.view(np.uint32), depth.view(np.uint32))


def test_exporter_points_match_the_device_backprojection():
    rng = np.random.default_rng(5)
    depth = rng.uniform(0.3, 6.0, size=(64, 80)).astype(np.float32)
    pose = [0.5, -1.0, 0.2, 0.05, -0.1, 0.2, 0.95]
    pts, cam = TA.unproject(depth, pose, max_range=10.0)
    f = TA.make_frame(pose, width=80, height=64)
    dev = B.backproject_f32(depth, f.fx, f.fy, f.cx, f.cy, np.array(f.R_wc[:], np.float32),
                            np.array(f.t_wc[:], np.float32)).reshape(-1, 3)
    assert pts.shape == dev.shape
    assert np.abs(pts - dev).max() < 1e-5 * (1.0 + np.abs(pts).max())
    assert np.array_equal(cam, np.float32(pose[:3]))
