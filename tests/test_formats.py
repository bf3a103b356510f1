"""OSNP / OSN1 files and the analytic point clouds (CPU): byte layouts checked
against an independent numpy rendering of the format definitions
(tools/tartanair_points.py:11-17, cpp/bench/bench_arms.cu:506-523), round
trips, every error path, and the generators against the oracle's restatement
of cpp/tests/analytic_scenes.hpp."""
import struct

import numpy as np
import pytest

import bindings as B
import formats as F


def osnp_bytes(points, cam, magic=b"OSNP", version=1, n=None):
    pts = np.asarray(points, np.float32).reshape(-1, 3)
    n = len(pts) if n is None else n
    return magic + struct.pack("<ii3f", version, n, *map(float, cam)) + pts.tobytes()


def test_osnp_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    pts = rng.normal(size=(1000, 3)).astype(np.float32)
    cam = np.array([0.5, -1.25, 2.0], np.float32)
    p = str(tmp_path / "a.bin")
    F.write_osnp(p, pts, cam)
    assert open(p, "rb").read() == osnp_bytes(pts, cam)
    q = str(tmp_path / "b.bin")
    open(q, "wb").write(osnp_bytes(pts[:7], cam))
    got, c = F.read_osnp(q)
    assert np.array_equal(got, pts[:7]) and np.array_equal(c, cam)


@pytest.mark.parametrize("blob,rc", [
    (osnp_bytes(np.zeros((2, 3)), [0, 0, 0], magic=b"OSNQ"), 3),
    (osnp_bytes(np.zeros((2, 3)), [0, 0, 0], version=2), 3),
    (osnp_bytes(np.zeros((2, 3)), [0, 0, 0], n=0), 3),
    (osnp_bytes(np.zeros((2, 3)), [0, 0, 0], n=3), 4),
    (b"OSNP\x01\x00", 3),
])
def test_osnp_rejects_bad_files(tmp_path, blob, rc):
    p = str(tmp_path / "bad.bin")
    open(p, "wb").write(blob)
    with pytest.raises(F.FormatError) as e:
        F.read_osnp(p)
    assert e.value.rc == rc
    with pytest.raises(F.FormatError) as e:
        F.read_osnp(str(tmp_path / "missing.bin"))
    assert e.value.rc == 2


def test_osn1_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(2)
    m = rng.normal(size=(999, 6)).astype(np.float32)
    p = str(tmp_path / "m.raw")
    F.write_osn1(p, m)
    raw = open(p, "rb").read()
    assert raw[:4] == b"OSN1" and struct.unpack("<I", raw[4:8])[0] == 999
    assert raw[8:] == m.tobytes()
    assert np.array_equal(F.read_osn1(p), m)
    F.write_osn1(p, np.zeros((0, 6), np.float32))
    assert F.read_osn1(p).shape == (0, 6)
    open(p, "wb").write(raw[:-4])
    with pytest.raises(F.FormatError) as e:
        F.read_osn1(p)
    assert e.value.rc == 4
    open(p, "wb").write(b"OSN2" + raw[4:])
    with pytest.raises(F.FormatError) as e:
        F.read_osn1(p)
    assert e.value.rc == 3


@pytest.mark.parametrize("r,nt,npf", [(0.5, 400, 800), (0.25, 37, 91), (2.0, 10, 3)])
def test_sphere_points_match_the_oracle(r, nt, npf):
    assert np.array_equal(F.sphere_points(r, nt, npf).view(np.uint32),
                          B.sphere_points(r, nt, npf).view(np.uint32))


@pytest.mark.parametrize("z0,he,n", [(0.5, 1.0, 566), (0.5, 0.3, 7), (-1.0, 2.5, 100)])
def test_plane_points_match_the_oracle(z0, he, n):
    assert np.array_equal(F.plane_points(z0, he, n).view(np.uint32),
                          B.plane_points(z0, he, n).view(np.uint32))
