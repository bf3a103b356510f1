"""TartanAir V2 ingest through the library's C ABI (include/osn_tartanair.h).

Trajectory files, NED pose composition (FP64 on the host, FP32 for the
device), depth PNG decoding, and the reference tool's world-point export:

    poses = load_poses("pose_lcam_front.txt")            # (n, 7): t, q (NED)
    depth = read_depth("000000_lcam_front_depth.png")    # (640, 640) float32 metres
    frame = make_frame(poses[0])                         # OsnTsdfHipFrame
    v.integrate_depth(torch.from_numpy(depth).cuda().view(-1), frame)

The conventions are the reference loader's (crates/tartanair/src/lib.rs,
tools/tartanair.py, tools/tartanair_points.py); see the header for citations.
The PNG decoder covers what the dataset stores: 8-bit RGBA, not interlaced,
any of the five scanline filters (PNG specification, ISO/IEC 15948, 9.2).
"""
from __future__ import annotations

import ctypes as C
import struct
import zlib

import numpy as np

import osn_tsdf_hip as _H

WIDTH = HEIGHT = 640
FX = FY = CX = CY = 320.0

ERRORS = {1: "bad arguments", 2: "cannot open", 3: "a pose line without exactly 7 numbers",
          5: "buffer too small"}


def _lib():
    L = _H.lib()
    if not getattr(L, "_ta_bound", False):
        P, I32, D, F = C.c_void_p, C.c_int32, C.c_double, C.c_float
        L.osn_ta_pose_rotation.restype, L.osn_ta_pose_rotation.argtypes = I32, [P, P]
        L.osn_ta_frame.restype = I32
        L.osn_ta_frame.argtypes = [P, I32, I32, D, D, D, D, F, P]
        L.osn_ta_load_poses.restype, L.osn_ta_load_poses.argtypes = I32, [C.c_char_p, P, I32, P]
        L.osn_ta_depth_from_rgba.restype = None
        L.osn_ta_depth_from_rgba.argtypes = [P, C.c_int64, P]
        L._ta_bound = True
    return L


class TartanAirError(RuntimeError):
    def __init__(self, what, rc):
        super().__init__(f"{what}: {ERRORS.get(rc, rc)}")
        self.rc = rc


def _pose(pose) -> np.ndarray:
    p = np.ascontiguousarray(pose, dtype=np.float64).reshape(7)
    return p


def pose_rotation(pose) -> np.ndarray:
    """Camera-to-world rotation (3, 3) of a NED pose (tx, ty, tz, qx, qy, qz, qw)."""
    p = _pose(pose)
    R = np.empty(9, np.float64)
    rc = _lib().osn_ta_pose_rotation(p.ctypes.data, R.ctypes.data)
    if rc:
        raise TartanAirError("pose", rc)
    return R.reshape(3, 3)


def make_frame(pose, width=WIDTH, height=HEIGHT, fx=FX, fy=FY, cx=CX, cy=CY,
               max_range=0.0) -> _H.Frame:
    """The depth-path frame descriptor of a NED pose."""
    p = _pose(pose)
    f = _H.Frame()
    rc = _lib().osn_ta_frame(p.ctypes.data, int(width), int(height), float(fx), float(fy),
                             float(cx), float(cy), float(max_range), C.addressof(f))
    if rc:
        raise TartanAirError("frame", rc)
    return f


def load_poses(path: str) -> np.ndarray:
    """(n, 7) float64 poses of a trajectory file."""
    L = _lib()
    n = C.c_int32(0)
    rc = L.osn_ta_load_poses(str(path).encode(), None, 0, C.addressof(n))
    if rc not in (0, 5):
        raise TartanAirError(str(path), rc)
    out = np.empty((max(n.value, 1), 7), np.float64)
    rc = L.osn_ta_load_poses(str(path).encode(), out.ctypes.data, n.value, C.addressof(n))
    if rc:
        raise TartanAirError(str(path), rc)
    return out[: n.value]


def depth_from_rgba(rgba: np.ndarray) -> np.ndarray:
    """Planar depth (float32 metres) from RGBA samples (..., 4) in file order."""
    a = np.ascontiguousarray(rgba, dtype=np.uint8)
    if a.shape[-1] != 4:
        raise ValueError("expected 4 channels")
    out = np.empty(a.shape[:-1], np.float32)
    _lib().osn_ta_depth_from_rgba(a.ctypes.data, out.size, out.ctypes.data)
    return out


def _paeth(a, b, c):
    p = a.astype(np.int16) + b - c
    pa, pb, pc = np.abs(p - a), np.abs(p - b), np.abs(p - c)
    return np.where((pa <= pb) & (pa <= pc), a, np.where(pb <= pc, b, c)).astype(np.uint8)


def read_png_rgba(path: str) -> np.ndarray:
    """(height, width, 4) uint8 samples of an 8-bit RGBA, non-interlaced PNG."""
    with open(path, "rb") as f:
        data = f.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"{path}: not a PNG")
    pos, idat, ihdr = 8, [], None
    while pos + 8 <= len(data):
        length, kind = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + length]
        if kind == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
        pos += 12 + length
    if ihdr is None:
        raise ValueError(f"{path}: no IHDR")
    w, h, depth, color, _, _, interlace = ihdr
    if depth != 8 or color != 6 or interlace != 0:
        raise ValueError(f"{path}: need 8-bit RGBA without interlace")
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8)
    stride = w * 4
    if raw.size != h * (stride + 1):
        raise ValueError(f"{path}: truncated image data")
    rows = raw.reshape(h, stride + 1)
    out = np.empty((h, stride), np.uint8)
    prev = np.zeros(stride, np.uint8)
    for y in range(h):
        ftype, line = rows[y, 0], rows[y, 1:].copy()
        if ftype == 1:    # Sub: bytes of the pixel to the left
            for x in range(4, stride, 4):
                line[x:x + 4] += line[x - 4:x]
        elif ftype == 2:  # Up
            line += prev
        elif ftype == 3:  # Average
            for x in range(0, stride, 4):
                left = line[x - 4:x].astype(np.uint16) if x else np.zeros(4, np.uint16)
                line[x:x + 4] += ((left + prev[x:x + 4]) >> 1).astype(np.uint8)
        elif ftype == 4:  # Paeth
            for x in range(0, stride, 4):
                left = line[x - 4:x] if x else np.zeros(4, np.uint8)
                upl = prev[x - 4:x] if x else np.zeros(4, np.uint8)
                line[x:x + 4] += _paeth(left, prev[x:x + 4], upl)
        elif ftype != 0:
            raise ValueError(f"{path}: unknown filter {ftype}")
        out[y] = line
        prev = line
    return out.reshape(h, w, 4)


def read_depth(path: str) -> np.ndarray:
    """A TartanAir depth PNG as (height, width) float32 metres."""
    return depth_from_rgba(read_png_rgba(path))


def unproject(depth: np.ndarray, pose, stride: int = 1, max_range: float = 1e30):
    """World points (float32 (n, 3)) and camera centre of one frame, in FP64 as
    the reference's OSNP exporter computes them (tools/tartanair_points.py:
    80-100): pixels with finite depth in (0.05, max_range), camera point
    ((u - cx) z / fx, (v - cy) z / fy, z), world = R_wc @ it + t."""
    h, w = depth.shape
    ys, xs = np.mgrid[0:h:stride, 0:w:stride]
    z = depth[::stride, ::stride].astype(np.float64)
    ok = np.isfinite(z) & (z > 0.05) & (z < max_range)
    xs, ys, z = xs[ok], ys[ok], z[ok]
    cam = np.stack([(xs - CX) * z / FX, (ys - CY) * z / FY, z], axis=1)
    p = _pose(pose)
    world = cam @ pose_rotation(p).T + p[:3]
    return world.astype(np.float32), p[:3].astype(np.float32)
