"""OSNP point clouds, OSN1 mesh dumps and the analytic point clouds, through
the library's C ABI (include/osn_io.h): the reference harness's formats
(cpp/bench/bench_arms.cu:148-170, 506-523; tools/tartanair_points.py:11-17)."""
from __future__ import annotations

import ctypes as C

import numpy as np

import osn_tsdf_hip as _H

ERRORS = {1: "bad arguments", 2: "cannot open or write", 3: "bad header", 4: "truncated",
          5: "buffer too small"}


def _lib():
    L = _H.lib()
    if not getattr(L, "_io_bound", False):
        P, I32, F = C.c_void_p, C.c_int32, C.c_float
        for name, args in (("osn_osnp_info", [C.c_char_p, P, P]),
                           ("osn_osnp_read", [C.c_char_p, P, I32, P, P]),
                           ("osn_osnp_write", [C.c_char_p, P, I32, P]),
                           ("osn_osn1_info", [C.c_char_p, P]),
                           ("osn_osn1_read", [C.c_char_p, P, I32, P]),
                           ("osn_osn1_write", [C.c_char_p, P, I32])):
            fn = getattr(L, name)
            fn.restype, fn.argtypes = I32, args
        L.osn_points_sphere.restype = None
        L.osn_points_sphere.argtypes = [F, I32, I32, P]
        L.osn_points_plane.restype = None
        L.osn_points_plane.argtypes = [F, F, I32, P]
        L._io_bound = True
    return L


class FormatError(RuntimeError):
    def __init__(self, what, rc):
        super().__init__(f"{what}: {ERRORS.get(rc, rc)}")
        self.rc = rc


def read_osnp(path: str):
    """Returns (points float32 (n, 3), cam float32 (3,))."""
    L = _lib()
    n = C.c_int32(0)
    cam = np.zeros(3, np.float32)
    rc = L.osn_osnp_info(path.encode(), C.addressof(n), cam.ctypes.data)
    if rc:
        raise FormatError(path, rc)
    pts = np.empty((n.value, 3), np.float32)
    rc = L.osn_osnp_read(path.encode(), pts.ctypes.data, n.value, C.addressof(n), cam.ctypes.data)
    if rc:
        raise FormatError(path, rc)
    return pts, cam


def write_osnp(path: str, points, cam) -> None:
    pts = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    c = np.ascontiguousarray(cam, dtype=np.float32).reshape(3)
    rc = _lib().osn_osnp_write(path.encode(), pts.ctypes.data, len(pts), c.ctypes.data)
    if rc:
        raise FormatError(path, rc)


def read_osn1(path: str) -> np.ndarray:
    """Returns posnor float32 (n_vertices, 6)."""
    L = _lib()
    n = C.c_int32(0)
    rc = L.osn_osn1_info(path.encode(), C.addressof(n))
    if rc:
        raise FormatError(path, rc)
    out = np.empty((n.value, 6), np.float32)
    rc = L.osn_osn1_read(path.encode(), out.ctypes.data, n.value, C.addressof(n))
    if rc:
        raise FormatError(path, rc)
    return out


def write_osn1(path: str, posnor) -> None:
    a = np.ascontiguousarray(posnor, dtype=np.float32).reshape(-1, 6)
    rc = _lib().osn_osn1_write(path.encode(), a.ctypes.data, len(a))
    if rc:
        raise FormatError(path, rc)


def sphere_points(radius: float, n_theta: int, n_phi: int) -> np.ndarray:
    out = np.empty(n_theta * n_phi * 3, np.float32)
    _lib().osn_points_sphere(radius, n_theta, n_phi, out.ctypes.data)
    return out


def plane_points(z0: float, half_extent: float, n: int) -> np.ndarray:
    out = np.empty(n * n * 3, np.float32)
    _lib().osn_points_plane(z0, half_extent, n, out.ctypes.data)
    return out
