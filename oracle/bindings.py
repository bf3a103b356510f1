"""ctypes bindings for the checker libraries under oracle/.

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as part of the product path.

Two libraries:
  * liboracle.so         -- this repository's sequential C restatement
                            (oracle/tsdf_oracle.c), cited function by function
                            against the reference sources there;
  * _ref/libosn_ref.so   -- the reference CPU arm itself, compiled from
                            cpp/src/{volume,marching_tets,c_api}.cpp by
                            oracle/Makefile when the reference checkout exists.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libosn_ref.so")

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)


def _ptr(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def build() -> None:
    """Compile liboracle.so (and _ref when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "all"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        _lib = _declare(C.CDLL(ORACLE_SO))
    return _lib


def load_variant(path: str):
    """A separately built copy of the oracle (for instance one compiled with a
    planted mistake, -DORC_PLANT=n), with the same signatures declared."""
    return _declare(C.CDLL(path))


def _declare(L):
    L.orc_create.restype = C.c_void_p
    L.orc_create.argtypes = [C.c_float, C.c_int32, C.c_float, C.c_float]
    L.orc_destroy.argtypes = [C.c_void_p]
    for name in ("orc_allocate_blocks",):
        getattr(L, name).argtypes = [C.c_void_p, f32p, C.c_int32, C.c_float, C.c_float,
                                     C.c_float, C.c_float]
        getattr(L, name).restype = C.c_int32
    for name in ("orc_update_voxels", "orc_integrate"):
        getattr(L, name).argtypes = [C.c_void_p, f32p, u8p, f32p, C.c_int32, C.c_float,
                                     C.c_float, C.c_float, C.c_float]
        getattr(L, name).restype = C.c_int32
    L.orc_count_visits.argtypes = [C.c_void_p, f32p, C.c_int32, C.c_float, C.c_float,
                                   C.c_float, C.c_float]
    L.orc_count_visits.restype = C.c_int64
    L.orc_touched_keys.argtypes = [C.c_void_p, f32p, C.c_int32, C.c_float, C.c_float,
                                   C.c_float, C.c_float, i64p, C.c_int32]
    L.orc_touched_keys.restype = C.c_int32
    L.orc_extract_mesh.argtypes = [C.c_void_p, f32p, u8p, C.c_int32, C.c_float, C.c_float]
    L.orc_extract_mesh.restype = C.c_int32
    L.orc_block_count.argtypes = [C.c_void_p]
    L.orc_block_count.restype = C.c_int32
    L.orc_drop_count.argtypes = [C.c_void_p]
    L.orc_drop_count.restype = C.c_uint64
    L.orc_table_size.argtypes = [C.c_void_p]
    L.orc_table_size.restype = C.c_uint32
    L.orc_trunc.argtypes = [C.c_void_p]
    L.orc_trunc.restype = C.c_float
    L.orc_block_keys.argtypes = [C.c_void_p, i64p, C.c_int32]
    L.orc_block_keys.restype = C.c_int32
    L.orc_find_key.argtypes = [C.c_void_p, C.c_int64]
    L.orc_find_key.restype = C.c_int32
    L.orc_block_voxels.argtypes = [C.c_void_p, C.c_int32, f32p, f32p, f32p, f32p, f32p]
    L.orc_block_voxels.restype = C.c_int32
    L.orc_load_block.argtypes = [C.c_void_p, C.c_int64, f32p, f32p, C.c_int32]
    L.orc_load_block.restype = C.c_int32
    L.orc_pack_coord.argtypes = [C.c_int32] * 3
    L.orc_pack_coord.restype = C.c_int64
    L.orc_unpack_key.argtypes = [C.c_int64, i32p]
    L.orc_hash_block.argtypes = [C.c_int32] * 3 + [C.c_uint32]
    L.orc_hash_block.restype = C.c_uint32
    L.orc_floor_div.argtypes = [C.c_int32, C.c_int32]
    L.orc_floor_div.restype = C.c_int32
    L.orc_scene_sphere.argtypes = [C.c_float, C.c_int32, C.c_int32, f32p]
    L.orc_scene_plane.argtypes = [C.c_float, C.c_float, C.c_int32, f32p]
    L.orc_backproject_f32.argtypes = [f32p, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                      C.c_float, C.c_float, f32p, f32p, C.c_float, f32p]
    L.orc_backproject_f32.restype = C.c_int32
    L.orc_proj_observe_f32.argtypes = [f32p, f32p, C.c_float, C.c_float, C.c_float,
                                       C.c_float, C.c_int32, C.c_int32, f32p, C.c_float,
                                       C.c_float] + [C.c_int32] * 6 + [f32p]
    L.orc_proj_observe_f32.restype = C.c_int32
    L.orc_proj_create.restype = C.c_void_p
    L.orc_proj_create.argtypes = [C.c_float, C.c_float, C.c_int32, C.c_int32]
    L.orc_proj_destroy.argtypes = [C.c_void_p]
    L.orc_proj_integrate.argtypes = [C.c_void_p, i64p, C.c_int32, f32p, C.c_int32,
                                     C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float,
                                     f64p, f64p, f32p, f32p]
    L.orc_proj_integrate.restype = C.c_int64
    L.orc_proj_block_count.argtypes = [C.c_void_p]
    L.orc_proj_block_count.restype = C.c_int32
    L.orc_proj_block.argtypes = [C.c_void_p, C.c_int64, f64p, i32p]
    L.orc_proj_block.restype = C.c_int32
    return L


# ---- scene helpers --------------------------------------------------------

def sphere_points(radius: float, n_theta: int, n_phi: int) -> np.ndarray:
    out = np.empty(n_theta * n_phi * 3, np.float32)
    lib().orc_scene_sphere(radius, n_theta, n_phi, _ptr(out, f32p))
    return out


def plane_points(z0: float, half_extent: float, n: int) -> np.ndarray:
    out = np.empty(n * n * 3, np.float32)
    lib().orc_scene_plane(z0, half_extent, n, _ptr(out, f32p))
    return out


def pack_coord(x: int, y: int, z: int) -> int:
    return lib().orc_pack_coord(x, y, z)


def backproject_f32(depth: np.ndarray, fx, fy, cx, cy, R: np.ndarray, t: np.ndarray,
                    max_range: float = 0.0) -> np.ndarray:
    depth = np.ascontiguousarray(depth, np.float32)
    H, W = depth.shape
    out = np.empty(W * H * 3, np.float32)
    R = np.ascontiguousarray(R, np.float32).reshape(9)
    t = np.ascontiguousarray(t, np.float32).reshape(3)
    n = lib().orc_backproject_f32(_ptr(depth, f32p), W, H, fx, fy, cx, cy, _ptr(R, f32p),
                                  _ptr(t, f32p), max_range, _ptr(out, f32p))
    return out[: n * 3].copy()


class Volume:
    """Sequential reference-semantics volume (weighted sums, Teschner table)."""

    def __init__(self, voxel: float, pool: int, trunc: float = -1.0, weight_cap: float = 32.0):
        self._L = lib()
        self.h = self._L.orc_create(voxel, pool, trunc, weight_cap)
        if not self.h:
            raise ValueError("orc_create failed")
        self.voxel = voxel
        self.pool = pool

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_destroy(self.h)
            self.h = None

    @property
    def trunc(self) -> float:
        return self._L.orc_trunc(self.h)

    def allocate(self, pts, cam, radius=0.0) -> int:
        pts = np.ascontiguousarray(pts, np.float32)
        return self._L.orc_allocate_blocks(self.h, _ptr(pts, f32p), pts.size // 3, *map(float, cam),
                                           radius)

    def update(self, pts, cam, radius=0.0, colors=None, weights=None) -> int:
        pts = np.ascontiguousarray(pts, np.float32)
        return self._L.orc_update_voxels(self.h, _ptr(pts, f32p), _ptr(colors, u8p),
                                         _ptr(weights, f32p), pts.size // 3, *map(float, cam),
                                         radius)

    def integrate(self, pts, cam, radius=0.0) -> int:
        rc = self.allocate(pts, cam, radius)
        return rc or self.update(pts, cam, radius)

    def count_visits(self, pts, cam, radius=0.0) -> int:
        pts = np.ascontiguousarray(pts, np.float32)
        return self._L.orc_count_visits(self.h, _ptr(pts, f32p), pts.size // 3, *map(float, cam),
                                        radius)

    def touched_keys(self, pts, cam, radius=0.0) -> np.ndarray:
        pts = np.ascontiguousarray(pts, np.float32)
        cap = max(self.block_count, 1)
        out = np.empty(cap, np.int64)
        n = self._L.orc_touched_keys(self.h, _ptr(pts, f32p), pts.size // 3, *map(float, cam),
                                     radius, _ptr(out, i64p), cap)
        if n < 0:
            raise RuntimeError("touched_keys overflow")
        return out[:n].copy()

    @property
    def block_count(self) -> int:
        return self._L.orc_block_count(self.h)

    @property
    def drop_count(self) -> int:
        return self._L.orc_drop_count(self.h)

    def keys(self) -> np.ndarray:
        n = self.block_count
        out = np.empty(max(n, 1), np.int64)
        got = self._L.orc_block_keys(self.h, _ptr(out, i64p), n)
        return out[:got].copy()

    def block_voxels(self, key: int):
        """(tsdf_sum, weight, r, g, b) for a block key, each 512 float32 (z,y,x order)."""
        bi = self._L.orc_find_key(self.h, int(key))
        if bi < 0:
            return None
        arrs = [np.empty(512, np.float32) for _ in range(5)]
        self._L.orc_block_voxels(self.h, bi, *[_ptr(a, f32p) for a in arrs])
        return arrs

    def load_block(self, key: int, tsdf, weight, mean_mode: int = 1) -> int:
        tsdf = np.ascontiguousarray(tsdf, np.float32)
        weight = np.ascontiguousarray(weight, np.float32)
        return self._L.orc_load_block(self.h, int(key), _ptr(tsdf, f32p), _ptr(weight, f32p),
                                      mean_mode)

    def count_mesh(self, min_weight=0.5, iso=0.0) -> int:
        return self._L.orc_extract_mesh(self.h, None, None, 0, min_weight, iso)

    def extract(self, min_weight=0.5, iso=0.0, cap=None, rgb=False):
        if cap is None:
            cap = max(self.count_mesh(min_weight, iso), 1)
        posnor = np.empty(cap * 6, np.float32)
        col = np.empty(cap * 3, np.uint8) if rgb else None
        n = self._L.orc_extract_mesh(self.h, _ptr(posnor, f32p), _ptr(col, u8p), cap,
                                     min_weight, iso)
        if n < 0:
            return None
        out = posnor[: n * 6].reshape(n, 6).copy()
        return (out, col[: n * 3].reshape(n, 3).copy()) if rgb else out


class ProjOracle:
    """FP64 projective-update oracle over touched block lists."""

    def __init__(self, voxel: float, trunc: float, w_max: int, cap_blocks: int, L=None):
        self._L = L if L is not None else lib()
        self.h = self._L.orc_proj_create(voxel, trunc, w_max, cap_blocks)
        if not self.h:
            raise ValueError("orc_proj_create failed")

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_proj_destroy(self.h)
            self.h = None

    def integrate(self, touched, depth, fx, fy, cx, cy, Rcw64, c64, Rcw32, c32) -> int:
        touched = np.ascontiguousarray(touched, np.int64)
        depth = np.ascontiguousarray(depth, np.float32)
        H, W = depth.shape
        Rcw64 = np.ascontiguousarray(Rcw64, np.float64).reshape(9)
        c64 = np.ascontiguousarray(c64, np.float64).reshape(3)
        Rcw32 = np.ascontiguousarray(Rcw32, np.float32).reshape(9)
        c32 = np.ascontiguousarray(c32, np.float32).reshape(3)
        return self._L.orc_proj_integrate(self.h, _ptr(touched, i64p), touched.size,
                                          _ptr(depth, f32p), W, H, fx, fy, cx, cy,
                                          _ptr(Rcw64, f64p), _ptr(c64, f64p), _ptr(Rcw32, f32p),
                                          _ptr(c32, f32p))

    def block(self, key: int):
        t = np.empty(512, np.float64)
        w = np.empty(512, np.int32)
        if self._L.orc_proj_block(self.h, int(key), _ptr(t, f64p), _ptr(w, i32p)):
            return None
        return t, w


# ---- the compiled reference CPU arm ----------------------------------------

_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    """The reference CPU arm's own osn_tsdf_* C ABI (cpp/include/osn_tsdf/c_api.h:14-46)."""
    global _ref
    if _ref is None:
        L = C.CDLL(REF_SO)
        L.osn_tsdf_create.restype = C.c_void_p
        L.osn_tsdf_create.argtypes = [C.c_float, C.c_int32, C.c_float, C.c_float]
        L.osn_tsdf_destroy.argtypes = [C.c_void_p]
        L.osn_tsdf_integrate.argtypes = [C.c_void_p, f32p, u8p, f32p, C.c_int32, C.c_int32,
                                         C.c_float, C.c_float, C.c_float, C.c_float]
        L.osn_tsdf_integrate.restype = C.c_int32
        L.osn_tsdf_allocate_blocks.argtypes = [C.c_void_p, f32p, C.c_int32, C.c_float,
                                               C.c_float, C.c_float, C.c_float]
        L.osn_tsdf_allocate_blocks.restype = C.c_int32
        L.osn_tsdf_update_voxels.argtypes = [C.c_void_p, f32p, u8p, f32p, C.c_int32, C.c_int32,
                                             C.c_float, C.c_float, C.c_float, C.c_float]
        L.osn_tsdf_update_voxels.restype = C.c_int32
        L.osn_tsdf_extract_mesh.argtypes = [C.c_void_p, f32p, u8p, C.c_int32, C.c_float,
                                            C.c_float]
        L.osn_tsdf_extract_mesh.restype = C.c_int32
        L.osn_tsdf_block_count.argtypes = [C.c_void_p]
        L.osn_tsdf_block_count.restype = C.c_int32
        L.osn_tsdf_drop_count.argtypes = [C.c_void_p]
        L.osn_tsdf_drop_count.restype = C.c_uint64
        L.osn_tsdf_set_thread_count.argtypes = [C.c_void_p, C.c_int32]
        L.osn_tsdf_thread_count.argtypes = [C.c_void_p]
        L.osn_tsdf_thread_count.restype = C.c_int32
        _ref = L
    return _ref


class RefVolume:
    """The compiled reference CPU arm, driven through its own C ABI."""

    def __init__(self, voxel: float, pool: int, trunc: float = -1.0, weight_cap: float = 32.0,
                 threads: int = 0):
        self._L = ref_lib()
        self.h = self._L.osn_tsdf_create(voxel, pool, trunc, weight_cap)
        if not self.h:
            raise ValueError("osn_tsdf_create failed")
        if threads:
            self._L.osn_tsdf_set_thread_count(self.h, threads)

    def __del__(self):
        if getattr(self, "h", None):
            self._L.osn_tsdf_destroy(self.h)
            self.h = None

    @property
    def threads(self) -> int:
        return self._L.osn_tsdf_thread_count(self.h)

    def allocate(self, pts, cam, radius=0.0) -> int:
        pts = np.ascontiguousarray(pts, np.float32)
        return self._L.osn_tsdf_allocate_blocks(self.h, _ptr(pts, f32p), pts.size // 3,
                                                *map(float, cam), radius)

    def update(self, pts, cam, radius=0.0) -> int:
        pts = np.ascontiguousarray(pts, np.float32)
        return self._L.osn_tsdf_update_voxels(self.h, _ptr(pts, f32p), None, None, pts.size // 3,
                                              0, *map(float, cam), radius)

    def integrate(self, pts, cam, radius=0.0) -> int:
        pts = np.ascontiguousarray(pts, np.float32)
        return self._L.osn_tsdf_integrate(self.h, _ptr(pts, f32p), None, None, pts.size // 3, 0,
                                          *map(float, cam), radius)

    @property
    def block_count(self) -> int:
        return self._L.osn_tsdf_block_count(self.h)

    @property
    def drop_count(self) -> int:
        return self._L.osn_tsdf_drop_count(self.h)

    def extract(self, min_weight=0.5, iso=0.0, cap=4 << 20):
        posnor = np.empty(cap * 6, np.float32)
        n = self._L.osn_tsdf_extract_mesh(self.h, _ptr(posnor, f32p), None, cap, min_weight, iso)
        if n < 0:
            return None
        return posnor[: n * 6].reshape(n, 6).copy()
