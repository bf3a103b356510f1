"""Python host binding of the gfx950 TSDF library (libosn_tsdf_hip.so).

A thin ctypes layer over the C ABI in include/osn_tsdf_hip.h. It mirrors the
reference's volume interface -- create(voxel, pool, trunc, weight_cap),
allocate / update / integrate over device points with a camera centre and far
radius, extract_mesh into host buffers, block_count / drop_count getters, reset
-- with the same argument meaning and return codes (reference
cpp/include/osn_tsdf/c_api.h:80-126, cpp/src/c_api_cuda.cu:151-246), plus the
depth-frame, projective-storage and measurement entry points.

Device buffers are passed as torch tensors on the GPU (PyTorch supplies device
memory only; every kernel is the library's own). The library is required: if it
is not built, loading raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libosn_tsdf_hip.so")

STORE_SUMS, STORE_F32, STORE_F16, STORE_F64 = 0, 1, 2, 3
HASH_REFERENCE, HASH_MIX = 0, 1
STAGES = ("allocate", "assign", "update", "extract")


class DeviceView(C.Structure):
    """Layout of OsnTsdfDeviceView (reference c_api.h:61-78), 104 bytes."""
    _fields_ = [("table", C.c_void_p), ("block_count", C.c_void_p), ("drop_count", C.c_void_p),
                ("block_coord", C.c_void_p), ("tsdf", C.c_void_p), ("weight", C.c_void_p),
                ("r", C.c_void_p), ("g", C.c_void_p), ("b", C.c_void_p),
                ("hash_mask", C.c_uint32), ("scratch_base", C.c_uint32),
                ("pool_capacity", C.c_int32), ("block_dim", C.c_int32),
                ("voxel_size_m", C.c_float), ("trunc_m", C.c_float), ("weight_cap", C.c_float)]


class Config(C.Structure):
    _fields_ = [("voxel_size_m", C.c_float), ("pool_capacity_blocks", C.c_int32),
                ("trunc_m", C.c_float), ("weight_cap", C.c_float), ("storage", C.c_int32),
                ("table_slots", C.c_uint32), ("hash_kind", C.c_int32),
                ("probe_limit", C.c_uint32), ("lds_set", C.c_int32),
                ("update_threads", C.c_int32), ("update_grid", C.c_int32),
                ("alloc_grid", C.c_int32)]


class Frame(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("fx", C.c_float),
                ("fy", C.c_float), ("cx", C.c_float), ("cy", C.c_float),
                ("R_wc", C.c_float * 9), ("t_wc", C.c_float * 3),
                ("R_wc64", C.c_double * 9), ("t_wc64", C.c_double * 3),
                ("max_range_m", C.c_float), ("radius_m", C.c_float)]


def make_frame(width, height, fx, fy, cx, cy, R_wc64, t_wc64, max_range=0.0, radius=0.0) -> Frame:
    """A frame descriptor from an FP64 world-from-camera pose (cast once to FP32)."""
    R64 = np.asarray(R_wc64, np.float64).reshape(9)
    t64 = np.asarray(t_wc64, np.float64).reshape(3)
    f = Frame()
    f.width, f.height = int(width), int(height)
    f.fx, f.fy, f.cx, f.cy = float(fx), float(fy), float(cx), float(cy)
    f.R_wc[:] = [float(np.float32(x)) for x in R64]
    f.t_wc[:] = [float(np.float32(x)) for x in t64]
    f.R_wc64[:] = [float(x) for x in R64]
    f.t_wc64[:] = [float(x) for x in t64]
    f.max_range_m = float(max_range)
    f.radius_m = float(radius)
    return f


_lib = None
_VP = C.c_void_p
_F = C.c_float
_I = C.c_int32


def lib():
    """Load the library; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it (what-irregularity-costs_amd/build.py) "
                           "before running the GPU path")
    L = C.CDLL(LIB_PATH)
    sig = {
        "osn_tsdf_hip_create": (_VP, [_F, _I, _F, _F]),
        "osn_tsdf_hip_create_ex": (_VP, [C.POINTER(Config)]),
        "osn_tsdf_hip_destroy": (None, [_VP]),
        "osn_tsdf_hip_valid": (_I, [_VP]),
        "osn_tsdf_hip_device_view": (_I, [_VP, C.POINTER(DeviceView)]),
        "osn_tsdf_hip_allocate_blocks": (_I, [_VP, _VP, _I, _F, _F, _F, _F]),
        "osn_tsdf_hip_update_voxels": (_I, [_VP, _VP, _I, _F, _F, _F, _F]),
        "osn_tsdf_hip_allocate_counting_cas": (_I, [_VP, _VP, _I, _F, _F, _F, _F]),
        "osn_tsdf_hip_allocate_no_cas": (_I, [_VP, _VP, _I, _F, _F, _F, _F]),
        "osn_tsdf_hip_allocate_slot_index": (_I, [_VP, _VP, _I, _F, _F, _F, _F]),
        "osn_tsdf_hip_allocate_probe_depth": (_I, [_VP, _VP, _I, _F, _F, _F, _F, _VP, _VP, _VP]),
        "osn_tsdf_hip_extract_mesh": (_I, [_VP, _VP, _VP, _I, _F, _F]),
        "osn_tsdf_hip_reset": (_I, [_VP]),
        "osn_tsdf_hip_block_count": (_I, [_VP]),
        "osn_tsdf_hip_drop_count": (C.c_uint64, [_VP]),
        "osn_tsdf_hip_synchronize": (None, [_VP]),
        "osn_tsdf_hip_allocate_depth": (_I, [_VP, _VP, C.POINTER(Frame)]),
        "osn_tsdf_hip_update_depth": (_I, [_VP, _VP, C.POINTER(Frame)]),
        "osn_tsdf_hip_integrate_depth": (_I, [_VP, _VP, C.POINTER(Frame)]),
        "osn_tsdf_hip_backproject": (_I, [_VP, C.POINTER(Frame), _VP]),
        "osn_tsdf_hip_insert_blocks": (_I, [_VP, _VP, _I]),
        "osn_tsdf_hip_touched_count": (_I, [_VP]),
        "osn_tsdf_hip_block_keys": (_I, [_VP, _VP, _I]),
        "osn_tsdf_hip_read_voxels": (_I, [_VP, _I, _I, _VP, _VP]),
        "osn_tsdf_hip_read_voxels_f64": (_I, [_VP, _I, _I, _VP, _VP]),
        "osn_tsdf_hip_extract_mesh_device": (_I, [_VP, _VP, _VP, _I, _F, _F, _VP]),
        "osn_tsdf_hip_extract_mesh_device_noalloc": (_I, [_VP, _VP, _VP, _I, _F, _F, _VP]),
        "osn_tsdf_hip_probe_stats": (_I, [_VP, _VP, _VP, _VP, _VP]),
        "osn_tsdf_hip_set_timing": (_I, [_VP, _I]),
        "osn_tsdf_hip_read_timing": (_I, [_VP, _VP, _VP]),
        "osn_tsdf_hip_stream": (_VP, [_VP]),
        "osn_tsdf_hip_batch_create": (_VP, [_VP, _I]),
        "osn_tsdf_hip_batch_destroy": (None, [_VP]),
        "osn_tsdf_hip_batch_set_schedule": (_I, [_VP, _I, _VP, _VP]),
        "osn_tsdf_hip_batch_run": (_I, [_VP, _I, _I]),
        "osn_tsdf_hip_batch_synchronize": (None, [_VP]),
        "osn_tsdf_hip_batch_tune": (_I, [_VP, _I, _I, _I, _I]),
        "osn_tsdf_hip_batch_set_timing": (_I, [_VP, _I]),
        "osn_tsdf_hip_batch_read_timing": (_I, [_VP, _VP, _VP]),
        "osn_tsdf_hip_batch_set_extract": (_I, [_VP, _VP, _VP, _VP, _F, _F, _I]),
        "osn_tsdf_hip_batch_set_alloc_mode": (_I, [_VP, _I]),
        "osn_tsdf_hip_insert_blocks_mode": (_I, [_VP, _VP, _I, _I]),
        "osn_tsdf_hip_cas_attempts": (_I, [_VP, _VP]),
        "osn_tsdf_hip_cas_ceiling": (_I, [_I, _I, _VP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _dptr(t) -> int:
    """Device pointer of a torch tensor (or an int already)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    if not t.is_cuda:
        raise ValueError("device buffer expected")
    if not t.is_contiguous():
        raise ValueError("contiguous device buffer expected")
    return t.data_ptr()


def _hptr(a: np.ndarray) -> int:
    return a.ctypes.data


class HipVolume:
    """A TSDF volume on the GPU.

    HipVolume(voxel, pool, trunc, weight_cap) is the reference-shaped volume
    (weighted-sum storage, reference hash, table = next_pow2(2 * pool)); pass
    storage=/table_slots=/hash_kind=/... for the projective variants.
    """

    def __init__(self, voxel: float, pool: int, trunc: float = -1.0, weight_cap: float = 32.0, *,
                 storage: int | None = None, table_slots: int = 0, hash_kind: int = HASH_REFERENCE,
                 probe_limit: int = 0, lds_set: int = -1, update_threads: int = 0,
                 update_grid: int = 0, alloc_grid: int = 0):
        self._L = lib()
        if storage is None:
            self.h = self._L.osn_tsdf_hip_create(voxel, pool, trunc, weight_cap)
            self.storage = STORE_SUMS
        else:
            cfg = Config(voxel, pool, trunc, weight_cap, storage, table_slots, hash_kind,
                         probe_limit, lds_set, update_threads, update_grid, alloc_grid)
            self.h = self._L.osn_tsdf_hip_create_ex(C.byref(cfg))
            self.storage = storage
        if not self.h:
            raise RuntimeError("osn_tsdf_hip_create failed")
        self.voxel = voxel
        self.pool = pool

    def close(self):
        if getattr(self, "h", None):
            self._L.osn_tsdf_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # ---- reference-shaped operations ----
    def allocate(self, d_pos, cam, radius=0.0) -> int:
        n = 0 if d_pos is None else d_pos.numel() // 3
        return self._L.osn_tsdf_hip_allocate_blocks(self.h, _dptr(d_pos), n, *map(float, cam),
                                                    float(radius))

    def update(self, d_pos, cam, radius=0.0) -> int:
        n = 0 if d_pos is None else d_pos.numel() // 3
        return self._L.osn_tsdf_hip_update_voxels(self.h, _dptr(d_pos), n, *map(float, cam),
                                                  float(radius))

    def integrate(self, d_pos, cam, radius=0.0) -> int:
        return self.allocate(d_pos, cam, radius) or self.update(d_pos, cam, radius)

    def allocate_counting_cas(self, d_pos, cam, radius=0.0) -> int:
        return self._L.osn_tsdf_hip_allocate_counting_cas(self.h, _dptr(d_pos), d_pos.numel() // 3,
                                                          *map(float, cam), float(radius))

    def allocate_no_cas(self, d_pos, cam, radius=0.0) -> int:
        return self._L.osn_tsdf_hip_allocate_no_cas(self.h, _dptr(d_pos), d_pos.numel() // 3,
                                                    *map(float, cam), float(radius))

    def allocate_probe_depth(self, d_pos, cam, radius=0.0):
        lane = C.c_uint64(0)
        wave = C.c_uint64(0)
        hist = np.zeros(64, np.uint64)
        rc = self._L.osn_tsdf_hip_allocate_probe_depth(
            self.h, _dptr(d_pos), d_pos.numel() // 3, *map(float, cam), float(radius),
            C.addressof(lane), C.addressof(wave), _hptr(hist))
        return rc, lane.value, wave.value, hist

    def extract(self, min_weight=0.5, iso=0.0, cap=4 << 20, rgb=False):
        posnor = np.empty(cap * 6, np.float32)
        col = np.empty(cap * 3, np.uint8) if rgb else None
        n = self._L.osn_tsdf_hip_extract_mesh(self.h, _hptr(posnor),
                                              _hptr(col) if rgb else None, cap, min_weight, iso)
        if n < 0:
            return None
        out = posnor[: n * 6].reshape(n, 6)
        return (out, col[: n * 3].reshape(n, 3)) if rgb else out

    def extract_device(self, d_posnor, d_rgb, cap, min_weight, iso, d_count) -> int:
        return self._L.osn_tsdf_hip_extract_mesh_device(self.h, _dptr(d_posnor), _dptr(d_rgb),
                                                        cap, min_weight, iso, _dptr(d_count))

    def extract_device_noalloc(self, d_posnor, d_rgb, cap, min_weight, iso, d_count) -> int:
        """extract_device with the library's static scratch: no device allocation."""
        return self._L.osn_tsdf_hip_extract_mesh_device_noalloc(
            self.h, _dptr(d_posnor), _dptr(d_rgb), cap, min_weight, iso, _dptr(d_count))

    def reset(self) -> int:
        return self._L.osn_tsdf_hip_reset(self.h)

    @property
    def block_count(self) -> int:
        return self._L.osn_tsdf_hip_block_count(self.h)

    @property
    def drop_count(self) -> int:
        return self._L.osn_tsdf_hip_drop_count(self.h)

    @property
    def valid(self) -> bool:
        return bool(self._L.osn_tsdf_hip_valid(self.h))

    def synchronize(self):
        self._L.osn_tsdf_hip_synchronize(self.h)

    def device_view(self):
        v = DeviceView()
        rc = self._L.osn_tsdf_hip_device_view(self.h, C.byref(v))
        return rc, v

    # ---- depth frames ----
    def allocate_depth(self, d_depth, frame: Frame) -> int:
        return self._L.osn_tsdf_hip_allocate_depth(self.h, _dptr(d_depth), C.byref(frame))

    def update_depth(self, d_depth, frame: Frame) -> int:
        return self._L.osn_tsdf_hip_update_depth(self.h, _dptr(d_depth), C.byref(frame))

    def integrate_depth(self, d_depth, frame: Frame) -> int:
        return self._L.osn_tsdf_hip_integrate_depth(self.h, _dptr(d_depth), C.byref(frame))

    def insert_blocks(self, d_coords) -> int:
        return self._L.osn_tsdf_hip_insert_blocks(self.h, _dptr(d_coords), d_coords.numel() // 3)

    @property
    def touched_count(self) -> int:
        return self._L.osn_tsdf_hip_touched_count(self.h)

    def keys(self) -> np.ndarray:
        n = self.block_count
        out = np.empty(max(n, 1), np.int64)
        got = self._L.osn_tsdf_hip_block_keys(self.h, _hptr(out), n)
        if got < 0:
            raise RuntimeError("block_keys failed")
        return out[:got].copy()

    def read_voxels(self, first=0, n=None):
        if n is None:
            n = self.block_count
        t = np.empty(max(n, 1) * 512, np.float32)
        w = np.empty(max(n, 1) * 512, np.float32)
        if n == 0:
            return t[:0], w[:0]
        rc = self._L.osn_tsdf_hip_read_voxels(self.h, first, n, _hptr(t), _hptr(w))
        if rc:
            raise RuntimeError(f"read_voxels rc={rc}")
        return t.reshape(n, 512), w.reshape(n, 512)

    def read_voxels_f64(self, first=0, n=None):
        if n is None:
            n = self.block_count
        t = np.empty(n * 512, np.float64)
        w = np.empty(n * 512, np.float64)
        rc = self._L.osn_tsdf_hip_read_voxels_f64(self.h, first, n, _hptr(t), _hptr(w))
        if rc:
            raise RuntimeError(f"read_voxels_f64 rc={rc}")
        return t.reshape(n, 512), w.reshape(n, 512)

    def probe_stats(self):
        hist = np.zeros(64, np.uint64)
        lane = C.c_uint64(0)
        wave = C.c_uint64(0)
        look = C.c_uint64(0)
        self._L.osn_tsdf_hip_probe_stats(self.h, _hptr(hist), C.addressof(lane),
                                         C.addressof(wave), C.addressof(look))
        return hist, lane.value, wave.value, look.value

    def set_timing(self, on: bool):
        self._L.osn_tsdf_hip_set_timing(self.h, 1 if on else 0)

    def read_timing(self):
        ms = np.zeros(4, np.float64)
        calls = np.zeros(4, np.int64)
        self._L.osn_tsdf_hip_read_timing(self.h, _hptr(ms), _hptr(calls))
        return ms, calls

    @property
    def stream(self) -> int:
        return self._L.osn_tsdf_hip_stream(self.h)


def backproject(d_depth, frame: Frame, d_out) -> int:
    return lib().osn_tsdf_hip_backproject(_dptr(d_depth), C.byref(frame), _dptr(d_out))


class Batch:
    """Many projective volumes integrated together (one launch per stage per
    step, optionally replayed from a captured hipGraph)."""

    def __init__(self, volumes):
        self._L = lib()
        self.volumes = list(volumes)
        arr = (C.c_void_p * len(self.volumes))(*[v.h for v in self.volumes])
        self.h = self._L.osn_tsdf_hip_batch_create(arr, len(self.volumes))
        if not self.h:
            raise RuntimeError("osn_tsdf_hip_batch_create failed")

    def close(self):
        if getattr(self, "h", None):
            self._L.osn_tsdf_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_schedule(self, depths, frames) -> int:
        """depths[k][s]: device tensor or None; frames[k][s]: Frame."""
        n_steps = len(depths)
        n = len(self.volumes)
        ptrs = (C.c_void_p * (n_steps * n))()
        frs = (Frame * (n_steps * n))()
        for k in range(n_steps):
            for s in range(n):
                d = depths[k][s]
                ptrs[k * n + s] = None if d is None else _dptr(d)
                if d is not None:
                    frs[k * n + s] = frames[k][s]
        return self._L.osn_tsdf_hip_batch_set_schedule(self.h, n_steps, ptrs, frs)

    def run(self, n_steps: int, use_graph: bool = False) -> int:
        return self._L.osn_tsdf_hip_batch_run(self.h, n_steps, 1 if use_graph else 0)

    def synchronize(self):
        self._L.osn_tsdf_hip_batch_synchronize(self.h)

    def tune(self, alloc_grid=-1, update_grid=-1, update_threads=0, lds_set=-1) -> int:
        return self._L.osn_tsdf_hip_batch_tune(self.h, alloc_grid, update_grid, update_threads,
                                               lds_set)

    def set_timing(self, on: bool):
        self._L.osn_tsdf_hip_batch_set_timing(self.h, 1 if on else 0)

    def set_alloc_mode(self, mode: int) -> int:
        """0 normal, 1 counting CAS attempts, 2 probe-depth statistics."""
        return self._L.osn_tsdf_hip_batch_set_alloc_mode(self.h, int(mode))

    def read_timing(self):
        ms = np.zeros(4, np.float64)
        calls = np.zeros(4, np.int64)
        self._L.osn_tsdf_hip_batch_read_timing(self.h, _hptr(ms), _hptr(calls))
        return ms, calls

    def set_extract(self, d_posnor, vert_caps, d_counts, min_weight=0.5, iso=0.0,
                    every=1) -> int:
        """Mesh every scene after every `every`-th step into d_posnor[s]
        (device float tensors, 6 per vertex); d_counts is a device int32 tensor
        with one entry per scene (vertex count, -1 on overflow). every=0 stops."""
        n = len(self.volumes)
        if every == 0:
            return self._L.osn_tsdf_hip_batch_set_extract(self.h, None, None, None, 0.0, 0.0, 0)
        self._ex_keep = (list(d_posnor), d_counts)  # the buffers must outlive their use
        ptrs = (C.c_void_p * n)(*[_dptr(t) for t in d_posnor])
        caps = (C.c_int32 * n)(*[int(c) for c in vert_caps])
        return self._L.osn_tsdf_hip_batch_set_extract(self.h, ptrs, caps, _dptr(d_counts),
                                                     float(min_weight), float(iso), int(every))


def insert_blocks_mode(vol: HipVolume, d_coords, mode: int) -> int:
    return lib().osn_tsdf_hip_insert_blocks_mode(vol.h, _dptr(d_coords), d_coords.numel() // 3, mode)


def cas_attempts(vol: HipVolume) -> int:
    out = C.c_uint64(0)
    if lib().osn_tsdf_hip_cas_attempts(vol.h, C.addressof(out)):
        raise RuntimeError("cas_attempts failed")
    return out.value


def cas_ceiling(log2_slots: int, reps: int = 5) -> float:
    """Milliseconds for 2^log2_slots uncontended, scattered 64-bit CASes."""
    ms = C.c_double(0.0)
    if lib().osn_tsdf_hip_cas_ceiling(log2_slots, reps, C.addressof(ms)):
        raise RuntimeError("cas_ceiling failed")
    return ms.value
