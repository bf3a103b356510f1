"""Python face of the mesh-comparison gate (include/osn_mesh_metrics.h).

Same functions and defaults as the reference's mesh-metrics crate
(crates/mesh-metrics/src/lib.rs): sample_surface, compare, gate and
Tolerance.for_voxel. Meshes are float32 arrays of shape (n_vertices, k) with
x, y, z in the first three columns (k = 3, or 6 for position + normal records);
3 consecutive vertices make a triangle.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

import osn_tsdf_hip as _H


class _Stats(C.Structure):
    _fields_ = [("mean", C.c_double), ("rms", C.c_double), ("p95", C.c_double),
                ("p99", C.c_double), ("max", C.c_double), ("n", C.c_int64)]


class _Agreement(C.Structure):
    _fields_ = [("a_to_b", _Stats), ("b_to_a", _Stats), ("mean_surface", C.c_double),
                ("hausdorff", C.c_double), ("hausdorff_p99", C.c_double)]


class _Tol(C.Structure):
    _fields_ = [("mean_surface_m", C.c_double), ("hausdorff_p99_m", C.c_double),
                ("vertex_count_rel", C.c_double)]


class _Gate(C.Structure):
    _fields_ = [("agreement", _Agreement), ("n_vertices_a", C.c_int64),
                ("n_vertices_b", C.c_int64), ("vertex_count_rel", C.c_double),
                ("passed", C.c_int32), ("failures", C.c_int32)]


FAIL_MEAN, FAIL_P99, FAIL_COUNT = 1, 2, 4


def _lib():
    L = _H.lib()
    if not getattr(L, "_mesh_bound", False):
        P, I32, I64, U64, D = C.c_void_p, C.c_int32, C.c_int64, C.c_uint64, C.c_double
        L.osn_mesh_tolerance_for_voxel.restype = _Tol
        L.osn_mesh_tolerance_for_voxel.argtypes = [D]
        L.osn_mesh_sample_surface.restype = I32
        L.osn_mesh_sample_surface.argtypes = [P, I64, I32, I64, U64, P, I64, P]
        L.osn_mesh_nearest_sq.restype = I32
        L.osn_mesh_nearest_sq.argtypes = [P, I64, I32, P, I64, D, I32, P]
        L.osn_mesh_compare.restype = I32
        L.osn_mesh_compare.argtypes = [P, I64, I32, P, I64, I32, I64, D, U64, I32,
                                       C.POINTER(_Agreement)]
        L.osn_mesh_gate.restype = I32
        L.osn_mesh_gate.argtypes = [P, I64, I32, P, I64, I32, _Tol, I64, D, I32,
                                    C.POINTER(_Gate)]
        L._mesh_bound = True
    return L


def _mesh(m):
    a = np.ascontiguousarray(m, dtype=np.float32)
    if a.ndim == 1:
        a = a.reshape(-1, 3)
    if a.shape[1] < 3 or a.shape[0] % 3:
        raise ValueError("triangle soup of (3 * n_triangles, >= 3) float32 expected")
    return a, a.shape[0], a.shape[1]


@dataclass
class Stats:
    mean: float
    rms: float
    p95: float
    p99: float
    max: float
    n: int


@dataclass
class Agreement:
    a_to_b: Stats
    b_to_a: Stats
    mean_surface: float
    hausdorff: float
    hausdorff_p99: float


@dataclass
class Tolerance:
    mean_surface_m: float
    hausdorff_p99_m: float
    vertex_count_rel: float

    @staticmethod
    def for_voxel(voxel_m: float) -> "Tolerance":
        t = _lib().osn_mesh_tolerance_for_voxel(float(voxel_m))
        return Tolerance(t.mean_surface_m, t.hausdorff_p99_m, t.vertex_count_rel)


@dataclass
class GateResult:
    agreement: Agreement
    n_vertices_a: int
    n_vertices_b: int
    vertex_count_rel: float
    passed: bool
    failures: int


def _stats(s):
    return Stats(s.mean, s.rms, s.p95, s.p99, s.max, s.n)


def _agreement(a):
    return Agreement(_stats(a.a_to_b), _stats(a.b_to_a), a.mean_surface, a.hausdorff,
                     a.hausdorff_p99)


def sample_surface(mesh, n_samples: int, seed: int) -> np.ndarray:
    a, nv, k = _mesh(mesh)
    cap = max(int(n_samples), nv)
    out = np.empty((cap, 3), np.float64)
    n = C.c_int64(0)
    if _lib().osn_mesh_sample_surface(a.ctypes.data, nv, k, int(n_samples), seed & (2**64 - 1),
                                      out.ctypes.data, cap, C.addressof(n)):
        raise ValueError("osn_mesh_sample_surface: bad arguments")
    return out[: n.value]


def nearest_sq(mesh, queries, cell: float, threads: int = 0) -> np.ndarray:
    a, nv, k = _mesh(mesh)
    q = np.ascontiguousarray(queries, dtype=np.float64).reshape(-1, 3)
    out = np.empty(len(q), np.float64)
    if _lib().osn_mesh_nearest_sq(a.ctypes.data, nv, k, q.ctypes.data, len(q), float(cell),
                                  threads, out.ctypes.data):
        raise ValueError("osn_mesh_nearest_sq: bad arguments")
    return out


def compare(a, b, n_samples: int, cell: float, seed: int, threads: int = 0) -> Agreement:
    A, na, ka = _mesh(a)
    B, nb, kb = _mesh(b)
    r = _Agreement()
    if _lib().osn_mesh_compare(A.ctypes.data, na, ka, B.ctypes.data, nb, kb, int(n_samples),
                               float(cell), seed & (2**64 - 1), threads, C.byref(r)):
        raise ValueError("osn_mesh_compare: bad arguments")
    return _agreement(r)


def gate(a, b, tol: Tolerance, n_samples: int, cell: float, threads: int = 0) -> GateResult:
    A, na, ka = _mesh(a)
    B, nb, kb = _mesh(b)
    r = _Gate()
    t = _Tol(tol.mean_surface_m, tol.hausdorff_p99_m, tol.vertex_count_rel)
    if _lib().osn_mesh_gate(A.ctypes.data, na, ka, B.ctypes.data, nb, kb, t, int(n_samples),
                            float(cell), threads, C.byref(r)):
        raise ValueError("osn_mesh_gate: bad arguments")
    return GateResult(_agreement(r.agreement), r.n_vertices_a, r.n_vertices_b,
                      r.vertex_count_rel, bool(r.passed), r.failures)
