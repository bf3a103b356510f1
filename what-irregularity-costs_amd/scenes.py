"""Seeded synthetic scenes and camera trajectories (host side, FP64).

Scenes are an axis-aligned room seen from inside plus spheres and boxes; the
library's renderer (osn_scene_render_depth) ray-casts them into planar depth
frames on the GPU. Trajectories are composed in FP64 on the host; a frame
descriptor casts each pose to FP32 once (osn_tsdf_hip.make_frame).

Configurations (sizes as the workload descriptions state them):
  small   -- 4x4x3 m room, 8 spheres r ~ U[0.1, 0.5] m, 640x480, 2 cm voxels,
             trunc 4 voxels, table 2^18, 20 frames;
  precise -- 6x6x3 m room, 20 spheres, 10 boxes with extents U[0.2, 1.0] m,
             640x480, 1 cm voxels, trunc 4 cm, table 2^22, 500 frames, depth
             noise N(0, 0.0012 z^2) and 1% dropouts;
  large   -- 12x12x4 m room, 60 mixed primitives, 1280x960, 5 mm voxels,
             trunc 2 cm, table 2^25, 1000 frames, same noise model;
  batched -- 24 independent scenes of uneven size (batched_workload).
Intrinsics: fx = fy = W/2, cx = W/2, cy = H/2 (fx = 320 at 640x480).
The camera moves 0.02 m and turns 1 degree of yaw per frame along a circle
inside the room, with a slow +-10 degree pitch oscillation.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

MAX_PRIMS = 64


class OsnScene(C.Structure):
    _fields_ = [("room_lo", C.c_float * 3), ("room_hi", C.c_float * 3),
                ("n_spheres", C.c_int32), ("n_boxes", C.c_int32),
                ("spheres", (C.c_float * 4) * MAX_PRIMS), ("boxes", (C.c_float * 6) * MAX_PRIMS)]


@dataclass
class Scene:
    room_lo: np.ndarray
    room_hi: np.ndarray
    spheres: np.ndarray = field(default_factory=lambda: np.zeros((0, 4)))
    boxes: np.ndarray = field(default_factory=lambda: np.zeros((0, 6)))
    path_centre: np.ndarray = None
    path_radius: float = 1.0

    def to_c(self) -> OsnScene:
        s = OsnScene()
        s.room_lo[:] = [float(x) for x in self.room_lo]
        s.room_hi[:] = [float(x) for x in self.room_hi]
        s.n_spheres = len(self.spheres)
        s.n_boxes = len(self.boxes)
        for i, sp in enumerate(self.spheres):
            s.spheres[i][:] = [float(x) for x in sp]
        for i, b in enumerate(self.boxes):
            s.boxes[i][:] = [float(x) for x in b]
        return s


def _dist_to_path(p, centre, radius):
    """Distance from p to the horizontal camera circle."""
    d = p[:2] - centre[:2]
    rho = math.hypot(d[0], d[1])
    return math.hypot(rho - radius, p[2] - centre[2])


def make_scene(seed: int, room=(4.0, 4.0, 3.0), n_spheres=8, n_boxes=0, r_range=(0.1, 0.5),
               box_range=(0.2, 1.0), clearance=0.35) -> Scene:
    rng = np.random.default_rng(seed)
    lo = np.array([-room[0] / 2, -room[1] / 2, 0.0])
    hi = np.array([room[0] / 2, room[1] / 2, room[2]])
    centre = np.array([0.0, 0.0, room[2] / 2])
    radius = 0.25 * min(room[0], room[1])
    spheres, boxes = [], []
    tries = 0
    while len(spheres) < n_spheres and tries < 100000:
        tries += 1
        r = rng.uniform(*r_range)
        m = r + 0.2
        c = rng.uniform(lo + m, hi - m)
        if _dist_to_path(c, centre, radius) < r + clearance:
            continue
        spheres.append([c[0], c[1], c[2], r])
    tries = 0
    while len(boxes) < n_boxes and tries < 100000:
        tries += 1
        ext = rng.uniform(box_range[0], box_range[1], size=3)
        c = rng.uniform(lo + ext / 2 + 0.05, hi - ext / 2 - 0.05)
        if _dist_to_path(c, centre, radius) < 0.5 * np.linalg.norm(ext) + clearance:
            continue
        boxes.append(list(c - ext / 2) + list(c + ext / 2))
    return Scene(lo, hi, np.array(spheres, np.float64).reshape(-1, 4),
                 np.array(boxes, np.float64).reshape(-1, 6), centre, radius)


def trajectory(seed: int, n_frames: int, scene: Scene, step_m=0.02, yaw_deg=1.0):
    """FP64 world-from-camera poses (R_wc 3x3, t_wc 3) along the scene's path.
    Camera axes: x right, y down, z forward; world z up."""
    rng = np.random.default_rng(seed + 7919)
    phi0 = rng.uniform(0, 2 * np.pi)
    psi0 = rng.uniform(0, 2 * np.pi)
    omega = step_m / scene.path_radius
    poses = []
    for k in range(n_frames):
        phi = phi0 + k * omega
        t = scene.path_centre + scene.path_radius * np.array([math.cos(phi), math.sin(phi), 0.0])
        psi = psi0 + math.radians(yaw_deg) * k
        theta = math.radians(10.0) * math.sin(2 * math.pi * k / 200.0)
        fwd = np.array([math.cos(theta) * math.cos(psi), math.cos(theta) * math.sin(psi),
                        math.sin(theta)])
        right = np.array([fwd[1], -fwd[0], 0.0])
        right /= np.linalg.norm(right)
        down = np.cross(fwd, right)
        R = np.stack([right, down, fwd], axis=1)
        poses.append((R, t))
    return poses


def intrinsics(width: int, height: int):
    return width / 2.0, width / 2.0, width / 2.0, height / 2.0


SMALL = dict(room=(4.0, 4.0, 3.0), n_spheres=8, n_boxes=0, width=640, height=480, voxel=0.02,
             trunc=0.08, table=1 << 18, frames=20, noise_k=0.0, dropout=0.0)
PRECISE = dict(room=(6.0, 6.0, 3.0), n_spheres=20, n_boxes=10, width=640, height=480,
               voxel=0.01, trunc=0.04, table=1 << 22, frames=500, noise_k=0.0012, dropout=0.01)
LARGE = dict(room=(12.0, 12.0, 4.0), n_spheres=30, n_boxes=30, width=1280, height=960,
             voxel=0.005, trunc=0.02, table=1 << 25, frames=1000, noise_k=0.0012, dropout=0.01)


BATCH_SIZES = ((320, 240), (640, 480), (1280, 720))


def batched_workload(seed: int, n_scenes: int = 24, frame_scale: float = 1.0) -> list:
    """The batched workload: independent scenes of uneven size. Per scene, an
    image size drawn from 320x240 / 640x480 / 1280x720 (intrinsics as above), a
    frame count U{50..400} (times frame_scale, at least 1), a table capacity
    2^U{16..22} slots (pool half of it), a room with sides U[4, 6] m and height
    3 m (between the small and precise rooms), U{8..20} spheres and U{0..10}
    boxes; 1 cm voxels, trunc 4 cm, the precise noise model. Returns one
    configuration dict per scene, with its scene seed and its unscaled
    frame count."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenes):
        W, H = BATCH_SIZES[int(rng.integers(len(BATCH_SIZES)))]
        frames = int(rng.integers(50, 401))
        table = 1 << int(rng.integers(16, 23))
        side = float(rng.uniform(4.0, 6.0))
        out.append(dict(room=(side, side, 3.0), n_spheres=int(rng.integers(8, 21)),
                        n_boxes=int(rng.integers(0, 11)), width=W, height=H, voxel=0.01,
                        trunc=0.04, table=table, frames=max(1, int(round(frames * frame_scale))),
                        frames_full=frames, noise_k=0.0012, dropout=0.01,
                        seed=int(seed) * 1000 + s))
    return out


def build(cfg: dict, seed: int):
    scene = make_scene(seed, room=cfg["room"], n_spheres=cfg["n_spheres"], n_boxes=cfg["n_boxes"])
    poses = trajectory(seed, cfg["frames"], scene)
    return scene, poses


_render = None


def render_depth(scene_c: OsnScene, frame, noise_k: float, dropout: float, seed: int,
                 index: int, d_depth, stream: int | None = None) -> int:
    """Render into a device tensor (float32, H*W) with the library's renderer."""
    global _render
    if _render is None:
        import osn_tsdf_hip as H
        L = H.lib()
        fn = L.osn_scene_render_depth
        fn.restype = C.c_int32
        fn.argtypes = [C.POINTER(OsnScene), C.POINTER(H.Frame), C.c_float, C.c_float,
                       C.c_uint64, C.c_uint32, C.c_void_p, C.c_void_p]
        _render = fn
    return _render(C.byref(scene_c), C.byref(frame), noise_k, dropout, seed, index,
                   d_depth.data_ptr(), stream)
