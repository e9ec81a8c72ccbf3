"""Benchmark workloads: world shapes, generator seeds, camera poses.

World sizes and the camera recipe follow the reference:
  * world = max_bitw chunks of 2^root_bitw voxels per axis (src/chunkset.h:36-46);
  * camera: glm_perspective(fov, W/H, 0.1, far), view_translation = translate(-location),
    view_rotation = rotate(pitch, x) then rotate(yaw, y), inv_proj_view_rot =
    inverse(projection * view_rotation)  (src/game.c:458-491).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .abi import Camera, GenParams, Sphere, WorldDesc

SEED_BASE = 0x5EED0000


@dataclass(frozen=True)
class WorldConfig:
    name: str
    root_bitw: int
    max_bitw: tuple
    seed: int
    period: int
    hmin: int
    hmax: int

    def desc(self) -> WorldDesc:
        return WorldDesc.make(self.root_bitw, self.max_bitw)

    def gen(self) -> GenParams:
        return GenParams(self.seed, self.period, self.hmin, self.hmax)


# one 64^3 chunk, heights 8..56
CFG0 = WorldConfig("one 64^3 chunk", 6, (0, 0, 0), SEED_BASE + 0, 64, 8, 56)
# 2048 x 256 x 2048 in 64^3 chunks (32 x 4 x 32 = 4096 chunks), heights 0..255
CFG1 = WorldConfig("2048x256x2048 world", 6, (5, 2, 5), SEED_BASE + 1, 256, 0, 255)
# 16384 x 512 x 16384 (256 x 8 x 256 = 524,288 chunks, 128 GiB), heights 0..255
CFG3 = WorldConfig("16384x512x16384 world", 6, (8, 3, 8), SEED_BASE + 3, 256, 0, 255)
# 1024 x 256 x 1024 (16 x 4 x 16 = 1024 chunks)
CFG4 = WorldConfig("1024x256x1024 world", 6, (4, 2, 4), SEED_BASE + 4, 256, 0, 255)
SPLAT_SEED = SEED_BASE + 2
EDIT_SEED = SEED_BASE + 3


def small_world(root_bitw: int, max_bitw, seed: int, hmax: int | None = None) -> WorldConfig:
    """Test-sized world with the same generator; heights span the whole world."""
    Y = (1 << max_bitw[1]) << root_bitw
    R = 1 << root_bitw
    return WorldConfig(f"test {root_bitw}/{max_bitw}", root_bitw, tuple(max_bitw), seed,
                       max(4, 4 * R), 0, Y - 1 if hmax is None else hmax)


# ---- camera ------------------------------------------------------------------
def _perspective(fovy: float, aspect: float, near: float, far: float) -> np.ndarray:
    f = 1.0 / math.tan(fovy * 0.5)
    fn = 1.0 / (near - far)
    m = np.zeros((4, 4))
    m[0, 0] = f / aspect
    m[1, 1] = f
    m[2, 2] = (near + far) * fn
    m[3, 2] = -1.0
    m[2, 3] = 2.0 * near * far * fn
    return m


def _rotation(angle: float, axis) -> np.ndarray:
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    c, s = math.cos(angle), math.sin(angle)
    x, y, z = a
    r = np.eye(4)
    r[:3, :3] = c * np.eye(3) + (1 - c) * np.outer(a, a) + s * np.array(
        [[0, -z, y], [z, 0, -x], [-y, x, 0]])
    return r


def _translation(v) -> np.ndarray:
    t = np.eye(4)
    t[:3, 3] = v
    return t


def _colmajor(m: np.ndarray):
    return [float(v) for v in np.asarray(m, np.float32).T.reshape(-1)]


def make_camera(location, yaw_deg: float, pitch_deg: float, width: int, height: int,
                fov_deg: float = 70.0, far: float = 4096.0, near: float = 0.1) -> Camera:
    """Matrices as src/game.c:458-491 builds them (math row-major here, stored
    column-major as GL receives them).  A positive pitch looks down."""
    P = _perspective(math.radians(fov_deg), width / float(height), near, far)
    T = _translation([-location[0], -location[1], -location[2]])
    Rv = _rotation(math.radians(pitch_deg), (1, 0, 0)) @ _rotation(math.radians(yaw_deg), (0, 1, 0))
    PR = P @ Rv
    cam = Camera()
    for i, v in enumerate(_colmajor(T)):
        cam.view_translation[i] = v
    for i, v in enumerate(_colmajor(PR)):
        cam.proj_view_rot[i] = v
    for i, v in enumerate(_colmajor(np.linalg.inv(PR))):
        cam.inv_proj_view_rot[i] = v
    cam.width, cam.height, cam.far_plane = width, height, far
    return cam


def make_view_projection(location, yaw_deg: float, pitch_deg: float, width: int, height: int,
                         fov_deg: float = 70.0, far: float = 4096.0, near: float = 0.1) -> np.ndarray:
    """projection * view_rotation * view_translation of make_camera's pose,
    column-major float32[16] (game.c:470-483); its frustum planes and the
    camera location drive the octree LOD selection."""
    P = _perspective(math.radians(fov_deg), width / float(height), near, far)
    T = _translation([-location[0], -location[1], -location[2]])
    Rv = _rotation(math.radians(pitch_deg), (1, 0, 0)) @ _rotation(math.radians(yaw_deg), (0, 1, 0))
    return np.asarray(_colmajor(P @ (Rv @ T)), np.float32)


def splat_poses(terrain_height, n: int = 16, seed: int = SPLAT_SEED, lo: float = 256.0,
                hi: float = 1792.0):
    """Seeded poses: x, z in [lo, hi], 64..512 above the terrain, yaw in
    [0, 360), looking 10..60 degrees below the horizon.  terrain_height(x, z)
    returns the top of the column."""
    rng = np.random.default_rng(seed)
    poses = []
    for _ in range(n):
        x, z = rng.uniform(lo, hi, 2)
        y = terrain_height(int(x), int(z)) + rng.uniform(64.0, 512.0)
        yaw = rng.uniform(0.0, 360.0)
        pitch = rng.uniform(10.0, 60.0)
        poses.append(((float(x), float(y), float(z)), float(yaw), float(pitch)))
    return poses


def edit_burst(world: WorldDesc, n: int = 10000, seed: int = EDIT_SEED, rmin: int = 1,
               rmax: int = 32):
    """Seeded sphere edits: centre uniform in the world, radius uniform in
    [rmin, rmax], add (colour 1..63) or remove with p = 1/2."""
    rng = np.random.default_rng(seed)
    X, Y, Z = world.size
    out = []
    for _ in range(n):
        c = (int(rng.integers(0, X)), int(rng.integers(0, Y)), int(rng.integers(0, Z)))
        r = int(rng.integers(rmin, rmax + 1))
        add = bool(rng.integers(0, 2))
        v = int(rng.integers(1, 64)) if add else 0
        out.append(Sphere(c[0], c[1], c[2], r, v))
    return out
