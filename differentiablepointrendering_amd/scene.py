"""Synthetic scenes for the beta-surfel renderer.

Each generator is seeded (numpy ``default_rng``) and draws in a fixed,
documented order so that every scene can be regenerated bit for bit:

  plain scene      centers, quaternions (Shoemake u1,u2,u3), scales (u, v),
                   beta, opacity, albedo; then the target jitter.
  clustered scene  cluster means, cluster id per surfel, Gaussian offsets,
                   then as above from the quaternions on.

Tangents are the first two columns of the rotation matrix of the quaternion,
canonicalised face-up (normal z >= 0 by flipping the second tangent) as the
reference's PLY loader does (PLYPointLoader.ixx:75-101).

The light is the reference's unit quad mesh (Assets/meshes/rectangle.obj:
vertices (+-1, +-1, 0), faces (2,3,1) and (2,4,3)) scaled to 0.5 x 0.5,
rotated so that its normal points down (-y) and centred at y = 1.5.  A
radiance L maps to the reference's emitter power Phi = L * pi * A
(MaterialBakerMitsuba.ixx:42-66 stores power and colour, not radiance).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from ._abi import Camera, Mesh

QUAD_POSITIONS = np.array([[-1.0, -1.0, 0.0], [1.0, -1.0, 0.0], [-1.0, 1.0, 0.0], [1.0, 1.0, 0.0]],
                          dtype=np.float32)
QUAD_INDICES = np.array([[1, 2, 0], [1, 3, 2]], dtype=np.uint32)


def light_quad(y: float = 1.5, size: float = 0.5, radiance: float = 10.0,
               color=(1.0, 1.0, 1.0), name: str = "light") -> Mesh:
    """Area light facing -y.  object->world = T(0,y,0) Rx(90) S(size/2)."""
    h = size / 2.0
    o2w = np.array([[h, 0, 0, 0], [0, 0, -h, y], [0, h, 0, 0], [0, 0, 0, 1]], dtype=np.float64)
    w2o = np.linalg.inv(o2w)
    power = radiance * math.pi * size * size
    return Mesh(QUAD_POSITIONS.copy(), QUAD_INDICES.copy(), o2w.astype(np.float32), w2o.astype(np.float32),
                tuple(color), float(np.float32(power)), name)


def quad_mesh(name: str, center, rotation, half_size: float, base_color, power: float = 0.0) -> Mesh:
    """The unit quad (normal +z in object space) scaled by half_size,
    rotated by the 3x3 `rotation` (its third column is the world normal) and
    moved to `center`: a floor, wall or ceiling of test scenes."""
    R = np.asarray(rotation, dtype=np.float64).reshape(3, 3)
    o2w = np.eye(4)
    o2w[:3, :3] = R * half_size
    o2w[:3, 3] = np.asarray(center, dtype=np.float64)
    w2o = np.linalg.inv(o2w)
    return Mesh(QUAD_POSITIONS.copy(), QUAD_INDICES.copy(), o2w.astype(np.float32), w2o.astype(np.float32),
                tuple(float(c) for c in base_color), float(power), name)


# Rotations whose third column (the quad normal) is +y, -y or +z.
FACING_UP = ((1, 0, 0), (0, 0, 1), (0, -1, 0))
FACING_DOWN = ((1, 0, 0), (0, 0, -1), (0, 1, 0))
FACING_PLUS_Z = ((1, 0, 0), (0, 1, 0), (0, 0, 1))


def look_at_camera(name: str, eye, target, up, width: int, height: int, fovy_deg: float,
                   use_for_adjoint: bool = True) -> Camera:
    """Pinhole camera from a look-at (right-handed, camera looks down -z)."""
    eye = np.asarray(eye, dtype=np.float64)
    target = np.asarray(target, dtype=np.float64)
    up = np.asarray(up, dtype=np.float64)
    f = target - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    view = np.array([[s[0], s[1], s[2], -s @ eye],
                     [u[0], u[1], u[2], -u @ eye],
                     [-f[0], -f[1], -f[2], f @ eye],
                     [0, 0, 0, 1]], dtype=np.float64)
    inv_view = np.linalg.inv(view)
    return Camera(name, int(width), int(height), float(fovy_deg), view.astype(np.float32),
                  inv_view.astype(np.float32), eye.astype(np.float32), f.astype(np.float32), use_for_adjoint)


def quaternion_tangents(q: np.ndarray):
    """Tangents (first two rotation-matrix columns) of unit quaternions
    q = (w, x, y, z), canonicalised so that the normal has z >= 0."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    tu = np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)], axis=1)
    tv = np.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], axis=1)
    n = np.cross(tu, tv)
    flip = n[:, 2] < 0
    tv[flip] = -tv[flip]
    return tu, tv


def shoemake_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Uniform random unit quaternions (Shoemake 1992), (w, x, y, z)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    a, b = np.sqrt(1 - u1), np.sqrt(u1)
    q = np.stack([b * np.cos(2 * np.pi * u3), a * np.sin(2 * np.pi * u2), a * np.cos(2 * np.pi * u2),
                  b * np.sin(2 * np.pi * u3)], axis=1)
    return q


def _finish(rng, centers, n, scale_lo, scale_hi):
    q = shoemake_quaternions(rng, n)
    tu, tv = quaternion_tangents(q)
    lo, hi = math.log(scale_lo), math.log(scale_hi)
    scale = np.exp(rng.uniform(lo, hi, size=(n, 2)))
    beta = rng.uniform(-0.5, 0.5, size=n)
    opacity = rng.uniform(0.3, 0.9, size=n)
    albedo = rng.uniform(0.2, 0.8, size=(n, 3))
    pts = dict(position=centers, tangent_u=tu, tangent_v=tv, scale=scale, albedo=albedo,
               opacity=opacity, beta=beta, shape=np.zeros(n))
    return {k: np.ascontiguousarray(v, dtype=np.float32) for k, v in pts.items()}


def uniform_surfels(n: int, seed: int = 0, scale_lo: float = 0.01, scale_hi: float = 0.05,
                    jitter: float = 0.01):
    """Plain scene: centers uniform in [-1,1]^3.  Returns (points, target_points)."""
    rng = np.random.default_rng(seed)
    centers = rng.uniform(-1.0, 1.0, size=(n, 3))
    pts = _finish(rng, centers, n, scale_lo, scale_hi)
    target = dict(pts)
    target["position"] = np.ascontiguousarray(
        (pts["position"].astype(np.float64) + rng.normal(0.0, jitter, size=(n, 3))).astype(np.float32))
    return pts, target


def clustered_surfels(n: int, seed: int = 0, clusters: int = 64, sigma: float = 0.1,
                      scale_lo: float = 0.002, scale_hi: float = 0.02, jitter: float = 0.01):
    """Clustered scene: 64 Gaussian clusters (sigma 0.1), means uniform in
    [-1,1]^3.  Returns (points, target_points)."""
    rng = np.random.default_rng(seed)
    means = rng.uniform(-1.0, 1.0, size=(clusters, 3))
    cid = rng.integers(0, clusters, size=n)
    centers = means[cid] + rng.normal(0.0, sigma, size=(n, 3))
    pts = _finish(rng, centers, n, scale_lo, scale_hi)
    target = dict(pts)
    target["position"] = np.ascontiguousarray(
        (pts["position"].astype(np.float64) + rng.normal(0.0, jitter, size=(n, 3))).astype(np.float32))
    return pts, target


@dataclass
class SceneSpec:
    meshes: List[Mesh]
    cameras: List[Camera]
    points: Dict[str, np.ndarray]
    settings: Dict[str, object] = field(default_factory=dict)
    target_points: Optional[Dict[str, np.ndarray]] = None
    name: str = "scene"


def default_camera(width: int, height: int, name: str = "camera1") -> Camera:
    return look_at_camera(name, (0.0, 0.0, 4.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), width, height, 45.0)


def config_scene(config: int, n: Optional[int] = None, width: Optional[int] = None,
                 height: Optional[int] = None, seed: int = 0) -> SceneSpec:
    """Scenes of the benchmark configurations.

    config 0: 4,096 uniform surfels, 128x128, 4 spp, depth 4, 100k photons,
              gather radius 0.05.
    config 1: 1,048,576 clustered surfels, 800x800, 16 spp, depth 6,
              2M photons, gather radius 0.02.
    ``n``/``width``/``height`` shrink a configuration for tests.
    """
    if config == 0:
        n = 4096 if n is None else n
        w = 128 if width is None else width
        h = 128 if height is None else height
        pts, tgt = uniform_surfels(n, seed=seed)
        settings = dict(photons=100_000, bounces=4, forward_passes=1, gather_passes=4, adjoint_bounces=4,
                        adjoint_passes=4, seed=0, gather_radius=0.05)
    elif config == 1:
        n = 1_048_576 if n is None else n
        w = 800 if width is None else width
        h = 800 if height is None else height
        pts, tgt = clustered_surfels(n, seed=seed)
        settings = dict(photons=2_000_000, bounces=6, forward_passes=1, gather_passes=16, adjoint_bounces=6,
                        adjoint_passes=16, seed=0, gather_radius=0.02)
    else:
        raise ValueError(f"unknown config {config}")
    return SceneSpec([light_quad()], [default_camera(w, h)], pts, settings, tgt, name=f"config{config}")


def single_surfel_scene(position=(0.0, 0.0, 0.0), tangent_u=(1.0, 0.0, 0.0), tangent_v=(0.0, 1.0, 0.0),
                        scale=(0.5, 0.5), albedo=(0.8, 0.8, 0.8), opacity=0.7, beta=0.0,
                        width: int = 64, height: int = 64) -> SceneSpec:
    """One surfel under the light quad, seen by the default camera."""
    pts = dict(position=[position], tangent_u=[tangent_u], tangent_v=[tangent_v], scale=[scale],
               albedo=[albedo], opacity=[opacity], beta=[beta], shape=[0.0])
    pts = {k: np.ascontiguousarray(np.asarray(v, dtype=np.float32)) for k, v in pts.items()}
    return SceneSpec([light_quad()], [default_camera(width, height)], pts, {}, None, name="single")


# ------------------------------------------------- batched and served sets --
def batched_cameras(count: int = 32, seed: int = 2):
    """Views of the batched configuration: eyes uniform on the sphere of
    radius 4 looking at the origin, pixel area log-uniform between 128x96 and
    1024x768 (aspect 4:3), spp uniform in {4, 8, 16}.  Returns (cameras,
    spp)."""
    rng = np.random.default_rng(seed)
    lo, hi = math.log(128 * 96), math.log(1024 * 768)
    cams, spp = [], []
    for i in range(count):
        area = math.exp(rng.uniform(lo, hi))
        w = max(8, int(round(math.sqrt(area * 4.0 / 3.0))))
        h = max(6, int(round(w * 0.75)))
        d = rng.normal(size=3)
        eye = 4.0 * d / np.linalg.norm(d)
        up = (0.0, 1.0, 0.0) if abs(eye[1]) < 3.9 else (1.0, 0.0, 0.0)
        cams.append(look_at_camera(f"v{i}", eye, (0.0, 0.0, 0.0), up, w, h, 45.0))
        spp.append(int(rng.choice([4, 8, 16])))
    return cams, spp


# The served configuration's resident scenes: (clustered surfels, photons).
SERVED_SCENES = ((100_000, 1_000_000), (500_000, 2_000_000), (1_000_000, 3_000_000), (3_000_000, 4_000_000))


def served_requests(count: int, rate: float, seed: int = 1):
    """The request stream of the served configuration: Poisson arrivals at
    `rate` per second; each request picks a resident scene uniformly, is
    forward-only with p = 0.7 (else forward + adjoint), has an area
    log-uniform between 256x256 and 1920x1080 (aspect in [1, 16/9]), spp
    uniform in {1, 4, 16, 64} and an eye uniform on the radius-4 sphere.
    Returns [(arrival s, scene index, camera, spp, adjoint)]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    lo, hi = math.log(256 * 256), math.log(1920 * 1080)
    for i in range(count):
        t += rng.exponential(1.0 / rate)
        sc = int(rng.integers(0, len(SERVED_SCENES)))
        adjoint = bool(rng.random() >= 0.7)
        area = math.exp(rng.uniform(lo, hi))
        aspect = rng.uniform(1.0, 16.0 / 9.0)
        w = max(16, int(round(math.sqrt(area * aspect))))
        h = max(16, int(round(area / w)))
        spp = int(rng.choice([1, 4, 16, 64]))
        d = rng.normal(size=3)
        eye = 4.0 * d / np.linalg.norm(d)
        up = (0.0, 1.0, 0.0) if abs(eye[1]) < 3.9 else (1.0, 0.0, 0.0)
        cam = look_at_camera(f"r{i}", eye, (0.0, 0.0, 0.0), up, w, h, 45.0)
        out.append((t, sc, cam, spp, adjoint))
    return out
