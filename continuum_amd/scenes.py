"""Seeded synthetic scenes for the cone-visibility path, plus the reference's
host-side packing of lights and cameras.

Scenes (all Z-up, FP32 geometry):

* ``room_scene``  - a closed room 4 x 4 x 3 m, each of its 6 inward faces a
  32 x 32 quad grid (12,288 triangles), with two boxes of random size
  0.3-1.0 m on the floor, one 0.5 x 0.5 m one-sided light under the ceiling,
  a 256 x 256 pinhole view at 60 degrees.
* ``rock_field``  - 100 geodesic spheres (frequency 22, 9,680 triangles
  each) displaced by 4 octaves of 3D sinusoids, lognormal radii
  (mu 0, sigma 0.5), scattered over a 50 x 50 m ground grid of 180 x 180
  quads; a 0.53 degree sun disk and two square one-sided lights of 1-4 m^2;
  a 1280 x 720 view looking down into the field (every pixel a shading
  point).

Random numbers come from PCG32 (O'Neill), as the reference's experiments use
it (experiments/flatland/src/rng.rs:3-30).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

F32 = np.float32


class Pcg32:
    """PCG32 with the reference's seeding (experiments/flatland/src/rng.rs:9-17)."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int, stream: int):
        self.state = 0
        self.inc = ((stream << 1) | 1) & self.MASK
        self.next_u32()
        self.state = (self.state + seed) & self.MASK
        self.next_u32()

    def next_u32(self) -> int:
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & self.MASK
        xorshifted = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xorshifted >> rot) | (xorshifted << ((-rot) & 31))) & 0xFFFFFFFF

    def next_f32(self) -> float:
        return float(F32(self.next_u32() >> 8) * F32(1.0 / 16777216.0))


# ---- lights and cameras (the reference's packing)


def rect_light(corner, edge_u, edge_v, one_sided: bool = False, radiance=(1.0, 1.0, 1.0)) -> np.ndarray:
    """LightGpu of a rectangle (crates/continuum-gpu/src/lib.rs:723-728)."""
    return np.array([*corner, 1.0 if one_sided else 0.0, *edge_u, 0.0, *edge_v, 0.0, *radiance, 0.0], dtype=F32)


def sphere_light(center, radius: float, radiance=(1.0, 1.0, 1.0)) -> np.ndarray:
    """LightGpu of a sphere without a spot (lib.rs:729-744)."""
    return np.array([*center, 2.0, radius, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, 0.0, *radiance, 0.0], dtype=F32)


def _glam_normalize(v: np.ndarray) -> np.ndarray:
    v = v.astype(F32)
    n = F32(math.sqrt(float(F32(v[0] * v[0]) + F32(v[1] * v[1]) + F32(v[2] * v[2]))))
    return (v * (F32(1.0) / n)).astype(F32)


def _cross(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], dtype=F32)


def camera_uniform(eye, target, fov_degrees: float, width: int, height: int, radiance: float = 1.0) -> np.ndarray:
    """CameraUniform (lib.rs:3538-3553): eye, forward, right, up, params."""
    eye = np.asarray(eye, dtype=F32)
    forward = _glam_normalize(np.asarray(target, dtype=F32) - eye)
    right = _glam_normalize(_cross(forward, np.array([0, 0, 1], dtype=F32)))
    up = _cross(right, forward)
    half = F32(F32(fov_degrees) * F32(math.pi / 180.0)) * F32(0.5)
    return np.array(
        [*eye, 0, *forward, 0, *right, 0, *up, 0, width, height, math.tan(float(half)), radiance], dtype=F32)


# ---- meshes


def grid_quad(origin, du, dv, n: int, m: int) -> np.ndarray:
    """A parallelogram origin + [0,1]du + [0,1]dv as n x m quads, two
    triangles each, wound counter-clockwise about du x dv."""
    origin, du, dv = (np.asarray(x, dtype=np.float64) for x in (origin, du, dv))
    i = np.arange(n)[:, None]
    j = np.arange(m)[None, :]

    def p(a, b):
        return origin + (a[..., None] / n) * du + (b[..., None] / m) * dv

    a, b, c, d = p(i, j), p(i + 1, j), p(i + 1, j + 1), p(i, j + 1)
    t1 = np.stack([a, b, c], axis=-2)
    t2 = np.stack([a, c, d], axis=-2)
    return np.concatenate([t1.reshape(-1, 3, 3), t2.reshape(-1, 3, 3)]).astype(F32)


def box_mesh(lo, hi) -> np.ndarray:
    """Twelve outward triangles (crates/continuum/src/cone3.rs:462-515)."""
    lo = np.asarray(lo, dtype=F32)
    hi = np.asarray(hi, dtype=F32)

    def c(x, y, z):
        return [hi[0] if x else lo[0], hi[1] if y else lo[1], hi[2] if z else lo[2]]

    quads = [
        [c(0, 0, 0), c(0, 1, 0), c(1, 1, 0), c(1, 0, 0)],
        [c(0, 0, 1), c(1, 0, 1), c(1, 1, 1), c(0, 1, 1)],
        [c(0, 0, 0), c(1, 0, 0), c(1, 0, 1), c(0, 0, 1)],
        [c(0, 1, 0), c(0, 1, 1), c(1, 1, 1), c(1, 1, 0)],
        [c(0, 0, 0), c(0, 0, 1), c(0, 1, 1), c(0, 1, 0)],
        [c(1, 0, 0), c(1, 1, 0), c(1, 1, 1), c(1, 0, 1)],
    ]
    tris = []
    for q in quads:
        tris.append([q[0], q[1], q[2]])
        tris.append([q[0], q[2], q[3]])
    return np.array(tris, dtype=F32)


@dataclass
class SceneSpec:
    """Meshes (each (n, 3, 3) f32), lights (LightGpu rows), camera, view size,
    and the kernel options the configuration is quoted with."""

    name: str
    meshes: list
    lights: np.ndarray
    camera: np.ndarray
    width: int
    height: int
    prefilter: float = 0.0
    notes: dict = field(default_factory=dict)

    @property
    def n_tris(self) -> int:
        return int(sum(len(m) for m in self.meshes))


def room_scene(seed: int = 0xC0_0001, width: int = 256, height: int = 256) -> SceneSpec:
    """The closed room with two random boxes (12,312 triangles, 8 objects)."""
    X, Y, Z, N = 4.0, 4.0, 3.0, 32
    faces = [
        grid_quad([0, 0, 0], [X, 0, 0], [0, Y, 0], N, N),      # floor, normal +z
        grid_quad([0, 0, Z], [0, Y, 0], [X, 0, 0], N, N),      # ceiling, -z
        grid_quad([0, 0, 0], [0, Y, 0], [0, 0, Z], N, N),      # x = 0, +x
        grid_quad([X, 0, 0], [0, 0, Z], [0, Y, 0], N, N),      # x = 4, -x
        grid_quad([0, 0, 0], [0, 0, Z], [X, 0, 0], N, N),      # y = 0, +y
        grid_quad([0, Y, 0], [X, 0, 0], [0, 0, Z], N, N),      # y = 4, -y
    ]
    rng = Pcg32(seed, 7)
    boxes = []
    for _ in range(2):
        size = [0.3 + 0.7 * rng.next_f32() for _ in range(3)]
        x = 0.5 + rng.next_f32() * (3.0 - size[0])
        y = 0.5 + rng.next_f32() * (3.0 - size[1])
        boxes.append(box_mesh([x, y, 0.0], [x + size[0], y + size[1], size[2]]))
    light = rect_light([1.75, 2.25, 2.999], [0.5, 0, 0], [0, -0.5, 0], one_sided=True)
    cam = camera_uniform([0.4, 0.4, 1.6], [3.6, 3.6, 0.8], 60.0, width, height)
    return SceneSpec("room", faces + boxes, light[None, :], cam, width, height, prefilter=0.0)


def _icosphere_grid(freq: int) -> np.ndarray:
    """Unit geodesic sphere: each icosahedron face split into freq^2
    triangles, vertices projected to the sphere; (20 freq^2, 3, 3)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    tris = []
    for a, b, c in f:
        A, B, Cc = v[a], v[b], v[c]

        def P(i, j):
            return A + (B - A) * (i / freq) + (Cc - A) * (j / freq)

        for i in range(freq):
            for j in range(freq - i):
                tris.append([P(i, j), P(i + 1, j), P(i, j + 1)])
                if i + j + 1 < freq:
                    tris.append([P(i + 1, j), P(i + 1, j + 1), P(i, j + 1)])
    arr = np.array(tris)
    return arr / np.linalg.norm(arr, axis=-1, keepdims=True)


def rock_field(seed: int = 0xC0_0002, width: int = 1280, height: int = 720, rocks: int = 100) -> SceneSpec:
    """The 1M-triangle rock field (about 1.03M triangles, 101 objects)."""
    rng = Pcg32(seed, 11)
    unit = _icosphere_grid(22)  # 9,680 triangles
    half = 25.0
    meshes = [grid_quad([-half, -half, 0.0], [2 * half, 0, 0], [0, 2 * half, 0], 180, 180)]
    for _ in range(rocks):
        # lognormal radius: exp(N(0, 0.5)) by Box-Muller on two uniforms
        u1 = max(rng.next_f32(), 1e-7)
        u2 = rng.next_f32()
        radius = math.exp(0.5 * math.sqrt(-2.0 * math.log(u1)) * math.cos(2 * math.pi * u2))
        cx = -22.0 + 44.0 * rng.next_f32()
        cy = -22.0 + 44.0 * rng.next_f32()
        disp = np.zeros(unit.shape[:-1])
        for octave in range(4):
            freq = 2.0 ** (octave + 1)
            amp = 0.18 / (2.0 ** octave)
            for _w in range(3):
                d = np.array([rng.next_f32() - 0.5, rng.next_f32() - 0.5, rng.next_f32() - 0.5])
                d /= max(np.linalg.norm(d), 1e-6)
                phase = 2 * math.pi * rng.next_f32()
                disp += amp / 3.0 * np.sin(freq * (unit @ d) + phase)
        pts = unit * (radius * (1.0 + disp))[..., None]
        pts[..., 2] *= 0.7
        pts += np.array([cx, cy, 0.25 * radius])
        meshes.append(pts.astype(F32))
    sun_dir = np.array([-0.45, -0.3, 0.84])
    sun_dir /= np.linalg.norm(sun_dir)
    sun = sphere_light(sun_dir * 1000.0, 1000.0 * math.tan(math.radians(0.265)), radiance=(3.0, 2.9, 2.7))
    lights = [sun]
    for _ in range(2):
        area = 1.0 + 3.0 * rng.next_f32()
        s = math.sqrt(area)
        cx = -15.0 + 30.0 * rng.next_f32()
        cy = -15.0 + 30.0 * rng.next_f32()
        z = 6.0 + 4.0 * rng.next_f32()
        # One-sided, emitting down: edge_u x edge_v points to -z.
        lights.append(rect_light([cx - s / 2, cy + s / 2, z], [s, 0, 0], [0, -s, 0], one_sided=True,
                                 radiance=(10.0, 10.0, 10.0)))
    # Looking down into the field: every pixel sees the ground or a rock.
    cam = camera_uniform([0.0, -5.0, 19.0], [0.0, 1.0, 0.0], 60.0, width, height)
    return SceneSpec("rock_field", meshes, np.stack(lights), cam, width, height, prefilter=1.0,
                     notes={"eye": (0.0, -5.0, 19.0), "target": (0.0, 1.0, 0.0), "fov": 60.0})


def meshes_to_arrays(meshes) -> tuple[np.ndarray, np.ndarray]:
    """Concatenated triangles (n, 9) f32 and per-mesh counts (u32)."""
    counts = np.array([len(m) for m in meshes], dtype=np.uint32)
    if counts.sum() == 0:
        return np.zeros((0, 9), dtype=F32), counts
    tris = np.ascontiguousarray(np.concatenate([np.asarray(m, dtype=F32).reshape(-1, 9) for m in meshes]))
    return tris, counts
