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


def sun_light(center, toward, distance: float, irradiance=(1.0, 1.0, 1.0)) -> np.ndarray:
    """LightGpu of the reference's `Light::sun` (lib.rs:661-700, through
    `Light::distant`): a one-sided square 0.5 degrees across, `distance` from
    `center` toward `toward`, facing `center`; radiance irradiance / its
    solid angle. f64 as there, then f32."""
    size_angle = 2.0 * math.tan(math.radians(0.25))
    solid_angle = size_angle * size_angle
    t = np.asarray(toward, np.float64)
    t = t / np.linalg.norm(t)
    size = distance * math.sqrt(max(solid_angle, 1e-12))
    w = -t
    z = np.array([0.0, 0.0, 1.0])
    side = np.cross(z, w) if np.linalg.norm(np.cross(w, z)) > 1e-6 else np.array([1.0, 0.0, 0.0])
    a = side / np.linalg.norm(side) * size
    b = np.cross(w, a)
    b = b / np.linalg.norm(b) * size
    middle = np.asarray(center, np.float64) + t * distance
    solid = (size / distance) * (size / distance)
    radiance = np.asarray(irradiance, np.float64) / solid
    return rect_light(middle - (a + b) * 0.5, a, b, one_sided=True, radiance=tuple(radiance))


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


def quad_pair(a, du, dv) -> np.ndarray:
    """One parallelogram a + [0,1]du + [0,1]dv as the two triangles (a, b, c),
    (a, c, d) that `Surface::from_triangles` takes as one chart."""
    a, du, dv = (np.asarray(x, dtype=F32) for x in (a, du, dv))
    b, c, d = a + du, a + du + dv, a + dv
    return np.array([[a, b, c], [a, c, d]], dtype=F32)


def lit_room(seed: int = 0xC0_0005, boxes: int = 2, size: float = 1.0) -> SceneSpec:
    """A room of large quads (floor, ceiling and four walls, each one chart)
    with a static pillar and `boxes` dynamic boxes on the floor, lit by a
    one-sided ceiling rect: the lighting cache's workload. Meshes: 6 faces,
    the pillar, then the boxes; `notes['dynamic']` flags the boxes."""
    X, Y, Z = 4.0 * size, 4.0 * size, 3.0 * size
    faces = [
        quad_pair([0, 0, 0], [X, 0, 0], [0, Y, 0]),
        quad_pair([0, 0, Z], [0, Y, 0], [X, 0, 0]),
        quad_pair([0, 0, 0], [0, Y, 0], [0, 0, Z]),
        quad_pair([X, 0, 0], [0, 0, Z], [0, Y, 0]),
        quad_pair([0, 0, 0], [0, 0, Z], [X, 0, 0]),
        quad_pair([0, Y, 0], [X, 0, 0], [0, 0, Z]),
    ]
    pillar = box_mesh([0.6 * size, 2.8 * size, 0.0], [1.0 * size, 3.2 * size, 2.2 * size])
    rng = Pcg32(seed, 11)
    movers = []
    for _ in range(boxes):
        s = [(0.25 + 0.35 * rng.next_f32()) * size for _ in range(3)]
        x = (0.4 + rng.next_f32() * 2.8) * size
        y = (0.4 + rng.next_f32() * 2.2) * size
        movers.append(box_mesh([x, y, 0.0], [x + s[0], y + s[1], s[2]]))
    light = rect_light([1.6 * size, 2.4 * size, Z - 0.001], [0.8 * size, 0, 0], [0, -0.8 * size, 0], one_sided=True)
    cam = camera_uniform([0.4 * size, 0.4 * size, 1.6 * size], [3.6 * size, 3.6 * size, 0.8 * size], 60.0, 256, 256)
    meshes = faces + [pillar] + movers
    spec = SceneSpec("lit room", meshes, light[None, :], cam, 256, 256, prefilter=0.0)
    spec.notes["dynamic"] = [False] * 7 + [True] * boxes
    return spec


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


def _cylinder(center, radius: float, z0: float, z1: float, seg: int, rings: int) -> np.ndarray:
    """An outward cylinder side (no caps), seg x rings quads."""
    th = np.linspace(0.0, 2.0 * np.pi, seg + 1)
    zs = np.linspace(z0, z1, rings + 1)
    T, Zz = np.meshgrid(th, zs, indexing="ij")
    P = np.stack([center[0] + radius * np.cos(T), center[1] + radius * np.sin(T), Zz], -1)
    a, b, c, d = P[:-1, :-1], P[1:, :-1], P[1:, 1:], P[:-1, 1:]
    return np.concatenate([np.stack([a, b, c], -2).reshape(-1, 3, 3),
                           np.stack([a, c, d], -2).reshape(-1, 3, 3)]).astype(F32)


def _vault(x0: float, x1: float, y0: float, y1: float, z_spring: float, seg: int, rows: int, inward: bool) -> np.ndarray:
    """A half-cylinder arch over [x0, x1] (axis along y from y0 to y1),
    springing at z_spring; its inside faces down (toward the passage)."""
    r = 0.5 * (x1 - x0)
    cx = 0.5 * (x0 + x1)
    th = np.linspace(np.pi, 0.0, seg + 1)
    ys = np.linspace(y0, y1, rows + 1)
    T, Y = np.meshgrid(th, ys, indexing="ij")
    P = np.stack([cx + r * np.cos(T), Y, z_spring + r * np.sin(T)], -1)
    a, b, c, d = P[:-1, :-1], P[1:, :-1], P[1:, 1:], P[:-1, 1:]
    t1, t2 = np.stack([a, b, c], -2), np.stack([a, c, d], -2)
    if not inward:
        t1, t2 = t1[..., ::-1, :], t2[..., ::-1, :]
    return np.concatenate([t1.reshape(-1, 3, 3), t2.reshape(-1, 3, 3)]).astype(F32)


def _curtain(x0: float, x1: float, y: float, z_top: float, z_bot: float, nu: int, nv: int, sag: float,
             phase: float) -> np.ndarray:
    """A hanging curtain grid in the plane y (both faces: two sheets a
    hair apart), sagging toward +y by up to `sag`."""
    u = np.linspace(0.0, 1.0, nu + 1)
    v = np.linspace(0.0, 1.0, nv + 1)
    U, V = np.meshgrid(u, v, indexing="ij")
    X = x0 + (x1 - x0) * U
    Z = z_top + (z_bot - z_top) * V
    S = sag * np.sin(np.pi * U) * V * (0.8 + 0.2 * np.sin(6.0 * U + phase))
    out = []
    for side in (1.0, -1.0):
        P = np.stack([X, y + S + side * 0.002, Z], -1)
        a, b, c, d = P[:-1, :-1], P[1:, :-1], P[1:, 1:], P[:-1, 1:]
        t1, t2 = np.stack([a, b, c], -2), np.stack([a, c, d], -2)
        if side < 0:
            t1, t2 = t1[..., ::-1, :], t2[..., ::-1, :]
        out += [t1.reshape(-1, 3, 3), t2.reshape(-1, 3, 3)]
    return np.concatenate(out).astype(F32)


def courtyard(seed: int = 0xC0_0003, width: int = 1280, height: int = 720, detail: float = 1.0) -> SceneSpec:
    """The 3M-triangle courtyard: an open atrium 30 x 15 x 20 m; on each long
    side two floors of 12 arches (48 in all) on columns, a gallery slab
    between the floors; hanging curtains (about 40% of the triangles) under
    the upper arches. `detail` scales every tessellation (1.0: about 3M
    triangles). ``notes['curtains']`` lists the curtain meshes, which
    `courtyard_frame` displaces each frame."""
    rng = Pcg32(seed, 13)
    L, W, H = 30.0, 15.0, 20.0
    k = math.sqrt(detail)

    def n(x):
        return max(2, int(round(x * k)))

    meshes = []
    # ground and outer walls (inward)
    meshes.append(grid_quad([0, 0, 0], [L, 0, 0], [0, W, 0], n(450), n(225)))
    meshes.append(grid_quad([0, 0, 0], [0, W, 0], [0, 0, H], n(225), n(300)))
    meshes.append(grid_quad([L, 0, 0], [0, 0, H], [0, W, 0], n(300), n(225)))
    meshes.append(grid_quad([0, 0, 0], [0, 0, H], [L, 0, 0], n(300), n(450)))
    meshes.append(grid_quad([0, W, 0], [L, 0, 0], [0, 0, H], n(450), n(300)))
    bay = L / 12.0
    depth = 3.0
    curtains = []
    for floor, (z0, z1) in enumerate(((0.0, 10.0), (10.0, 20.0))):
        spring = z0 + 6.0
        for side, (ya, yb) in enumerate(((0.0, depth), (W - depth, W))):
            for i in range(12):
                xa, xb = i * bay + 0.35, (i + 1) * bay - 0.35
                meshes.append(_vault(xa, xb, ya, yb, spring, n(72), n(36), inward=True))
                meshes.append(_cylinder([i * bay + 0.175 + 0.0, ya + (depth if side == 0 else 0.0), 0], 0.3, z0,
                                        spring, n(60), n(45)))
            # the gallery slab over this floor's arcade (downward and upward faces)
            if floor == 0:
                y_in = depth if side == 0 else W - depth
                slab_lo = grid_quad([0, ya, 9.6], [0, yb - ya, 0], [L, 0, 0], n(90), n(450))
                slab_hi = grid_quad([0, ya, 10.0], [L, 0, 0], [0, yb - ya, 0], n(450), n(90))
                meshes += [slab_lo, slab_hi]
                del y_in
    # curtains under the upper arches, toward the courtyard
    for side, y in enumerate((depth + 0.4, W - depth - 0.4)):
        for i in range(12):
            xa, xb = i * bay + 0.45, (i + 1) * bay - 0.45
            sag = 0.2 + 0.6 * rng.next_f32()
            phase = 2.0 * math.pi * rng.next_f32()
            c = _curtain(xa, xb, y, 15.8, 10.6, n(112), n(112), sag if side == 0 else -sag, phase)
            curtains.append(len(meshes))
            meshes.append(c)
    # lights: a sun disk and three area lights of 0.5-8 m^2 with random colours
    sun_dir = np.array([0.35, 0.25, 0.9])
    sun_dir /= np.linalg.norm(sun_dir)
    lights = [sphere_light(sun_dir * 1000.0 + np.array([L / 2, W / 2, 0]), 1000.0 * math.tan(math.radians(0.265)),
                           radiance=(3.0, 2.9, 2.7))]
    for _ in range(3):
        area = 0.5 + 7.5 * rng.next_f32()
        s = math.sqrt(area)
        cx = 4.0 + 22.0 * rng.next_f32()
        cy = 4.5 + 6.0 * rng.next_f32()
        z = 8.0 + 9.0 * rng.next_f32()
        col = (0.3 + 0.7 * rng.next_f32(), 0.3 + 0.7 * rng.next_f32(), 0.3 + 0.7 * rng.next_f32())
        lights.append(rect_light([cx - s / 2, cy + s / 2, z], [s, 0, 0], [0, -s, 0], one_sided=True,
                                 radiance=tuple(10.0 * c for c in col)))
    eye, target = (2.5, 7.5, 3.0), (22.0, 6.0, 7.0)
    cam = camera_uniform(eye, target, 60.0, width, height)
    return SceneSpec("courtyard", meshes, np.stack(lights), cam, width, height, prefilter=1.0,
                     notes={"curtains": curtains, "eye": eye, "target": target, "fov": 60.0, "seed": seed})


def _boulder_grid(coef, nu: int, nv: int) -> np.ndarray:
    a1, a2, a3 = coef
    u = np.arange(nu + 1, dtype=np.float64) / nu
    v = np.arange(nv + 1, dtype=np.float64) / nv
    th, ph = np.meshgrid(2 * np.pi * u, np.pi * v, indexing="ij")
    r = (1.0 + a1 * np.sin(3 * th + 1) * np.sin(2 * ph) + a2 * np.cos(5 * th) * np.sin(3 * ph + 0.5)
         + a3 * np.sin(9 * th + 2 * ph))
    return np.stack([np.sin(ph) * np.cos(th) * r * 0.5, np.sin(ph) * np.sin(th) * r * 0.5, -np.cos(ph) * r * 0.33],
                    axis=-1).astype(F32)


def boulder(coef, nu: int = 100, nv: int = 50) -> np.ndarray:
    """A lumpy boulder about 1 m across, as experiments/room3d/src/field.rs:30-58
    builds it (a sphere with low-frequency bumps, squashed vertically, on a
    `Surface::parametric` nu x nv grid, degenerate pole triangles dropped,
    wound outward), with the three bump amplitudes `coef`."""
    grid = _boulder_grid(coef, nu, nv)
    a, b, c, d = grid[:-1, :-1], grid[1:, :-1], grid[1:, 1:], grid[:-1, 1:]
    tris = np.stack([np.stack([a, b, c], axis=2), np.stack([a, c, d], axis=2)], axis=2).reshape(-1, 3, 3)
    e1, e2 = tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]
    scale = np.maximum(np.linalg.norm(e1, axis=1), np.linalg.norm(e2, axis=1))
    keep = np.linalg.norm(np.cross(e1, e2), axis=1) > 1e-6 * scale * scale
    tris = tris[keep]
    t = tris[0]
    if np.dot(np.cross(t[1] - t[0], t[2] - t[0]), t[0]) < 0:
        tris = tris[:, [0, 2, 1]]
    return np.ascontiguousarray(tris, dtype=F32)


def boulder_field(seed: int = 0xC0_0005, count: int = 1_000_000, kinds: int = 10, half: int = 500,
                  spacing: float = 2.0, lights: int = 32):
    """The instanced serving world: a ground quad about 2.1 km across
    (geometry 0) and `kinds` seeded boulder shapes (geometries 1..kinds,
    about 9,800 triangles each), placed `count` times on a jittered 2 m grid
    (the `count` nearest the camera, as field.rs:60-124 does), each turned
    about z and scaled 0.6-1.4. Lights: a 0.5-degree sun panel 3 km away, then
    rectangles of 1-4 m^2 scattered over the field 6-12 m up.

    Returns (geometries as one (n, 3, 3) array, triangle counts, placements
    as a PLACEMENT_DTYPE-shaped record array, lights, eye, target). Random
    numbers come from numpy's PCG64 stream `seed`."""
    rng = np.random.default_rng(seed)
    h = 1050.0
    ground = quad_pair([-h, -h, 0.0], [2 * h, 0, 0], [0, 2 * h, 0])
    shapes = [boulder(rng.uniform([0.12, 0.06, 0.02], [0.30, 0.18, 0.10])) for _ in range(kinds)]
    geoms = [ground] + shapes
    eye = np.array([-990.0, -990.0, 1.7])
    target = np.array([0.0, 0.0, 1.0])
    n = 2 * half
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    jit = rng.uniform(-0.6, 0.6, (n, n, 2))
    x = (i - half + 0.5) * spacing + jit[..., 0]
    y = (j - half + 0.5) * spacing + jit[..., 1]
    angle = rng.uniform(0, 2 * np.pi, (n, n))
    scale = rng.uniform(0.6, 1.4, (n, n))
    kind = rng.integers(1, kinds + 1, (n, n))
    pos = np.stack([x, y, 0.2 * scale], axis=-1).reshape(-1, 3)
    d2 = ((pos[:, :2] - eye[:2]) ** 2).sum(1)
    ok = np.sqrt(d2) >= 4.0
    order = np.argsort(np.where(ok, d2, np.inf), kind="stable")[: min(count, int(ok.sum()))]
    m = len(order) + 1
    from .kernel import PLACEMENT_DTYPE  # the library's record layout

    pl = np.zeros(m, dtype=PLACEMENT_DTYPE)
    pl["rotation"][0] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    pl["scale"][0] = 1.0
    ang, sc = angle.reshape(-1)[order], scale.reshape(-1)[order]
    c, s = np.cos(ang).astype(F32), np.sin(ang).astype(F32)
    z, o = np.zeros_like(c), np.ones_like(c)
    pl["geometry"][1:] = kind.reshape(-1)[order]
    pl["rotation"][1:] = np.stack([c, -s, z, s, c, z, z, z, o], axis=1)
    pl["scale"][1:] = sc
    pl["translation"][1:] = pos[order]
    # the sun: 30 degrees up, azimuth 200 degrees, 3 km out, 0.5 degrees across
    el, az = math.radians(30.0), math.radians(200.0)
    toward = np.array([math.cos(el) * math.cos(az), math.cos(el) * math.sin(az), math.sin(el)])
    dist = 3000.0
    size = 2.0 * dist * math.tan(math.radians(0.25))
    w = -toward
    a = np.cross([0.0, 0.0, 1.0], w)
    a /= np.linalg.norm(a)
    b = np.cross(w, a)
    center = toward * dist
    out_lights = [rect_light(center - (a + b) * size / 2, a * size, b * size, one_sided=True)]
    for _ in range(lights - 1):
        area = rng.uniform(1.0, 4.0)
        s_ = math.sqrt(area)
        cx, cy = rng.uniform(-990, 990, 2)
        cz = rng.uniform(6.0, 12.0)
        out_lights.append(rect_light([cx - s_ / 2, cy + s_ / 2, cz], [s_, 0, 0], [0, -s_, 0], one_sided=True,
                                     radiance=tuple(rng.uniform(2.0, 10.0, 3))))
    tris9, counts = meshes_to_arrays(geoms)
    return tris9, counts, pl, np.stack(out_lights), eye, target


def courtyard_frame(spec: SceneSpec, frame: int, tris12: np.ndarray, ranges: np.ndarray) -> np.ndarray:
    """The courtyard's triangles (12 f32 each, flattened order) at `frame`:
    curtain vertices displaced across the curtain by a seeded sine field of
    time. Other geometry is unchanged."""
    out = tris12.copy()
    t = 0.08 * frame
    seed = spec.notes.get("seed", 0)
    for m in spec.notes["curtains"]:
        first, count = int(ranges[m][0]), int(ranges[m][1])
        blk = out[first:first + count].reshape(count, 3, 4)
        x, z = blk[..., 0], blk[..., 2]
        amp = 0.15 * np.clip((15.8 - z) / 5.2, 0.0, 1.0)
        blk[..., 1] += (amp * np.sin(1.3 * x + 0.9 * z + t + 0.001 * seed)).astype(F32)
        out[first:first + count] = blk.reshape(count, 12)
    return out


def meshes_to_arrays(meshes) -> tuple[np.ndarray, np.ndarray]:
    """Concatenated triangles (n, 9) f32 and per-mesh counts (u32)."""
    counts = np.array([len(m) for m in meshes], dtype=np.uint32)
    if counts.sum() == 0:
        return np.zeros((0, 9), dtype=F32), counts
    tris = np.ascontiguousarray(np.concatenate([np.asarray(m, dtype=F32).reshape(-1, 9) for m in meshes]))
    return tris, counts


def boulder_surface(coef, nu: int = 100, nv: int = 50):
    """`boulder` with its surface: one parametric chart over the nu x nv
    grid (continuum_amd.surfaces.parametric), smooth normals; its triangles
    equal boulder(coef, nu, nv)'s."""
    from .surfaces import parametric
    grid = _boulder_grid(coef, nu, nv)
    tris = boulder(coef, nu, nv)
    a, b, c = grid[0, 0], grid[1, 0], grid[1, 1]
    plain = parametric(grid, flip=False)
    flip = not np.array_equal(plain.tris[0], tris[0])
    return plain if not flip else parametric(grid, flip=True)


def vcache_world(seed: int = 0xC0_0006, boulders: int = 6, boxes: int = 4, half: float = 10.0):
    """A small instanced world for the view cache: a 2*half m ground quad
    (geometry 0), a box (geometry 1) and a boulder (geometry 2, 24 x 12
    grid) placed `boxes` and `boulders` times on the ground, the first box
    and the first boulder movable; a rect light 6 m up and a sphere light.

    Returns (tris9, counts, placements, surface over the geometries,
    lights, movable objects). Seeded with numpy's PCG64 stream `seed`."""
    from .kernel import PLACEMENT_DTYPE
    from .surfaces import concat, from_triangles
    rng = np.random.default_rng(seed)
    ground = quad_pair([-half, -half, 0.0], [2 * half, 0, 0], [0, 2 * half, 0])
    box = box_mesh([-0.5, -0.4, 0.0], [0.5, 0.4, 0.9])
    rock = boulder_surface((0.2, 0.1, 0.05), 24, 12)
    geoms = [ground, box, rock.tris]
    surface = concat([from_triangles(ground), from_triangles(box), rock])
    tris9, counts = meshes_to_arrays(geoms)
    n = 1 + boxes + boulders
    pl = np.zeros(n, dtype=PLACEMENT_DTYPE)
    pl[0]["rotation"] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    pl[0]["scale"] = 1.0
    for k in range(1, n):
        g = 1 if k <= boxes else 2
        ang = rng.uniform(0, 2 * np.pi)
        c, s_ = np.cos(ang), np.sin(ang)
        pl[k]["geometry"] = g
        pl[k]["rotation"] = np.array([c, -s_, 0, s_, c, 0, 0, 0, 1], np.float32)
        pl[k]["scale"] = rng.uniform(0.7, 1.5)
        z = 0.0 if g == 1 else 0.25 * float(pl[k]["scale"])
        pl[k]["translation"] = [rng.uniform(-0.6 * half, 0.6 * half), rng.uniform(-0.6 * half, 0.6 * half), z]
    lights = np.stack([rect_light([-1.0, -1.0, 6.0], [2.0, 0, 0], [0, 2.0, 0], one_sided=False),
                       sphere_light([6.0, -4.0, 5.0], 0.5)])
    return tris9, counts, pl, surface, lights, [1, 1 + boxes]


def boulder_field_surface(seed: int = 0xC0_0005, kinds: int = 10):
    """The surfaces of boulder_field's geometries (same seed, same shapes):
    the ground quad's one chart, then each boulder's parametric chart, as
    one continuum_amd.surfaces.Surface with charts numbered globally (chart
    g is geometry g's)."""
    from .surfaces import concat, from_triangles
    rng = np.random.default_rng(seed)
    h = 1050.0
    ground = quad_pair([-h, -h, 0.0], [2 * h, 0, 0], [0, 2 * h, 0])
    coefs = [rng.uniform([0.12, 0.06, 0.02], [0.30, 0.18, 0.10]) for _ in range(kinds)]
    return concat([from_triangles(ground)] + [boulder_surface(c) for c in coefs])
