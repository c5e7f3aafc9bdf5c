"""The oracle's interpolated shading normal for camera receivers
(crates/continuum-gpu/src/mask_cones.wgsl :2047-2058 `shading_normal`, with
`barycentric` :1780-1793; `per_pixel_hit_color` :2280 shades with it) against
what fixes it independently:

* at every pixel of a view of one triangle with three different corner
  normals, the normal is the f64 barycentric blend of the corners' normals
  at the hit point, normalized;
* corner normals that sum to nothing give the face normal;
* on a geodesic sphere with its exact corner normals (the radial
  directions) the shading normal is within the tessellation's error of the
  radial direction at the hit point, and much closer to it than the face
  normal is;
* a rotated, scaled placement turns the object-space normals into the world.
"""
import math

import numpy as np

from continuum_amd.kernel import PLACEMENT_DTYPE, Flat
from continuum_amd.scenes import _icosphere_grid, camera_uniform, meshes_to_arrays
from oracle import oracle as O

W, H = 48, 40


def _hit_points(cam, hits):
    """World hit points from the hits' distances (f64), per pixel."""
    eye = cam[0:3].astype(np.float64)
    f, r, u = (cam[4:7].astype(np.float64), cam[8:11].astype(np.float64), cam[12:15].astype(np.float64))
    wpx, hpx, tanh = float(cam[16]), float(cam[17]), float(cam[18])
    y, x = np.divmod(np.arange(W * H), W)
    sx = ((x + 0.5) / wpx * 2.0 - 1.0) * tanh * (wpx / hpx)
    sy = (1.0 - (y + 0.5) / hpx * 2.0) * tanh
    d = f[None] + r[None] * sx[:, None] + u[None] * sy[:, None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dist = hits[:, 2].view(np.float32).astype(np.float64)
    return eye[None] + d * dist[:, None]


def _one_triangle():
    tri = np.array([[[-1.0, -1.0, 0.0], [1.5, -1.0, 0.0], [-1.0, 1.5, 0.0]]], np.float32)
    corner = np.array([[0.0, 0.0, 1.0], [0.6, 0.0, 0.8], [0.0, -0.6, 0.8]], np.float64)
    return tri, corner


def test_the_normal_is_the_barycentric_blend_of_the_corners():
    tri, corner = _one_triangle()
    t9, c = meshes_to_arrays([tri])
    flat = Flat.build_arrays(t9, c)
    cam = camera_uniform([0.1, 0.2, 3.0], [0.0, 0.0, 0.0], 60.0, W, H)
    nrm = corner.astype(np.float32).reshape(1, 9)
    hits, recv, _ = O.primary(flat, cam, W, H, np.zeros((0, 16), np.float32), normals=nrm)
    lit = recv[:, 7] >= 0
    assert lit.sum() > 300
    p = _hit_points(cam, hits)[lit]
    v = tri[0].astype(np.float64)
    e1, e2 = v[1] - v[0], v[2] - v[0]
    d = p - v[0]
    m = np.array([[e1 @ e1, e1 @ e2], [e1 @ e2, e2 @ e2]])
    rhs = np.stack([d @ e1, d @ e2], axis=1)
    su = np.linalg.solve(m, rhs.T).T
    b = np.stack([1 - su[:, 0] - su[:, 1], su[:, 0], su[:, 1]], axis=1)
    want = b @ corner
    want /= np.linalg.norm(want, axis=1, keepdims=True)
    got = recv[lit, 4:7].astype(np.float64)
    assert np.abs(got - want).max() < 2e-5
    # the shading point itself still sits on the face normal's side
    flat_hits, flat_recv, _ = O.primary(flat, cam, W, H, np.zeros((0, 16), np.float32))
    assert np.array_equal(flat_hits, hits)
    assert np.array_equal(flat_recv[:, 0:3], recv[:, 0:3])
    assert np.allclose(flat_recv[lit, 4:7], [0.0, 0.0, 1.0], atol=1e-7)


def test_normals_that_sum_to_nothing_give_the_face_normal():
    tri, _ = _one_triangle()
    t9, c = meshes_to_arrays([tri])
    flat = Flat.build_arrays(t9, c)
    cam = camera_uniform([0.1, 0.2, 3.0], [0.0, 0.0, 0.0], 60.0, W, H)
    nrm = np.zeros((1, 9), np.float32)
    _, recv, _ = O.primary(flat, cam, W, H, np.zeros((0, 16), np.float32), normals=nrm)
    _, flat_recv, _ = O.primary(flat, cam, W, H, np.zeros((0, 16), np.float32))
    assert np.array_equal(recv, flat_recv)


def test_a_sphere_with_exact_normals_shades_radially():
    unit = _icosphere_grid(6)  # 720 triangles
    sphere = (unit * 2.0).astype(np.float32)
    t9, c = meshes_to_arrays([sphere])
    flat = Flat.build_arrays(t9, c)
    nrm = unit.astype(np.float32).reshape(-1, 9)  # radial corner normals
    cam = camera_uniform([0.3, -5.0, 1.2], [0.0, 0.0, 0.0], 45.0, W, H)
    hits, recv, _ = O.primary(flat, cam, W, H, np.zeros((0, 16), np.float32), normals=nrm)
    _, flat_recv, _ = O.primary(flat, cam, W, H, np.zeros((0, 16), np.float32))
    lit = recv[:, 7] >= 0
    assert lit.sum() > 500
    radial = _hit_points(cam, hits)[lit]
    radial /= np.linalg.norm(radial, axis=1, keepdims=True)
    err_shade = np.degrees(np.arccos(np.clip((recv[lit, 4:7] * radial).sum(1), -1, 1)))
    err_face = np.degrees(np.arccos(np.clip((flat_recv[lit, 4:7] * radial).sum(1), -1, 1)))
    # a frequency-6 geodesic sphere's faces span about 12 degrees
    assert err_shade.max() < 1.5 and err_face.max() > 4.0
    assert err_shade.mean() < 0.25 * err_face.mean()


def test_a_placement_turns_the_normals_into_the_world():
    tri, corner = _one_triangle()
    t9, c = meshes_to_arrays([tri])
    a = math.radians(35.0)
    rot = np.array([[math.cos(a), 0.0, math.sin(a)], [0.0, 1.0, 0.0], [-math.sin(a), 0.0, math.cos(a)]])
    pl = np.zeros(1, dtype=PLACEMENT_DTYPE)
    pl[0]["geometry"] = 0
    pl[0]["rotation"] = rot.astype(np.float32).reshape(9)
    pl[0]["scale"] = 1.7
    pl[0]["translation"] = [0.5, -0.2, 0.3]
    o = O.OracleInstancedFlat(t9, c, pl, object_prefilter=False)
    cam = camera_uniform([2.5, 0.3, 3.2], [0.5, -0.2, 0.3], 60.0, W, H)
    nrm = np.tile(np.array([0.0, 0.0, 1.0], np.float32), 3).reshape(1, 9)  # the face's own normal at every corner
    _, recv, _ = O.primary(o, cam, W, H, np.zeros((0, 16), np.float32), normals=nrm)
    lit = recv[:, 7] >= 0
    assert lit.sum() > 200
    want = rot @ np.array([0.0, 0.0, 1.0])
    assert np.abs(recv[lit, 4:7] - want[None]).max() < 1e-6
