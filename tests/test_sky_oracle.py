"""CPU: the oracle's sky light (WGSL :1323-1510: six cube-face mask cones,
meshes combined as independent occluders, the receiver's horizon culling
the walk) pinned against what fixes it independently of its code: an open
sky gives pi (Lambert's projected solid angle of a hemisphere) at any
normal; a square overhead whose edges fall on bin boundaries blocks exactly
its projected solid angle (the polygon form factor, in f64); a closed room
blocks everything; two objects over the same bins combine as
1 - (1 - a)(1 - b), and as max(a, b) with the maximum diagnostic; geometry
wholly below the horizon is never walked with the horizon cull and is
without it."""
import math

import numpy as np
import pytest

from continuum_amd.scenes import grid_quad, meshes_to_arrays, room_scene
from oracle import oracle as O

F32 = np.float32
OPTS = O.opts_array(hierarchy=2, prefilter=0.0)


def _flat(meshes):
    tris9, counts = meshes_to_arrays(meshes)
    return O.OracleFlat(tris9, counts)


def _far_away():
    """A triangle far below, facing away from every test point: culled, so
    the sky is open."""
    return np.array([[[0, 0, -1000], [0, 1, -1000], [1, 0, -1000]]], F32)


def _recv(point, normal):
    r = np.zeros((1, 8), F32)
    r[0, 0:3] = point
    r[0, 4:7] = np.asarray(normal, np.float64) / np.linalg.norm(normal)
    return r


def _form_factor(point, normal, poly):
    """Lambert's projected solid angle of a planar polygon seen from `point`
    (f64): half the sum over edges of the angle they subtend times the
    normal's component along the edge plane's normal."""
    n = np.asarray(normal, np.float64)
    n = n / np.linalg.norm(n)
    d = [np.asarray(v, np.float64) - np.asarray(point, np.float64) for v in poly]
    d = [v / np.linalg.norm(v) for v in d]
    s = 0.0
    for i in range(len(d)):
        a, b = d[i], d[(i + 1) % len(d)]
        g = np.cross(a, b)
        s += math.atan2(np.linalg.norm(g), float(np.dot(a, b))) * float(np.dot(n, g / np.linalg.norm(g)))
    return abs(0.5 * s)


@pytest.mark.parametrize("normal", [(0, 0, 1), (1, 2, 3), (-0.3, 0.1, -1.0), (1, 0, 0)])
def test_open_sky_is_pi(normal):
    flat = _flat([_far_away()])
    total, faces, full = O.sky(flat, _recv((0.3, -0.2, 0.1), normal), OPTS)
    assert abs(float(total[0]) - math.pi) < 2e-5 * math.pi
    assert np.all(faces >= 0) and np.all(full == 0)
    # faces wholly below the horizon give nothing
    n = np.asarray(normal, np.float64)
    for f in range(6):
        axis, sign = f // 2, (-1.0 if f & 1 else 1.0)
        corners_below = sign * n[axis] < -np.sum(np.abs(np.delete(n, axis)))
        if corners_below:
            assert faces[f, 0] == 0.0


def test_square_overhead_blocks_its_form_factor():
    """A square at height 1 over [-0.5, 0.5]^2 facing the point: in the +z
    face its projection spans bins 2..5 on both axes exactly (tan = +-0.5),
    so every bin is wholly covered or empty and the sky left is pi less the
    square's form factor up to rounding."""
    sq = grid_quad([-0.5, -0.5, 1.0], [0, 1, 0], [1, 0, 0], 4, 4)
    flat = _flat([sq])
    total, faces, full = O.sky(flat, _recv((0, 0, 0), (0, 0, 1)), OPTS)
    ff = _form_factor((0, 0, 0), (0, 0, 1), [(-0.5, -0.5, 1), (0.5, -0.5, 1), (0.5, 0.5, 1), (-0.5, 0.5, 1)])
    assert abs(float(total[0]) - (math.pi - ff)) < 2e-5 * math.pi
    # the +z face's full bins are the 4 x 4 block of columns and rows 2..5
    want = 0
    for row in range(2, 6):
        for col in range(2, 6):
            want |= 1 << (row * 8 + col)
    assert int(full[4, 0]) == want
    assert np.all(full[[0, 1, 2, 3, 5], 0] == 0)


def test_closed_room_blocks_the_sky():
    """The room's six walls as one mesh (inward, 32 x 32 quads each): every
    bin of every face is wholly covered, so nothing is left. (As six
    objects, bins split between two walls leak: independent occluders give
    1 - (1 - a)(1 - b) < 1 for a + b = 1, as the reference's sky does.)"""
    spec = room_scene()
    walls = np.concatenate([np.asarray(m, F32).reshape(-1, 3, 3) for m in spec.meshes[:6]])
    flat = _flat([walls])
    recv = np.concatenate([_recv((2.0, 2.0, 1.5), (0, 0, 1)), _recv((1.0, 3.0, 0.5), (0.3, -0.2, 1.0)),
                           _recv((3.5, 0.7, 2.5), (-1, 0, 0))])
    total, faces, full = O.sky(flat, recv, OPTS, O.sky_options(horizon_cull=False))
    assert np.all(total < 1e-5), total
    # With the horizon cull, triangles wholly below the horizon are not
    # walked: exact where the horizon runs along bin edges (axis-aligned
    # normals); a tilted normal's bins that straddle it keep only their
    # upper part's coverage against their clipped weight, as the reference.
    total, _, _ = O.sky(flat, recv, OPTS)
    assert total[0] < 1e-5 and total[2] < 1e-5 and 1e-5 < total[1] < 0.05, total


def _strip(x0, x1, z=1.0):
    """A strip x0..x1 (y in -0.5..0.5) at height z facing down."""
    return grid_quad([x0, -0.5, z], [0, 1.0, 0], [x1 - x0, 0, 0], 2, 2)


def test_independent_occluders_combine_as_one_minus_the_product():
    """Objects A (x -0.5..-0.15) and B (x -0.05..0.5) do not overlap but share
    the +z face's bin column [-0.25, 0): A covers 0.4 of its bins and B 0.2.
    Independent occluders give 1 - 0.6 * 0.8 = 0.52 there, the maximum 0.4."""
    a, b = _strip(-0.5, -0.15), _strip(-0.05, 0.5)
    p, n = (0, 0, 0), (0, 0, 1)
    _, ca = O.sky_face(_flat([a]), p, n, 4, OPTS)
    _, cb = O.sky_face(_flat([b]), p, n, 4, OPTS)
    _, cab = O.sky_face(_flat([a, b]), p, n, 4, OPTS)
    _, cmax = O.sky_face(_flat([a, b]), p, n, 4, OPTS, O.sky_options(max_combine=True))
    shared = [row * 8 + 3 for row in range(2, 6)]
    assert np.allclose(ca[shared], 0.4, atol=1e-5) and np.allclose(cb[shared], 0.2, atol=1e-5)
    assert np.allclose(cab, 1.0 - (1.0 - ca) * (1.0 - cb), atol=1e-6)
    assert np.allclose(cab[shared], 0.52, atol=1e-5)
    assert np.allclose(cmax, np.maximum(ca, cb), atol=1e-6)


def test_horizon_cull_skips_geometry_below_the_horizon():
    """A strip in the +x face wholly below a tilted receiver's horizon but in
    bins the horizon crosses: with the cull the sky equals the open sky bit
    for bit; without it the strip's coverage counts and takes light away."""
    normal = np.array([0.3, 0.0, 1.0])
    # the horizon crosses the +x face at z / x = -0.3, inside the bin row
    # [-0.5, -0.25); the strip's directions (2, y, z), z in -0.9..-0.7, lie
    # in that row wholly below it
    strip = grid_quad([2.0, -0.5, -0.9], [0, 0, 0.2], [0, 1.0, 0], 2, 2)
    strip_n = np.cross(strip[0, 1] - strip[0, 0], strip[0, 2] - strip[0, 0])
    assert strip_n[0] < 0  # faces the point at the origin
    assert np.all(strip.reshape(-1, 3) @ normal < 0)  # wholly below the horizon
    r = _recv((0, 0, 0), normal)
    open_total, open_faces, _ = O.sky(_flat([_far_away()]), r, OPTS)
    culled, culled_faces, _ = O.sky(_flat([strip, _far_away()]), r, OPTS)
    walked, walked_faces, _ = O.sky(_flat([strip, _far_away()]), r, OPTS, O.sky_options(horizon_cull=False))
    assert culled.view(np.uint32)[0] == open_total.view(np.uint32)[0]
    assert walked_faces[0, 0] < open_faces[0, 0] - 1e-4
