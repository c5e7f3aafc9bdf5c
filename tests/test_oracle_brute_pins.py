"""CPU: pins of the oracle's camera pass and of the builder's split rule,
by brute force and by the rules themselves, computed here without the
oracle's code.

* Camera pass (`oracle_primary`, following WGSL `camera_dir` :1724,
  `ray_triangle` :1539, `hit_triangle` :1575-1600): every pixel's ray is
  rebuilt in f64 from the 20-float camera uniform (lib.rs:3538-3553: the
  pixel centre, x right, y down) and intersected with every front-facing
  triangle; the oracle's (mesh, triangle) must be the nearest hit, up to a
  stated distance tie, and its distance must agree. Constructed scenes pin
  the tie rules: equal distances go to the lower mesh, then the lower
  triangle; back faces are invisible to camera rays; and the result does not
  depend on the traversal (no hierarchy, mesh hierarchy, triangle
  hierarchies; nearest-first or fixed child order).
* Split rule (`split`, hierarchy.rs:128-179): at every internal node of
  every triangle hierarchy the items divide at len / 2, on the longest axis
  of their padded boxes' centroids (ties to x, then y), with every left key
  (lo + hi on that axis) at most every right key; every node's box is the
  union of its items' padded boxes (`padded`, hierarchy.rs:762).
"""
import numpy as np
import pytest

from continuum_amd.kernel import Flat
from continuum_amd.scenes import box_mesh, camera_uniform, grid_quad, meshes_to_arrays, rock_field, room_scene
from oracle import oracle as O

F32 = np.float32
NONE = 0xFFFFFFFF


# ---- camera pass

def _rays(cam, w, h, pixels):
    """f64 directions of the given pixels' camera rays (WGSL :1724)."""
    c = cam.astype(np.float64)
    eye, fwd, right, up = c[0:3], c[4:7], c[8:11], c[12:15]
    rw, uw = c[11], c[15]
    W, H, t = c[16], c[17], c[18]
    x, y = pixels % w, pixels // w
    sx = ((x + 0.5 + rw) / W * 2.0 - 1.0) * t * (W / H)
    sy = (1.0 - (y + 0.5 + uw) / H * 2.0) * t
    d = fwd[None] + right[None] * sx[:, None] + up[None] * sy[:, None]
    return eye, d / np.linalg.norm(d, axis=1, keepdims=True)


def _brute(eye, dirs, tris12, tol):
    """Per ray: every front-facing triangle it meets within barycentric
    tolerance `tol` (Moller-Trumbore in f64): (t, u, v) arrays (n_rays,
    n_tris), t = inf where missed."""
    v = tris12.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    v0, e1, e2 = v[:, 0], v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    n = np.cross(e1, e2)
    s = eye[None] - v0
    q = np.cross(s, e1)
    out = np.full((len(dirs), len(v)), np.inf)
    for k, d in enumerate(dirs):
        p = np.cross(d[None], e2)
        det = np.einsum("ij,ij->i", e1, p)
        ok = (np.abs(det) > 1e-18) & (n @ d <= 0.0)
        inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
        u = np.einsum("ij,ij->i", s, p) * inv
        vv = (q @ d) * inv
        t = np.einsum("ij,ij->i", e2, q) * inv
        hit = ok & (u >= -tol) & (u <= 1 + tol) & (vv >= -tol) & (u + vv <= 1 + tol) & (t > 0)
        out[k] = np.where(hit, t, np.inf)
    return out


def _tri_mesh(flat):
    mesh = np.zeros(len(flat.tris), np.int64)
    for m, (first, count, _, _) in enumerate(flat.ranges):
        mesh[first:first + count] = m
    return mesh


def _check_view(flat, cam, w, h, pixels, min_exact):
    hits, _, _ = O.primary(flat, cam, w, h, np.zeros((0, 16), np.float32))
    eye, dirs = _rays(cam, w, h, pixels)
    strict = _brute(eye, dirs, flat.tris, -1e-4)   # hits well inside a triangle
    loose = _brute(eye, dirs, flat.tris, 1e-4)     # hits within a generous edge tolerance
    tmesh = _tri_mesh(flat)
    exact = 0
    for k, p in enumerate(pixels):
        m, t, dbits = int(hits[p, 0]), int(hits[p, 1]), int(hits[p, 2])
        best = float(strict[k].min())
        if m == NONE:
            # no hit: no triangle holds the ray well inside it
            assert not np.isfinite(best), f"pixel {p}: the oracle sees nothing, brute force hits {strict[k].argmin()}"
            exact += 1
            continue
        assert tmesh[t] == m, f"pixel {p}: triangle {t} is not in mesh {m}"
        dist = float(np.array([dbits], np.uint32).view(np.float32)[0])
        assert np.isfinite(loose[k, t]), f"pixel {p}: the oracle's triangle {t} is not on the ray"
        assert abs(dist - loose[k, t]) <= 1e-5 * loose[k, t], (p, dist, loose[k, t])
        # nothing strictly nearer than the oracle's hit, beyond a tie
        if np.isfinite(best):
            assert loose[k, t] <= best * (1 + 1e-5), f"pixel {p}: a nearer triangle {strict[k].argmin()} exists"
            exact += int(strict[k].argmin() == t)
        else:
            exact += 1
    assert exact >= min_exact * len(pixels), f"only {exact} of {len(pixels)} pixels pick the brute force's nearest"


def test_camera_pass_is_the_nearest_front_face_in_the_room():
    sc = room_scene(width=48, height=48)
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = O.OracleFlat(tris9, counts)
    _check_view(flat, sc.camera, 48, 48, np.arange(48 * 48), 0.98)


def test_camera_pass_is_the_nearest_front_face_on_the_rock_field():
    sc = rock_field(width=160, height=90, rocks=12)
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = O.OracleFlat(tris9, counts)
    pixels = np.arange(0, 160 * 90, 97)
    _check_view(flat, sc.camera, 160, 90, pixels, 0.97)


def _card_scene():
    """A wall (mesh 0), then two copies of one floor quad (meshes 1, 2),
    a mesh holding one quad twice (mesh 3, triangles repeated), and a card
    turned away from the camera in front of the wall (mesh 4)."""
    wall = grid_quad([0, 4, 0], [4, 0, 0], [0, 0, 3], 4, 3)           # faces -y, toward the camera
    floor = grid_quad([0, 0, 0], [4, 0, 0], [0, 4, 0], 2, 2)          # faces +z
    dup = np.concatenate([grid_quad([0.5, 1.0, 0.5], [1, 0, 0], [0, 1, 0], 1, 1)] * 2)
    # the card's winding reversed, so it faces +y (away from a camera at y = -3)
    card = grid_quad([2.5, 2.0, 0.6], [1, 0, 0], [0, 0, 1], 1, 1)[:, [0, 2, 1]]
    return [wall, floor, floor.copy(), dup, card]


def _facing(tri, d):
    n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
    return float(n @ d)


def test_camera_tie_rules_on_constructed_faces():
    meshes = _card_scene()
    cam = camera_uniform([2.0, -3.0, 2.0], [2.0, 2.0, 0.6], 60.0, 64, 48)
    d = np.array([0.0, 1.0, 0.0])
    assert _facing(meshes[0][0].astype(np.float64), d) < 0, "the wall faces the camera"
    assert _facing(meshes[4][0].astype(np.float64), d) > 0, "the card faces away"
    tris9, counts = meshes_to_arrays(meshes)
    flat = O.OracleFlat(tris9, counts)
    hits, _, _ = O.primary(flat, cam, 64, 48, np.zeros((0, 16), np.float32))
    seen = set(hits[:, 0].tolist())
    assert 2 not in seen, "a floor pixel went to the higher of two identical meshes"
    assert 1 in seen and 0 in seen
    assert 4 not in seen, "camera rays see back faces"
    first3 = int(flat.ranges[3][0])
    on_dup = hits[hits[:, 0] == 3]
    assert len(on_dup) > 0
    # mesh 3 holds triangles [a, b, a, b]: only the first copy is ever hit
    assert set(on_dup[:, 1].tolist()) <= {first3, first3 + 1}, "a repeated triangle won over its first copy"
    # identical meshes swapped: still the lower mesh index
    meshes2 = [meshes[0], meshes[2], meshes[1], meshes[3], meshes[4]]
    t2, c2 = meshes_to_arrays(meshes2)
    h2, _, _ = O.primary(O.OracleFlat(t2, c2), cam, 64, 48, np.zeros((0, 16), np.float32))
    assert 2 not in set(h2[:, 0].tolist())


@pytest.mark.parametrize("scene", ["room", "cards"])
def test_camera_hits_do_not_depend_on_the_traversal(scene):
    if scene == "room":
        sc = room_scene(width=64, height=64)
        meshes, cam, w, h = sc.meshes, sc.camera, 64, 64
    else:
        meshes = _card_scene()
        cam, w, h = camera_uniform([2.0, -3.0, 2.0], [2.0, 2.0, 0.6], 60.0, 64, 48), 64, 48
    tris9, counts = meshes_to_arrays(meshes)
    flat = O.OracleFlat(tris9, counts)
    ref = None
    for hierarchy in (0, 1, 2):
        for nearest in (0, 1):
            got, _, _ = O.primary(flat, cam, w, h, np.zeros((0, 16), np.float32), hierarchy=hierarchy,
                                  nearest_first=nearest)
            if ref is None:
                ref = got
            assert np.array_equal(got, ref), (hierarchy, nearest)


# ---- split rule

def _padded_boxes(tris12):
    v = tris12.reshape(-1, 3, 4)[:, :, :3].astype(F32)
    lo, hi = v.min(axis=1), v.max(axis=1)
    pad = F32(1e-5) * np.maximum(np.maximum(np.abs(lo), np.abs(hi)).max(axis=1), F32(1.0))
    return (lo - pad[:, None]).astype(F32), (hi + pad[:, None]).astype(F32)


def _check_splits(flat):
    lo, hi = _padded_boxes(flat.tris)
    blas = flat.blas
    nlo, nhi = blas[:, 0:3].view(np.float32), blas[:, 4:7].view(np.float32)
    checked = 0
    for first, count, root, _ in flat.ranges:
        if root == NONE:
            continue

        def items_of(n):
            a, c = int(blas[n, 8]), int(blas[n, 9])
            if c > 0:
                return list(flat.blas_items[a:a + c])
            return items_of(a) + items_of(a + 1)

        stack = [int(root)]
        while stack:
            n = stack.pop()
            its = np.array(items_of(n), np.int64)
            assert np.array_equal(nlo[n], lo[its].min(axis=0)) and np.array_equal(nhi[n], hi[its].max(axis=0)), n
            a, c = int(blas[n, 8]), int(blas[n, 9])
            if c > 0:
                assert len(its) <= 8
                continue
            assert len(its) > 8
            cen = ((lo[its] + hi[its]) * F32(0.5)).astype(F32)
            e = cen.max(axis=0) - cen.min(axis=0)
            axis = 0 if (e[0] >= e[1] and e[0] >= e[2]) else (1 if e[1] >= e[2] else 2)
            left, right = np.array(items_of(a)), np.array(items_of(a + 1))
            assert len(left) == len(its) // 2 and len(right) == len(its) - len(its) // 2, n
            kl = (lo[left, axis] + hi[left, axis]).astype(F32)
            kr = (lo[right, axis] + hi[right, axis]).astype(F32)
            assert kl.max() <= kr.min(), f"node {n}: a left key exceeds a right key on axis {axis}"
            # the other axes would not have split this way unless their keys happen to agree
            checked += 1
            stack += [a, a + 1]
    return checked


def test_split_rule_at_every_internal_node():
    rng = np.random.default_rng(3)
    meshes = room_scene().meshes + [
        rng.uniform(-5, 5, (777, 3, 3)).astype(F32),
        np.repeat(box_mesh([2, 2, 2], [2.5, 2.5, 2.5])[:2], 20, axis=0),    # many equal keys
        (rng.uniform(0, 1, (300, 3, 3)) * [8.0, 0.5, 0.2]).astype(F32),     # long in x
        (rng.uniform(0, 1, (300, 3, 3)) * [0.3, 0.3, 6.0]).astype(F32),     # long in z
    ]
    tris9, counts = meshes_to_arrays(meshes)
    for flat in (O.OracleFlat(tris9, counts), Flat.build_arrays(tris9, counts)):
        assert _check_splits(flat) > 100


def test_split_rule_catches_a_wrong_axis():
    """The check fails on a hierarchy split on the wrong axis: the mesh is
    long in x, so a split on z breaks the key order or the axis choice."""
    rng = np.random.default_rng(9)
    mesh = (rng.uniform(0, 1, (64, 3, 3)) * [8.0, 0.5, 0.2]).astype(F32)
    tris9, counts = meshes_to_arrays([mesh])
    flat = O.OracleFlat(tris9, counts)
    lo, hi = _padded_boxes(flat.tris)
    root_items = np.array(flat.blas_items)
    bad = Flat(flat.tris, flat.ranges, flat.blas.copy(), flat.blas_items.copy(), flat.blas_cov, flat.nodes,
               flat.items)
    # re-order the root's items by their z keys, keeping the tree shape
    order = np.argsort((lo[root_items, 2] + hi[root_items, 2]), kind="stable")
    bad.blas_items[:] = root_items[order]
    with pytest.raises(AssertionError):
        _check_splits(bad)
