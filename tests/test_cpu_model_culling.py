"""CPU: the oracle's restatement of the reference CPU model's culling
structures (test infrastructure for the cone3.rs model):

* `Bvh` (crates/continuum/src/bvh.rs:25-165): its cone walk's candidates
  equal brute force, mesh by mesh (the reference's own
  `candidates_match_brute_force`, bvh.rs:173-214, transcribed, and the same
  check from many more apexes, triangle sets compared, not only counts);
* `TileFrustum` (experiments/room3d/src/lib.rs:121-171): its planes in
  closed form (each side plane holds the tile's edge and the light's
  matching edge), every segment from the tile to the light inside, and its
  box and triangle tests against geometry inside and outside.
"""
import math

import numpy as np

from continuum_amd.scenes import box_mesh
from oracle import oracle as O

F32 = np.float32


def sphere_mesh(center, radius, segments, rings):
    """cone3.rs:420-441: a UV sphere, outward-facing."""
    c = np.asarray(center, np.float64)

    def at(s, r):
        theta, phi = 2 * math.pi * s / segments, math.pi * r / rings
        return c + np.array([math.sin(phi) * math.cos(theta), math.sin(phi) * math.sin(theta), math.cos(phi)]) * radius

    tris = []
    for r in range(rings):
        for s in range(segments):
            a, b, cc, d = at(s, r), at(s + 1, r), at(s + 1, r + 1), at(s, r + 1)
            if r > 0:
                tris.append([a, d, b])
            if r + 1 < rings:
                tris.append([b, d, cc])
    return np.array(tris, F32)


def _bvh_scene():
    meshes = []
    for i in range(20):
        x = 0.05 * i
        meshes.append(box_mesh([x, 0.3, 0.4 + 0.01 * i], [x + 0.03, 0.7, 0.42 + 0.01 * i]))
    meshes.append(sphere_mesh([0.5, 0.5, 0.6], 0.1, 24, 12))
    return meshes


LIGHT = np.array([0.4, 0.4, 1.0, 0, 0.2, 0, 0, 0, 0, 0.2, 0, 0, 1, 1, 1, 0], F32)


def _brute(meshes, apex):
    per, sets, off = [], [], 0
    for m in meshes:
        ok = O.cone3_may_touch(np.asarray(m, F32).reshape(-1, 9), apex, LIGHT)
        per.append(int(ok.sum()))
        sets.append(set((off + np.nonzero(ok)[0]).tolist()))
        off += len(m)
    return per, sets


def test_candidates_match_brute_force():
    """bvh.rs:173-214 from its three apexes."""
    meshes = _bvh_scene()
    for apex in ([0.5, 0.5, 0.0], [0.1, 0.9, 0.0], [0.8, 0.2, 0.0]):
        per, found, visited = O.bvh_candidates(meshes, apex, LIGHT)
        want, _ = _brute(meshes, apex)
        assert per.tolist() == want, apex
        assert sum(want) > 0 and visited > 0


def test_candidates_are_the_brute_force_sets_from_many_apexes():
    meshes = _bvh_scene()
    off = np.cumsum([0] + [len(m) for m in meshes])
    total_nodes = 2 * math.ceil(sum(len(m) for m in meshes) / 4)
    rng = np.random.default_rng(21)
    pruned = 0
    for _ in range(60):
        apex = rng.uniform([0.0, 0.0, 0.0], [1.0, 1.0, 0.35])
        res = O.bvh_candidates(meshes, apex, LIGHT)
        assert res is not None
        per, found, visited = res
        want_per, want_sets = _brute(meshes, apex)
        assert per.tolist() == want_per
        # grouped by mesh, in mesh order, each triangle once
        mesh_of = np.searchsorted(off, found, side="right") - 1
        assert np.all(np.diff(mesh_of) >= 0)
        assert len(set(found.tolist())) == len(found)
        for m, want in enumerate(want_sets):
            assert set(found[mesh_of == m].tolist()) == want
        pruned += visited < total_nodes
    assert pruned > 30  # the walk skips subtrees for most cones


def test_a_cone_that_meets_nothing_has_no_candidates():
    meshes = _bvh_scene()
    per, found, _ = O.bvh_candidates(meshes, [5.0, 5.0, 0.0], LIGHT)
    assert per.sum() == 0 and len(found) == 0
    # an apex in the light's plane has no cone (cone3.rs:197-203)
    assert O.bvh_candidates(meshes, [0.5, 0.5, 1.0], LIGHT) is None


def test_tile_frustum_planes_in_closed_form():
    tile = [0.25, 0.5, 0.75, 0.875]
    light = np.array([0.4, 0.45, 1.0, 0, 0.2, 0, 0, 0, 0, 0.25, 0, 0], F32)
    planes, _, _ = O.tile_frustum(tile, light)
    lx0, ly0, lx1, ly1, h = 0.4, 0.45, 0.6, 0.7, 1.0
    # side planes hold the tile edge at z = 0 and the light edge at z = h
    holds = [
        (0, [(tile[0], 0.3, 0.0), (lx0, 0.5, h)]),
        (1, [(tile[2], 0.9, 0.0), (lx1, 0.1, h)]),
        (2, [(0.1, tile[1], 0.0), (0.7, ly0, h)]),
        (3, [(0.3, tile[3], 0.0), (0.2, ly1, h)]),
        (4, [(0.3, 0.3, 0.0), (7.0, -2.0, 0.0)]),
        (5, [(0.3, 0.3, h), (-4.0, 9.0, h)]),
    ]
    for k, pts in holds:
        n, o = planes[k, :3].astype(np.float64), float(planes[k, 3])
        for p in pts:
            assert abs(n @ np.array(p) - o) <= 1e-6, (k, p)
    # every segment from a tile point to a light point is inside every plane
    rng = np.random.default_rng(4)
    a = np.column_stack([rng.uniform(tile[0], tile[2], 500), rng.uniform(tile[1], tile[3], 500), np.zeros(500)])
    b = np.column_stack([rng.uniform(lx0, lx1, 500), rng.uniform(ly0, ly1, 500), np.full(500, h)])
    t = rng.uniform(0, 1, (500, 1))
    pts = a + (b - a) * t
    assert np.all(pts @ planes[:, :3].T.astype(np.float64) <= planes[:, 3].astype(np.float64) + 1e-6)


def test_tile_frustum_box_and_triangle_tests():
    tile = [0.25, 0.5, 0.75, 0.875]
    light = np.array([0.4, 0.45, 1.0, 0, 0.2, 0, 0, 0, 0, 0.25, 0, 0], F32)
    rng = np.random.default_rng(8)
    # small boxes around points inside the prismatoid: all kept
    a = np.column_stack([rng.uniform(tile[0], tile[2], 300), rng.uniform(tile[1], tile[3], 300), np.zeros(300)])
    b = np.column_stack([rng.uniform(0.4, 0.6, 300), rng.uniform(0.45, 0.7, 300), np.ones(300)])
    p = a + (b - a) * rng.uniform(0.05, 0.95, (300, 1))
    inside = np.column_stack([p - 0.01, p + 0.01])
    # boxes wholly above the light, below the floor, or far to a side: culled
    outside = np.array([[0, 0, 1.1, 1, 1, 1.3], [0, 0, -0.5, 1, 1, -0.1], [2, 0, 0, 3, 1, 1], [0, -3, 0, 1, -2, 1],
                        [-3, 0, 0.2, -2, 1, 0.4], [0, 3, 0.2, 1, 4, 0.4]], F32)
    _, kept, _ = O.tile_frustum(tile, light, boxes=np.concatenate([inside, outside]))
    assert kept[:300].all()
    assert not kept[300:].any()
    # triangles: one spanning the column at mid height kept; one beside it culled
    tris = np.array([[0.3, 0.6, 0.5, 0.7, 0.6, 0.5, 0.5, 0.8, 0.5],
                     [2.0, 0.6, 0.5, 2.4, 0.6, 0.5, 2.2, 0.8, 0.5]], F32)
    _, _, tk = O.tile_frustum(tile, light, tris9=tris)
    assert tk.tolist() == [True, False]
