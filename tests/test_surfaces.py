"""CPU: the view cache's surface inputs (continuum::mesh::Surface,
crates/continuum/src/mesh.rs:20-376) as continuum_amd.surfaces restates
them: a box's faces are six parallelogram charts (mesh.rs:63-102), a
parametric surface is one chart over its grid with its triangles' chart
coordinates on the grid (mesh.rs:224-284), smooth normals sum the face
normals facing each corner's side in triangle order (mesh.rs:330-376,
checked against a direct double loop), and the worlds' surfaces line up
with their triangles."""
import numpy as np

from continuum_amd import surfaces as S
from continuum_amd.scenes import boulder, boulder_field, boulder_field_surface, boulder_surface, box_mesh

F32 = np.float32


def test_box_faces_are_parallelogram_charts():
    s = S.from_triangles(box_mesh([0, 0, 0], [1.0, 2.0, 3.0]))
    assert len(s.chart_size) == 6 and np.all(s.chart_grid == 1)
    assert np.array_equal(s.tri_chart, np.repeat(np.arange(6), 2))
    sizes = sorted(tuple(sorted(c)) for c in s.chart_size.tolist())
    assert sizes == sorted([(1.0, 2.0), (1.0, 2.0), (1.0, 3.0), (1.0, 3.0), (2.0, 3.0), (2.0, 3.0)])
    # corners map to chart corners: a + u (b - a) + v (d - a)
    for k in range(0, 12, 2):
        t, u = s.tris[k], s.tris[k + 1]
        a, b, d = t[0], t[1], u[2]
        for tri, uv in ((t, s.tri_uv[k]), (u, s.tri_uv[k + 1])):
            for c in range(3):
                p = a + uv[2 * c] * (b - a) + uv[2 * c + 1] * (d - a)
                assert np.allclose(p, tri[c], atol=1e-6)
    # flat normals, outward
    n = s.normals.reshape(-1, 3, 3)
    assert np.allclose(np.linalg.norm(n, axis=2), 1.0, atol=1e-6)


def _smooth_direct(tris):
    e1, e2 = tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]
    face = np.cross(e1, e2).astype(F32)
    keys = [tuple(np.round((p * F32(1e5)).astype(F32)).astype(np.int64)) for p in tris.reshape(-1, 3)]
    out = np.zeros((len(keys), 3), F32)
    for c, k in enumerate(keys):
        acc = np.zeros(3, F32)
        n = face[c // 3]
        for u, k2 in enumerate(keys):
            if k2 == k:
                f = face[u // 3]
                if F32(F32(f[0] * n[0]) + F32(f[1] * n[1])) + F32(f[2] * n[2]) > 0:
                    acc = (acc + f).astype(F32)
        out[c] = S._normalize_or_zero(acc)
    return out.reshape(-1, 9)


def test_smooth_normals_equal_the_direct_sum():
    tris = boulder((0.2, 0.1, 0.05), 12, 6)
    got = S.smooth_normals(tris)
    want = _smooth_direct(tris)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    # each corner's normal lies within 90 degrees of its own face
    face = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    assert np.all(np.einsum("tcj,tj->tc", got.reshape(-1, 3, 3), face) > 0)


def test_parametric_chart_covers_its_grid():
    s = boulder_surface((0.2, 0.1, 0.05), 24, 12)
    assert np.array_equal(s.tris, boulder((0.2, 0.1, 0.05), 24, 12))
    assert list(s.chart_grid[0]) == [24, 12] and np.all(s.tri_chart == 0)
    uv = s.tri_uv.reshape(-1, 3, 2)
    assert uv.min() == 0.0 and uv.max() == 1.0
    # chart coordinates are grid points: (i / 24, j / 12)
    assert np.allclose(uv[..., 0] * 24, np.round(uv[..., 0] * 24), atol=1e-5)
    assert np.allclose(uv[..., 1] * 12, np.round(uv[..., 1] * 12), atol=1e-5)
    # the chart's size is the longest row and column of the grid
    assert s.chart_size[0, 0] > 2.5 and s.chart_size[0, 1] > 1.0


def test_boulder_field_surface_lines_up():
    tris9, counts, pl, lights, eye, _ = boulder_field(count=50)
    s = boulder_field_surface()
    assert np.array_equal(s.tris.reshape(-1, 9), tris9.reshape(-1, 9))
    off = np.concatenate([[0], np.cumsum(counts)]).astype(int)
    for g in range(len(counts)):
        assert np.all(s.tri_chart[off[g]:off[g + 1]] == g)
