"""Surfaces for the view cache: each triangle's chart, its corners' chart
coordinates and shading normals, and each chart's size and cell grid, as
continuum::mesh::Surface holds them (crates/continuum/src/mesh.rs:20-39).

`from_triangles` restates Surface::from_triangles (mesh.rs:63-102: a
parallelogram pair shares one chart, any other triangle has its own),
`parametric` Surface::parametric (mesh.rs:224-284: one chart over an
nu x nv grid, degenerate triangles dropped), `smooth_normals` mesh.rs:330-
376. Arithmetic is f32 in glam's order. `concat` numbers charts globally,
geometries in order, as the view cache's chart tables do (vcache.rs:704).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

F32 = np.float32


@dataclass
class Surface:
    tris: np.ndarray        # (n, 3, 3) f32
    tri_chart: np.ndarray   # (n,) u32, chart within this surface
    tri_uv: np.ndarray      # (n, 6) f32
    normals: np.ndarray     # (n, 9) f32
    chart_size: np.ndarray  # (c, 2) f32
    chart_grid: np.ndarray  # (c, 2) u32


def _len(v) -> F32:
    v = np.asarray(v, dtype=F32)
    return F32(np.sqrt(F32(F32(v[0] * v[0]) + F32(v[1] * v[1])) + F32(v[2] * v[2])))


def _cross(a, b):
    return np.array([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]], dtype=F32)


def _normalize_or_zero(v):
    n = _len(v)
    r = F32(1.0) / n if n > 0 else F32(np.inf)
    if not np.isfinite(r) or r <= 0:
        return np.zeros(3, dtype=F32)
    return (np.asarray(v, dtype=F32) * r).astype(F32)


def from_triangles(tris) -> Surface:
    """Surface::from_triangles (mesh.rs:63-102) with flat normals."""
    tris = np.asarray(tris, dtype=F32).reshape(-1, 3, 3)
    charts, grids, tri_chart, uvs, normals = [], [], [], [], []
    i = 0
    while i < len(tris):
        t = tris[i]
        n = _normalize_or_zero(_cross(t[1] - t[0], t[2] - t[0]))
        pair = False
        if i + 1 < len(tris):
            u = tris[i + 1]
            if np.array_equal(u[0], t[0]) and np.array_equal(u[1], t[2]):
                a, b, c, d = t[0], t[1], t[2], u[2]
                scale = max(_len(c - a), F32(1e-12))
                pair = _len((a + c) - (b + d)) <= F32(1e-5) * scale
        chart = len(charts)
        if pair:
            a, b, d = t[0], t[1], tris[i + 1][2]
            charts.append([_len(b - a), _len(d - a)])
            grids.append([1, 1])
            for tri, uv in ((t, [0, 0, 1, 0, 1, 1]), (tris[i + 1], [0, 0, 1, 1, 0, 1])):
                tri_chart.append(chart)
                uvs.append(uv)
                normals.append(np.tile(_normalize_or_zero(_cross(tri[1] - tri[0], tri[2] - tri[0])), 3))
            i += 2
        else:
            charts.append([_len(t[1] - t[0]), _len(t[2] - t[0])])
            grids.append([1, 1])
            tri_chart.append(chart)
            uvs.append([0, 0, 1, 0, 0, 1])
            normals.append(np.tile(n, 3))
            i += 1
    return Surface(tris, np.array(tri_chart, np.uint32), np.array(uvs, F32).reshape(-1, 6),
                   np.array(normals, F32).reshape(-1, 9), np.array(charts, F32).reshape(-1, 2),
                   np.array(grids, np.uint32).reshape(-1, 2))


def smooth_normals(tris) -> np.ndarray:
    """mesh.rs:330-376: each corner's normal sums the face normals facing its
    own side of every triangle with a corner at the same position (to 1e-5),
    in triangle order, normalized. (n, 9) f32."""
    tris = np.asarray(tris, dtype=F32).reshape(-1, 3, 3)
    e1, e2 = tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]
    face = np.stack([e1[:, 1] * e2[:, 2] - e1[:, 2] * e2[:, 1], e1[:, 2] * e2[:, 0] - e1[:, 0] * e2[:, 2],
                     e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]], axis=1).astype(F32)
    keys = np.round((tris.reshape(-1, 3) * F32(1e5)).astype(F32)).astype(np.int64)
    order = np.lexsort((np.arange(len(keys)), keys[:, 2], keys[:, 1], keys[:, 0]))
    sk = keys[order]
    starts = np.concatenate([[0], np.nonzero(np.any(sk[1:] != sk[:-1], axis=1))[0] + 1, [len(sk)]])
    sizes = np.diff(starts)
    group_of_sorted = np.repeat(np.arange(len(sizes)), sizes)
    rank = np.arange(len(sk)) - starts[group_of_sorted]
    members = np.full((len(sizes), int(sizes.max()) if len(sizes) else 1), -1, np.int64)
    members[group_of_sorted, rank] = order
    group = np.empty(len(keys), np.int64)
    group[order] = group_of_sorted
    n = face[np.arange(len(keys)) // 3]
    acc = np.zeros((len(keys), 3), F32)
    # members in corner order, each summed in turn (f32, as the reference's iterator sum)
    for j in range(members.shape[1]):
        m = members[group, j]
        f = face[np.maximum(m, 0) // 3]
        d = ((f[:, 0] * n[:, 0]).astype(F32) + (f[:, 1] * n[:, 1]).astype(F32)).astype(F32) + (f[:, 2] * n[:, 2])
        take = (m >= 0) & (d.astype(F32) > 0)
        acc = np.where(take[:, None], (acc + f).astype(F32), acc)
    ln = np.sqrt(((acc[:, 0] * acc[:, 0]).astype(F32) + (acc[:, 1] * acc[:, 1]).astype(F32)).astype(F32)
                 + (acc[:, 2] * acc[:, 2]).astype(F32)).astype(F32)
    with np.errstate(divide="ignore"):
        r = (F32(1.0) / ln).astype(F32)
    ok = np.isfinite(r) & (r > 0)
    out = np.where(ok[:, None], (acc * r[:, None]).astype(F32), F32(0.0))
    return out.reshape(-1, 9).astype(F32)


def parametric(grid: np.ndarray, keep: np.ndarray | None = None, flip: bool = False) -> Surface:
    """Surface::parametric (mesh.rs:224-284) over grid points (nu+1, nv+1, 3):
    triangles [a, b, c], [a, c, d] of each quad (i outer, j inner), chart
    coordinates (i/nu, j/nv), degenerate ones dropped; `flip` reverses each
    triangle's winding (corners 0, 2, 1) as a caller turning the surface
    outward does. Smooth normals."""
    grid = np.asarray(grid, dtype=F32)
    nu, nv = grid.shape[0] - 1, grid.shape[1] - 1
    seg_u = np.sqrt(np.sum((grid[1:] - grid[:-1]) ** 2, axis=-1, dtype=F32), dtype=F32)
    seg_v = np.sqrt(np.sum((grid[:, 1:] - grid[:, :-1]) ** 2, axis=-1, dtype=F32), dtype=F32)
    along_u = max(F32(sum(seg_u[i, j] for i in range(nu))) for j in range(nv + 1))
    along_v = max(F32(sum(seg_v[i, j] for j in range(nv))) for i in range(nu + 1))
    a, b, c, d = grid[:-1, :-1], grid[1:, :-1], grid[1:, 1:], grid[:-1, 1:]
    tris = np.stack([np.stack([a, b, c], axis=2), np.stack([a, c, d], axis=2)], axis=2).reshape(-1, 3, 3)
    ii, jj = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    u0, u1, v0, v1 = ii / nu, (ii + 1) / nu, jj / nv, (jj + 1) / nv
    uv1 = np.stack([u0, v0, u1, v0, u1, v1], axis=-1)
    uv2 = np.stack([u0, v0, u1, v1, u0, v1], axis=-1)
    uv = np.stack([uv1, uv2], axis=2).reshape(-1, 6).astype(F32)
    if keep is None:
        e1, e2 = tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]
        scale = np.maximum(np.linalg.norm(e1, axis=1), np.linalg.norm(e2, axis=1))
        keep = np.linalg.norm(np.cross(e1, e2), axis=1) > 1e-6 * scale * scale
    tris, uv = tris[keep], uv[keep]
    if flip:
        tris = tris[:, [0, 2, 1]]
        uv = uv[:, [0, 1, 4, 5, 2, 3]]
    tris = np.ascontiguousarray(tris, dtype=F32)
    return Surface(tris, np.zeros(len(tris), np.uint32), np.ascontiguousarray(uv, dtype=F32), smooth_normals(tris),
                   np.array([[along_u, along_v]], F32), np.array([[nu, nv]], np.uint32))


def concat(surfaces) -> Surface:
    """Geometries in order, charts numbered globally."""
    first = np.cumsum([0] + [len(s.chart_size) for s in surfaces])
    return Surface(np.concatenate([s.tris for s in surfaces]),
                   np.concatenate([s.tri_chart + np.uint32(f) for s, f in zip(surfaces, first)]).astype(np.uint32),
                   np.concatenate([s.tri_uv for s in surfaces]), np.concatenate([s.normals for s in surfaces]),
                   np.concatenate([s.chart_size for s in surfaces]), np.concatenate([s.chart_grid for s in surfaces]))
