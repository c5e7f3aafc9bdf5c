"""Host-side lightmap texels for the lighting cache: the reference's
`Lightmap::from_meshes` (crates/continuum-gpu/src/lightmap.rs:95-190) over
`Surface::from_triangles` charts (crates/continuum/src/mesh.rs:63-104).

Every chart of a static mesh gets a grid of texels about `texel_size`
apart, padded to whole 8 x 8 blocks of 64 consecutive receivers; a texel on
no triangle (within 1.5 texels) keeps the chart's first vertex and a zero
normal. Each block's bounds are those of its real texels. Arithmetic is
f32, in the reference's order where it decides a texel's placement.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence

import numpy as np

BLOCK = 8  # lightmap.rs:80
F32 = np.float32


def _len(v) -> np.float32:
    return F32(math.sqrt(float(F32(v[0] * v[0]) + F32(v[1] * v[1]) + F32(v[2] * v[2]))))


def _normalize_or_zero(v) -> np.ndarray:
    n = _len(v)
    if not (n > 0 and math.isfinite(float(n))):
        return np.zeros(3, F32)
    out = v * F32(F32(1.0) / n)
    return out if np.all(np.isfinite(out)) else np.zeros(3, F32)


def _flat_normal(t) -> np.ndarray:
    return _normalize_or_zero(np.cross(t[1] - t[0], t[2] - t[0]).astype(F32))


def surface_charts(tris: np.ndarray):
    """Surface::from_triangles: a chart per planar quad (two triangles
    sharing the diagonal) or per triangle; (charts sizes, chart of each
    triangle, uv (n, 3, 2), flat normals (n, 3))."""
    tris = np.asarray(tris, F32).reshape(-1, 3, 3)
    sizes, chart, uvs, normals = [], [], [], []
    i = 0
    while i < len(tris):
        t = tris[i]
        n = _flat_normal(t)
        c = len(sizes)
        quad = False
        if i + 1 < len(tris):
            u = tris[i + 1]
            if np.array_equal(u[0], t[0]) and np.array_equal(u[1], t[2]):
                a, b, cc, d = t[0], t[1], t[2], u[2]
                scale = max(float(_len(cc - a)), 1e-12)
                quad = float(_len((a + cc) - (b + d))) <= 1e-5 * scale
        if quad:
            u = tris[i + 1]
            a, b, d = t[0], t[1], u[2]
            sizes.append((_len(b - a), _len(d - a)))
            for tri, uv in ((t, [(0, 0), (1, 0), (1, 1)]), (u, [(0, 0), (1, 1), (0, 1)])):
                chart.append(c)
                uvs.append(uv)
                normals.append(_flat_normal(tri))
            i += 2
        else:
            sizes.append((_len(t[1] - t[0]), _len(t[2] - t[0])))
            chart.append(c)
            uvs.append([(0, 0), (1, 0), (0, 1)])
            normals.append(n)
            i += 1
    return np.array(sizes, F32).reshape(-1, 2), np.array(chart, np.int64), np.array(uvs, F32).reshape(-1, 3, 2), \
        np.array(normals, F32).reshape(-1, 3)


def _closest_barycentric(c, q):
    """lightmap.rs:452-474: distance from c to triangle q (0 inside) and the
    nearest point's barycentric coordinates."""
    e1, e2, d = q[1] - q[0], q[2] - q[0], c - q[0]
    area = F32(e1[0] * e2[1] - e1[1] * e2[0])
    s = F32(F32(d[0] * e2[1] - d[1] * e2[0]) / area)
    t = F32(F32(e1[0] * d[1] - e1[1] * d[0]) / area)
    if s >= -1e-6 and t >= -1e-6 and s + t <= 1.0 + 1e-6:
        return F32(0.0), np.array([F32(1.0) - s - t, s, t], F32)
    best = (math.inf, np.zeros(3, F32))
    for a, b in ((0, 1), (1, 2), (2, 0)):
        edge = q[b] - q[a]
        u = F32(min(max(float(F32(np.dot(c - q[a], edge)) / F32(np.dot(edge, edge))), 0.0), 1.0))
        p = q[a] + edge * u
        dist = float(np.hypot(*(p - c)))
        if dist < best[0]:
            w = np.zeros(3, F32)
            w[a] = F32(1.0) - u
            w[b] = u
            best = (dist, w)
    return F32(best[0]), best[1]


def _place(tris, normals, uvs, nu: int, nv: int, reach: float):
    """lightmap.rs:404-450: grid cells placed on the chart's triangles, as
    (point, normal) or None, row-major."""
    scale = np.array([nu, nv], F32)
    best = [(math.inf, None)] * (nu * nv)
    for k in range(len(tris)):
        tri, q = tris[k], uvs[k] * scale
        if abs(float((q[1] - q[0])[0] * (q[2] - q[0])[1] - (q[1] - q[0])[1] * (q[2] - q[0])[0])) < 1e-12:
            continue
        lo = q.min(axis=0) - F32(reach) - F32(0.5)
        hi = q.max(axis=0) + F32(reach) - F32(0.5)
        i0, j0 = max(math.ceil(lo[0]), 0), max(math.ceil(lo[1]), 0)
        i1, j1 = min(math.floor(hi[0]), nu - 1), min(math.floor(hi[1]), nv - 1)
        if i1 < 0 or j1 < 0:
            continue
        geometric = _flat_normal(tri)
        for j in range(j0, j1 + 1):
            for i in range(i0, i1 + 1):
                d, b = _closest_barycentric(np.array([i + 0.5, j + 0.5], F32), q)
                cell = j * nu + i
                if d <= reach and d < best[cell][0]:
                    point = tri[0] * b[0] + tri[1] * b[1] + tri[2] * b[2]
                    n = _normalize_or_zero(normals[k] * b[0] + normals[k] * b[1] + normals[k] * b[2])
                    if not n.any():
                        n = geometric
                    best[cell] = (float(d), (point.astype(F32), n.astype(F32)))
    return [c for _, c in best]


@dataclass
class Lightmap:
    """Texels (receivers, 8 floats each, 64 per block) and block bounds."""

    texels: np.ndarray  # (n_blocks * 64, 8)
    blocks: np.ndarray  # (n_blocks, 6): lo, hi
    charts: List[tuple]  # (mesh, first texel, nu, nv, blocks per row)
    irradiance: Optional[np.ndarray] = None  # per texel, once baked

    @property
    def n_blocks(self) -> int:
        return len(self.blocks)

    def real(self) -> np.ndarray:
        """Texels that lie on a triangle (a nonzero normal)."""
        return np.any(self.texels[:, 4:7] != 0, axis=1)

    @classmethod
    def from_meshes(cls, meshes: Sequence, texel_size: float, is_static: Callable[[int], bool]) -> "Lightmap":
        texels: List[np.ndarray] = []
        blocks: List[np.ndarray] = []
        charts: List[tuple] = []
        for m, mesh in enumerate(meshes):
            if not is_static(m):
                continue
            tris = np.asarray(mesh, F32).reshape(-1, 3, 3)
            if len(tris) == 0:
                continue
            sizes, chart, uvs, normals = surface_charts(tris)
            for c in range(len(sizes)):
                sel = np.nonzero(chart == c)[0]
                nu = min(max(math.ceil(float(sizes[c][0] / F32(texel_size))), 1), 1024)
                nv = min(max(math.ceil(float(sizes[c][1] / F32(texel_size))), 1), 1024)
                nbu, nbv = -(-nu // BLOCK), -(-nv // BLOCK)
                cells = _place(tris[sel], normals[sel], uvs[sel], nu, nv, 1.5)
                charts.append((m, 64 * len(blocks), nu, nv, nbu))
                pad = tris[sel[0]][0]
                for bj in range(nbv):
                    for bi in range(nbu):
                        lo = np.full(3, np.finfo(F32).max, F32)
                        hi = np.full(3, -np.finfo(F32).max, F32)
                        block = np.zeros((64, 8), F32)
                        for k in range(64):
                            i, j = bi * BLOCK + k % BLOCK, bj * BLOCK + k // BLOCK
                            cell = cells[j * nu + i] if i < nu and j < nv else None
                            if cell is not None:
                                p = (cell[0] + cell[1] * F32(1e-4)).astype(F32)
                                lo, hi = np.minimum(lo, p), np.maximum(hi, p)
                                block[k, 0:3], block[k, 4:7] = p, cell[1]
                            else:
                                block[k, 0:3] = pad
                        texels.append(block)
                        blocks.append(np.concatenate([lo, hi]).astype(F32))
        tex = np.concatenate(texels) if texels else np.zeros((0, 8), F32)
        return cls(tex.astype(F32), np.array(blocks, F32).reshape(-1, 6), charts)
