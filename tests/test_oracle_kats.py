"""The oracle against the reference's own known-answer tests and its GPU-vs-CPU
checks, with the constants those tests fix:

* crates/continuum/src/mask.rs:82-113 (ConeMask range / interval / OR),
* crates/continuum/src/coverage.rs:130-222 (CoverageMask),
* crates/continuum/src/cone3.rs:498-624 (the CPU mask cone against
  stratified ray sampling, blocking, leaks, geometry beyond the light),
* crates/continuum-gltf/tests/gpu.rs:294-336 (`cones_match_the_cpu`: the GPU
  semantics within 1e-4 of the light's weight of the CPU model) and
  :362-392 (`a_large_wall_shadows_receivers_close_to_it`, <= 5e-3).
"""
import math

import numpy as np
import pytest

from continuum_amd.scenes import box_mesh, meshes_to_arrays, rect_light
from oracle import oracle as O

FULL = (1 << 64) - 1


def test_mask_range_edges():
    L = O.lib()
    assert L.oracle_mask_range(0, 0) == 1
    assert L.oracle_mask_range(63, 63) == 1 << 63
    assert L.oracle_mask_range(0, 63) == FULL
    assert L.oracle_mask_range(5, 4) == 0
    assert L.oracle_mask_range(64, 70) == 0
    assert bin(L.oracle_mask_range(60, 99)).count("1") == 4


def test_mask_interval_uses_bin_centers():
    L = O.lib()
    assert L.oracle_mask_from_interval(0.0, 0.4 / 64.0) == 0
    assert L.oracle_mask_from_interval(0.0, 0.6 / 64.0) == 1
    assert L.oracle_mask_from_interval(0.0, 1.0) == FULL
    assert bin(L.oracle_mask_from_interval(0.5, 0.0)).count("1") == 32
    assert L.oracle_mask_from_interval(-1.0, 2.0) == FULL


def test_mask_correlated_occluders_do_not_double_count():
    L = O.lib()
    top = L.oracle_mask_from_interval(0.0, 0.5)
    bottom = L.oracle_mask_from_interval(0.5, 1.0)
    assert 1.0 - bin(top | top).count("1") / 64 == 0.5
    assert 1.0 - bin(top | bottom).count("1") / 64 == 0.0


def _visible(intervals):
    a = np.array(intervals, dtype=np.float32).reshape(-1)
    return O.lib().oracle_coverage_visible(a.ctypes.data, len(intervals))


def test_coverage_single_occluder_is_exact():
    for a, b in [(0.1, 0.37), (0.0, 1.0), (0.505, 0.507), (0.25, 0.5), (0.99, 1.0)]:
        assert abs(_visible([(a, b)]) - (1.0 - (b - a))) < 1e-6


def test_coverage_correlated_and_order():
    assert abs(_visible([(0.0, 0.5), (0.0, 0.5)]) - 0.5) < 1e-6
    assert abs(_visible([(0.0, 0.5), (0.5, 1.0)])) < 1e-6
    assert _visible([(0.1, 0.3013), (0.3, 0.6), (0.2, 0.25)]) == _visible([(0.2, 0.25), (0.3, 0.6), (0.1, 0.3013)])


def test_coverage_moving_one_occluder_is_one_lipschitz():
    seed = [12345]

    def nxt():
        seed[0] = (seed[0] * 1664525 + 1013904223) & 0xFFFFFFFF
        return np.float32(seed[0] >> 8) / np.float32(16777216.0)

    for _ in range(2000):
        others = []
        for _ in range(4):
            a = nxt()
            others.append((a, min(a + nxt() * np.float32(0.2), 1.0)))
        a = nxt()
        a0, b0 = a, min(a + nxt() * np.float32(0.3), 1.0)
        da, db = (nxt() - 0.5) * np.float32(0.02), (nxt() - 0.5) * np.float32(0.02)
        a1, b1 = min(max(a0 + da, 0.0), 1.0), min(max(b0 + db, 0.0), 1.0)
        change = abs(_visible(others + [(a0, b0)]) - _visible(others + [(min(a1, b1), max(a1, b1))]))
        assert change <= abs(a1 - a0) + abs(b1 - b0) + 1e-5


def test_coverage_open_length():
    L = O.lib()

    def ol(iv, t0, t1):
        a = np.array(iv, dtype=np.float32).reshape(-1)
        return L.oracle_coverage_open_length(a.ctypes.data if a.size else 0, len(iv), t0, t1)

    assert abs(ol([], 0.1, 0.3) - 0.2) < 1e-6
    assert abs(ol([(0.0, 0.5)], 0.25, 0.75) - 0.25) < 1e-6
    assert abs(ol([(0.0, 0.5)], 0.1, 0.4)) < 1e-6
    assert abs(ol([(0.0, 0.5), (0.5, 0.505)], 0.5, 0.52) - 0.02) < 1e-6


# ---- cone3.rs tests: a 0.2 x 0.2 light at height 1 over (0.5, 0.5)
FLOOR_LIGHT = rect_light([0.4, 0.4, 1.0], [0.2, 0, 0], [0, 0.2, 0])


def _r6(points):
    r = np.zeros((len(points), 6), np.float32)
    r[:, :3] = points
    r[:, 5] = 1.0
    return r


def test_cone3_bin_weights_sum_to_projected_solid_angle():
    pts = [[0.5, 0.5, 0.0], [0.1, 0.8, 0.0], [0.9, 0.3, 0.2]]
    _, total = O.cone3_visible(_r6(pts), FLOOR_LIGHT, [])
    for p, t in zip(pts, total):
        ref = O.sampled_visible(np.array(p, np.float32), FLOOR_LIGHT, [box_mesh([9, 9, 9], [9.1, 9.1, 9.1])], 256)
        assert abs(t - ref) / ref < 1e-3


def test_cone3_box_blocking_everything():
    vis, _ = O.cone3_visible(_r6([[0.5, 0.5, 0.0]]), FLOOR_LIGHT, [box_mesh([0.2, 0.2, 0.4], [0.8, 0.8, 0.5])],
                             cull=False)
    assert vis[0] < 1e-6


def test_cone3_matches_sampled_visibility():
    meshes = [box_mesh([0.35, 0.3, 0.4], [0.5, 0.7, 0.45]), box_mesh([0.55, 0.45, 0.6], [0.62, 0.52, 0.7])]
    pts = [[0.5, 0.5, 0.0], [0.42, 0.55, 0.0], [0.7, 0.4, 0.0]]
    vis, total = O.cone3_visible(_r6(pts), FLOOR_LIGHT, meshes, cull=False)
    for p, v, t in zip(pts, vis, total):
        ref = O.sampled_visible(np.array(p, np.float32), FLOOR_LIGHT, meshes, 128)
        assert abs(v - ref) < 0.02 * t


def test_cone3_internal_edges_do_not_leak_and_beyond_light():
    p = _r6([[0.5, 0.5, 0.0]])
    vis, total = O.cone3_visible(p, FLOOR_LIGHT, [box_mesh([0.0, 0.0, 0.5], [0.5, 1.0, 0.55])], cull=False)
    # The left half of the window's bins: the light's left half seen from the
    # centre, by the same Lambert formula (the oracle's weights, half the window).
    w = np.zeros(64, np.float32)
    t = np.zeros(1, np.float32)
    apex = np.array([0.5, 0.5, 0.0], np.float32)
    n = np.array([0, 0, 1], np.float32)
    O.lib().oracle_bin_weights(apex.ctypes.data, n.ctypes.data, FLOOR_LIGHT.ctypes.data, 8.0, w.ctypes.data,
                               t.ctypes.data)
    left = sum(w[b] for b in range(64) if b % 8 < 4)
    assert abs(vis[0] - (total[0] - left)) < 1e-4 * total[0]
    vis, total = O.cone3_visible(p, FLOOR_LIGHT, [box_mesh([0.2, 0.2, 1.1], [0.8, 0.8, 1.2])], cull=False)
    assert abs(vis[0] - total[0]) < 1e-6


def _floor_receivers(size=48):
    k = np.arange(size * size)
    r6 = np.zeros((size * size, 6), np.float32)
    r6[:, 0] = ((k % size) + 0.5) / size
    r6[:, 1] = ((k // size) + 0.5) / size
    r6[:, 5] = 1.0
    recv = np.zeros((len(r6), 8), np.float32)
    recv[:, :3] = r6[:, :3]
    recv[:, 4:7] = r6[:, 3:]
    return r6, recv


CPU_SCENE = [box_mesh([0.30, 0.10, 0.55], [0.70, 0.35, 0.58]), box_mesh([0.62, 0.40, 0.30], [0.85, 0.60, 0.33]),
             box_mesh([0.15, 0.45, 0.20], [0.25, 0.55, 0.30]), box_mesh([0.40, 0.65, 0.45], [0.48, 0.90, 0.62])]


@pytest.mark.parametrize("hierarchy", [0, 1, 2])
def test_wgsl_semantics_match_the_cpu_model(hierarchy):
    """gpu.rs:294-336: the GPU kernel's semantics within 1e-4 of the light's
    weight of the CPU model at every receiver."""
    r6, recv = _floor_receivers()
    tris9, counts = meshes_to_arrays(CPU_SCENE)
    flat = O.OracleFlat(tris9, counts)
    vis, full, touched, _, capped, ovf = O.irradiance(flat, recv, FLOOR_LIGHT[None, :], O.opts_array(hierarchy=hierarchy))
    cpu, total = O.cone3_visible(r6, FLOOR_LIGHT, CPU_SCENE)
    worst = float(np.max(np.abs(cpu - vis) / total))
    assert worst <= 1e-4, worst
    assert ovf == 0 and capped.sum() == 0
    # full bins are touched bins
    assert np.all((full & ~touched) == 0)


def test_a_large_wall_shadows_receivers_close_to_it():
    """gpu.rs:362-392, with the sun of `sun_behind_the_wall` (Light::sun,
    lib.rs:667-700, computed here in f64 as the reference does)."""
    toward = np.array([-0.6, -0.3, 0.75])
    toward /= np.linalg.norm(toward)
    distance = 900.0
    size = distance * (2.0 * math.tan(math.radians(0.25)))
    w = -toward
    side = np.cross([0, 0, 1.0], w)
    a = side / np.linalg.norm(side) * size
    b = np.cross(w, a)
    b = b / np.linalg.norm(b) * size
    middle = toward * distance
    sun = rect_light(middle - (a + b) * 0.5, a, b, one_sided=True)
    c = [np.array([0.0, y, z]) for y, z in [(-10, -10), (10, -10), (10, 10), (-10, 10)]]
    wall = np.array([[c[0], c[1], c[2]], [c[0], c[2], c[3]]], np.float32)
    recv = np.zeros((400, 8), np.float32)
    for k in range(400):
        i, j = k % 20, k // 20
        n = np.array([-0.05 * i, 0.0, 1.0], np.float32)
        n = n * (np.float32(1.0) / np.float32(np.sqrt(np.float32(n @ n))))
        recv[k, :3] = [0.02 + 0.03 * i, -1.0 + 0.1 * j, 0.5]
        recv[k, 4:7] = n
    opts = O.opts_array(hierarchy=0)
    empty = O.OracleFlat(np.zeros((0, 9), np.float32), np.zeros(0, np.uint32))
    open_, *_ = O.irradiance(empty, recv, sun[None, :], opts)
    t9, cn = meshes_to_arrays([wall])
    shadowed, *_ = O.irradiance(O.OracleFlat(t9, cn), recv, sun[None, :], opts)
    assert np.all(open_ > 0)
    assert float(np.max(shadowed / np.maximum(open_, 1e-12))) <= 5e-3
