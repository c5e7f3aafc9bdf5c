"""GPU: instanced worlds (per-object transforms, the object hierarchy over
world bounds, prefiltered object-level splats) against the oracle, bit for
bit: camera hits, masks and irradiance, for each hierarchy and with object
splats on and off."""
import numpy as np
import pytest

from continuum_amd import _lib
from continuum_amd.kernel import PLACEMENT_DTYPE, Flat, Scene
from continuum_amd.scenes import boulder, box_mesh, camera_uniform, meshes_to_arrays, quad_pair, rect_light, sphere_light
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _world(n=300, seed=7):
    rng = np.random.default_rng(seed)
    geoms = [quad_pair([-30, -30, 0], [60, 0, 0], [0, 60, 0]), box_mesh([-0.5, -0.3, 0.0], [0.5, 0.3, 0.8]),
             boulder((0.2, 0.1, 0.05), 40, 20), boulder((0.1, 0.15, 0.08), 30, 15)]
    tris9, counts = meshes_to_arrays(geoms)
    pl = np.zeros(n + 1, dtype=PLACEMENT_DTYPE)
    pl[0]["rotation"] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    pl[0]["scale"] = 1.0
    for k in range(1, n + 1):
        if k % 5:
            ang = rng.uniform(0, 2 * np.pi)
            c, s = np.cos(ang), np.sin(ang)
            rot = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])
        else:
            a = rng.normal(size=3)
            a /= np.linalg.norm(a)
            ang = rng.uniform(0, 2 * np.pi)
            K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])
            rot = np.eye(3) + np.sin(ang) * K + (1 - np.cos(ang)) * K @ K
        sc = rng.uniform(0.6, 1.6)
        pl[k]["geometry"] = 1 + k % 3
        pl[k]["rotation"] = rot.astype(np.float32).reshape(9)
        pl[k]["scale"] = sc
        # half rest on the ground, half float between it and the big light
        z = 0.2 * sc if k % 2 else rng.uniform(2.0, 25.0)
        pl[k]["translation"] = [rng.uniform(-25, 25), rng.uniform(-25, 25), z]
    lights = np.stack([rect_light([-20.0, 20.0, 30.0], [40.0, 0, 0], [0, -40.0, 0], one_sided=True),
                       sphere_light([300.0, -200.0, 400.0], 4.0),
                       rect_light([8.0, -6.0, 3.0], [1.0, 0, 0], [0, 0, 1.0], one_sided=False)])
    return tris9, counts, pl, lights


@pytest.mark.parametrize("prefilter", [0.0, 1.0, 3.0])
def test_instanced_views_match_the_oracle(device, prefilter):
    tris9, counts, pl, lights = _world()
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True)
    o = O.OracleInstancedFlat(tris9, counts, pl, object_prefilter=True)
    scene = Scene(device, f)
    cam = camera_uniform([-24.0, -24.0, 2.5], [5.0, 5.0, 0.5], 55.0, 96, 64)
    for hierarchy in (1, 2):
        for caps in ((4096, 512), (O.NONE, O.NONE)):
            opts = _lib.default_options(hierarchy=hierarchy, prefilter=prefilter, object_cov=f.object_cov,
                                        light_cone_cap=caps[0], wide_cone_cap=caps[1])
            scene.counters(reset=True)
            hits, vis, full, touched = scene.render_view(cam, 96, 64, lights, opts)
            gpu_capped = int(scene.counters(reset=True)[0][0])
            rh, recv, _ = O.primary(o, cam, 96, 64, lights, hierarchy=hierarchy)
            assert np.array_equal(hits, rh), hierarchy
            rv, rf, rt, _, capped, _ = O.irradiance(
                o, recv, lights, O.opts_array(hierarchy=hierarchy, prefilter=prefilter, object_cov=f.object_cov,
                                              light_cone_cap=caps[0], wide_cone_cap=caps[1]))
            bad = (vis.view(np.uint32) != rv.view(np.uint32)) | (full != rf) | (touched != rt)
            if caps[0] == O.NONE:
                assert capped.sum() == 0 and gpu_capped == 0 and not bad.any(), (hierarchy, int(bad.sum()))
            else:
                # A walk that reaches its node cap keeps what it found; the
                # packet's shared nearest-first order can move a lane's walk
                # across the cap where its lone walk is not (the only packet
                # difference). Every other query is bit-exact.
                assert int((bad & (capped == 0)).sum()) <= gpu_capped, hierarchy
                assert (bad & (capped == 0)).mean() < 1e-3, hierarchy
            assert (rv > 0).sum() > 1000


def test_object_splats_change_results_only_where_prefiltered(device):
    """With object splats off (object_cov NONE) the walk descends every
    object; with them on, far objects are splatted: the two differ, and each
    equals the oracle with the same setting (checked above)."""
    tris9, counts, pl, lights = _world()
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True)
    scene = Scene(device, f)
    cam = camera_uniform([-24.0, -24.0, 2.5], [5.0, 5.0, 0.5], 55.0, 96, 64)
    on = scene.render_view(cam, 96, 64, lights, _lib.default_options(prefilter=3.0, object_cov=f.object_cov))
    off = scene.render_view(cam, 96, 64, lights, _lib.default_options(prefilter=3.0))
    assert np.array_equal(on[0], off[0])          # camera hits do not splat
    assert not np.array_equal(on[1], off[1])


def test_moves_equal_a_fresh_build_and_the_oracle(device):
    """Movers (renderer.rs:442 set_transform, instanced.rs:1226 apply_moves):
    after moving some of them, every query equals a scene built afresh with
    the moved placements, and the oracle's, bit for bit (caps off)."""
    tris9, counts, pl, lights = _world()
    movable = list(range(1, 120, 3))
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True, movable=movable)
    scene = Scene(device, f)
    rng = np.random.default_rng(11)
    moved = pl.copy()
    for m in movable[::2]:
        ang = rng.uniform(0, 2 * np.pi)
        c, s = np.cos(ang), np.sin(ang)
        rot = np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]], np.float32)
        t = (moved[m]["translation"] + rng.uniform(-4, 4, 3) * np.array([1, 1, 0.3])).astype(np.float32)
        scene.set_transform(m, rot, t)
        moved[m]["rotation"] = rot.reshape(9)
        moved[m]["translation"] = t
    scene.apply_moves()
    fresh = Flat.build_instanced(tris9, counts, moved, object_prefilter=True, movable=movable)
    fresh_scene = Scene(device, fresh)
    o = O.OracleInstancedFlat(tris9, counts, moved, object_prefilter=True, movable=movable)
    cam = camera_uniform([-24.0, -24.0, 2.5], [5.0, 5.0, 0.5], 55.0, 96, 64)
    for prefilter in (0.0, 3.0):
        kw = dict(hierarchy=2, prefilter=prefilter, object_cov=f.object_cov, light_cone_cap=O.NONE,
                  wide_cone_cap=O.NONE)
        got = scene.render_view(cam, 96, 64, lights, _lib.default_options(**kw))
        ref = fresh_scene.render_view(cam, 96, 64, lights, _lib.default_options(**kw))
        for a, b in zip(got, ref):
            assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32))
        rh, recv, _ = O.primary(o, cam, 96, 64, lights, hierarchy=2)
        assert np.array_equal(got[0], rh)
        rv, rf, rt, _, capped, _ = O.irradiance(o, recv, lights, O.opts_array(
            hierarchy=2, prefilter=prefilter, object_cov=f.object_cov, light_cone_cap=O.NONE, wide_cone_cap=O.NONE))
        assert np.array_equal(got[1].view(np.uint32), rv.view(np.uint32)) and np.array_equal(got[2], rf)
    with pytest.raises(Exception):
        scene.set_transform(0, np.eye(3), [0, 0, 0])    # the ground is static


def _cloth(nu=24, nv=16, phase=0.0):
    """A 2 m x 1.5 m sheet on an nu x nv grid, rippled by `phase`."""
    u = np.linspace(-1.0, 1.0, nu + 1)
    v = np.linspace(-0.75, 0.75, nv + 1)
    x, y = np.meshgrid(u, v, indexing="ij")
    z = 0.6 + 0.15 * np.sin(3 * x + phase) * np.cos(2 * y - phase)
    g = np.stack([x, y, z], axis=-1).astype(np.float32)
    a, b, c, d = g[:-1, :-1], g[1:, :-1], g[1:, 1:], g[:-1, 1:]
    return np.stack([np.stack([a, b, c], axis=2), np.stack([a, c, d], axis=2)], axis=2).reshape(-1, 3, 3)


def test_poses_match_the_oracle(device):
    """Deforming geometry (instanced.rs:1154 apply_poses): a sheet placed by
    two movable objects takes a new pose; its hierarchy is refit on the GPU
    (bit for bit the oracle's `refit_triangle_hierarchy`), the movers'
    subtree follows its new bounds, and every query equals the oracle's on
    the refit scene. A geometry with a static instance refuses poses."""
    geoms = [quad_pair([-12, -12, 0], [24, 0, 0], [0, 24, 0]), box_mesh([-0.5, -0.3, 0.0], [0.5, 0.3, 0.8]),
             _cloth()]
    tris9, counts = meshes_to_arrays(geoms)
    rng = np.random.default_rng(4)
    n = 14
    pl = np.zeros(n, dtype=PLACEMENT_DTYPE)
    pl["rotation"][:] = [1, 0, 0, 0, 1, 0, 0, 0, 1]
    pl["scale"][:] = 1.0
    for k in range(1, n):
        pl[k]["geometry"] = 2 if k in (3, 9) else 1
        pl[k]["translation"] = [rng.uniform(-8, 8), rng.uniform(-8, 8), 0.0]
    movable = [3, 9, 5]
    lights = np.stack([rect_light([-2.0, -2.0, 5.0], [4.0, 0, 0], [0, 4.0, 0]), sphere_light([30.0, -20.0, 40.0], 2.0)])
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True, movable=movable)
    scene = Scene(device, f)
    posed = _cloth(phase=1.3).reshape(-1, 9)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(int)
    print("pose: sheet of", len(posed), "triangles on objects 3 and 9", flush=True)
    scene.set_pose(2, posed)
    scene.apply_moves()
    with pytest.raises(Exception):
        scene.set_pose(1, tris9[off[1]:off[2]])  # boxes 1, 2, ... are static
    # the oracle's scene: the new triangles, the original hierarchy refit, and
    # the object hierarchy over the new bounds
    new9 = tris9.copy()
    new9[off[2]:off[3]] = posed
    fresh = Flat.build_instanced(new9, counts, pl, object_prefilter=True, movable=movable)
    o = O.OracleInstancedFlat(tris9, counts, pl, object_prefilter=True, movable=movable)
    nb = len(o.blas)
    blas_r, cov_r = O.refit(o, fresh.tris)
    got_blas, got_cov = scene.read_blas()
    cols = [0, 1, 2, 4, 5, 6, 8, 9]
    assert np.array_equal(got_blas[:, cols], blas_r[:, cols])
    assert np.array_equal(got_cov[:nb].view(np.uint32), cov_r[:nb].view(np.uint32))

    class Assembled:
        pass

    a = Assembled()
    a.tris, a.ranges, a.blas, a.blas_items = fresh.tris, fresh.ranges, blas_r, o.blas_items
    a.blas_cov = np.ascontiguousarray(np.concatenate([cov_r[:nb], fresh.blas_cov[nb:]]))
    a.nodes, a.items, a.xforms = fresh.nodes, fresh.items, fresh.xforms
    cam = camera_uniform([-10.0, -9.0, 3.0], [0.0, 0.0, 0.5], 60.0, 96, 64)
    for prefilter in (0.0, 2.0):
        kw = dict(hierarchy=2, prefilter=prefilter, object_cov=f.object_cov, light_cone_cap=O.NONE,
                  wide_cone_cap=O.NONE)
        hits, vis, full, touched = scene.render_view(cam, 96, 64, lights, _lib.default_options(**kw))
        rh, recv, _ = O.primary(a, cam, 96, 64, lights, hierarchy=2)
        assert np.array_equal(hits, rh)
        rv, rf, rt, _, capped, _ = O.irradiance(a, recv, lights, O.opts_array(**kw))
        assert np.array_equal(vis.view(np.uint32), rv.view(np.uint32)) and np.array_equal(full, rf)
        assert np.array_equal(touched, rt)
    # the sheet's pixels see the new pose: its hits are not the rest pose's
    rest = Scene(device, f).render_view(cam, 96, 64, lights, _lib.default_options(hierarchy=2))
    assert not np.array_equal(rest[0], hits)


def test_blended_and_transmissive_objects_match_the_oracle(device):
    """Objects placed alpha-blended (clear > 0), as thin glass (clear < 0:
    each triangle's weight from Schlick's reflectance at the cone's centre
    direction, WGSL :608-614) and as glass volumes (clear <= -2), with and
    without object splats, against the oracle bit for bit (caps off)."""
    tris9, counts, pl, lights = _world()
    clears = np.array([0.0, 0.35, -0.4, -1.0, -2.6, -3.0], np.float32)
    pl["clear"][1:] = clears[np.arange(1, len(pl)) % len(clears)]
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True)
    o = O.OracleInstancedFlat(tris9, counts, pl, object_prefilter=True)
    assert np.array_equal(f.blas_cov.view(np.uint32), o.blas_cov.view(np.uint32))
    scene = Scene(device, f)
    cam = camera_uniform([-24.0, -24.0, 2.5], [5.0, 5.0, 0.5], 55.0, 96, 64)
    for prefilter in (0.0, 2.0):
        opts = _lib.default_options(hierarchy=2, prefilter=prefilter, object_cov=f.object_cov,
                                    light_cone_cap=O.NONE, wide_cone_cap=O.NONE)
        hits, vis, full, touched = scene.render_view(cam, 96, 64, lights, opts)
        rh, recv, _ = O.primary(o, cam, 96, 64, lights)
        assert np.array_equal(hits, rh)
        rv, rf, rt, _, capped, _ = O.irradiance(
            o, recv, lights, O.opts_array(hierarchy=2, prefilter=prefilter, object_cov=f.object_cov,
                                          light_cone_cap=O.NONE, wide_cone_cap=O.NONE))
        assert capped.sum() == 0
        bad = (vis.view(np.uint32) != rv.view(np.uint32)) | (full != rf) | (touched != rt)
        assert not bad.any(), (prefilter, int(bad.sum()))
        # partial coverage (blended or glass) shows as touched bins that are not full
        assert ((rt & ~rf) != 0).sum() > 1000
