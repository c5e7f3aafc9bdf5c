"""GPU: the batched form. Many views of uneven size and light count in one
camera pass and one cone launch give, view by view, exactly the results of
each view's own launch and of the oracle, whatever the tuning (tile shape,
persistent grid) and on every replay of the captured hipGraph."""
import numpy as np
import pytest

from continuum_amd import _lib
from continuum_amd.kernel import Batch, Flat, Scene
from continuum_amd.scenes import camera_uniform, meshes_to_arrays, rect_light, room_scene, sphere_light
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _views():
    sc = room_scene()
    lights = [sc.lights[0], rect_light([0.5, 0.5, 2.5], [0.6, 0, 0], [0, 0.4, 0], one_sided=False),
              sphere_light([3.2, 1.0, 2.4], 0.15), rect_light([3.0, 3.0, 2.99], [0.3, 0, 0], [0, -0.3, 0], True)]
    rng = np.random.default_rng(4)
    views = []
    sizes = [(64, 64), (40, 23), (96, 54), (8, 8), (17, 90), (130, 7)]
    for k, (w, h) in enumerate(sizes):
        eye = rng.uniform([0.3, 0.3, 0.5], [3.7, 3.7, 2.5])
        tgt = rng.uniform([0.3, 0.3, 0.0], [3.7, 3.7, 2.0])
        lf = k % 3
        lc = 1 + (k % (len(lights) - lf))
        views.append((camera_uniform(eye, tgt, 50.0 + 5 * k, w, h), w, h, lf, lc))
    return sc, np.stack(lights), views


def test_batch_equals_single_views_and_the_oracle(device):
    sc, lights, views = _views()
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    scene = Scene(device, flat)
    opts = _lib.default_options(hierarchy=2)
    batch = Batch(scene, views, lights)
    assert batch.queries == sum(w * h * lc for _, w, h, _, lc in views)
    batch.run(opts, Batch.tuning(8, 0))
    for k, (cam, w, h, lf, lc) in enumerate(views):
        hits, vis, full, touched = batch.read(k)
        sh, sv, sf, st = scene.render_view(cam, w, h, lights[lf:lf + lc], opts)
        assert np.array_equal(hits, sh), k
        assert np.array_equal(vis.view(np.uint32), sv.view(np.uint32)), k
        assert np.array_equal(full, sf) and np.array_equal(touched, st), k
        rh, rrecv, _ = O.primary(flat, cam, w, h, lights[lf:lf + lc])
        rv, rf, rt, *_ = O.irradiance(flat, rrecv, lights[lf:lf + lc], O.opts_array(hierarchy=2))
        assert np.array_equal(hits, rh), k
        assert np.array_equal(vis.view(np.uint32), rv.view(np.uint32)) and np.array_equal(full, rf), k


def test_every_tuning_and_every_replay_agree(device):
    sc, lights, views = _views()
    tris9, counts = meshes_to_arrays(sc.meshes)
    scene = Scene(device, Flat.build_arrays(tris9, counts))
    opts = _lib.default_options(hierarchy=2)
    batch = Batch(scene, views, lights)
    batch.run(opts, Batch.tuning(8, 0))
    ref = [batch.read(k) for k in range(len(views))]

    def same(tag):
        for k in range(len(views)):
            got = batch.read(k)
            for a, b in zip(got, ref[k]):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (tag, k)

    for tw in (1, 4, 8, 16, 64):
        for pb in (0, 7, 256):
            batch.run(opts, Batch.tuning(tw, pb))
            same((tw, pb))
    # occupancy (one wave per SIMD: LDS reserved) and the ray stack's depth
    for wv, rs in ((1, 64), (2, 32), (1, 32), (0, 0)):
        for tw, pb in ((16, 0), (8, 512)):
            batch.run(opts, Batch.tuning(tw, pb, wv, rs))
            same((tw, pb, wv, rs))
    best, ms = batch.autotune(opts)
    assert best.tile_w in (4, 8, 16, 32, 64) and best.waves_per_simd in (1, 2) and best.ray_stack in (32, 64)
    assert ms > 0
    batch.run(opts, best)
    same("autotuned")
    # Graph replays, with eager work (a single-view query) and synchronisation between them.
    batch.capture(opts, Batch.tuning(16, 64))
    for r in range(4):
        batch.replay()
        same(("replay", r))
        scene.render_view(views[0][0], 8, 8, lights[:1], opts)


def test_light_edits_change_only_what_they_touch(device):
    """A recolour changes no query (results are per unit radiance); a moved
    light changes its own queries, and equals a fresh batch with that light."""
    sc, lights, views = _views()
    tris9, counts = meshes_to_arrays(sc.meshes)
    scene = Scene(device, Flat.build_arrays(tris9, counts))
    opts = _lib.default_options(hierarchy=2)
    batch = Batch(scene, views, lights)
    batch.run(opts, Batch.tuning(8, 0))
    before = [batch.read(k) for k in range(len(views))]
    recolored = lights.copy()
    recolored[:, 12:15] *= np.float32(3.0)
    batch.set_lights(recolored)
    batch.run(opts, Batch.tuning(8, 0))
    for k in range(len(views)):
        for a, b in zip(batch.read(k), before[k]):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    moved = lights.copy()
    moved[1, 0:3] += np.float32(0.3)
    batch.set_lights(moved)
    batch.run(opts, Batch.tuning(8, 0))
    fresh = Batch(scene, views, moved)
    fresh.run(opts, Batch.tuning(8, 0))
    for k in range(len(views)):
        for a, b in zip(batch.read(k), fresh.read(k)):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_a_new_layout_drops_the_captured_graph(device):
    """A captured graph holds the view table and the item buffers of its
    layout; laying the batch out for another tile width (a run or an
    autotune) drops the graph, so a replay is refused until the batch is
    captured again, and the new capture replays to the same bits."""
    from continuum_amd._lib import CnqError

    sc, lights, views = _views()
    tris9, counts = meshes_to_arrays(sc.meshes)
    scene = Scene(device, Flat.build_arrays(tris9, counts))
    opts = _lib.default_options(hierarchy=2)
    batch = Batch(scene, views, lights)
    batch.run(opts, Batch.tuning(16, 0))
    ref = [batch.read(k) for k in range(len(views))]
    batch.capture(opts, Batch.tuning(16, 64))
    batch.replay()
    batch.run(opts, Batch.tuning(8, 0))
    batch.run(opts, Batch.tuning(16, 0))  # back at the captured width: still a new layout
    with pytest.raises(CnqError):
        batch.replay()
    batch.capture(opts, Batch.tuning(16, 64))
    for _ in range(2):
        batch.replay()
        for k in range(len(views)):
            for a, b in zip(batch.read(k), ref[k]):
                assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
