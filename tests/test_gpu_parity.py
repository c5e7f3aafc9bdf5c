"""GPU parity: the HIP path through the C ABI against the oracle on the same
seeded inputs. Integer results (masks, hit indices) and the FP32 results
(irradiance bits, coverage bits, refit boxes) must match bit for bit: the
kernels and the oracle evaluate the same operations in the same order, IEEE
single precision without fused multiply-adds, and the same arctangent.
Where a tolerance appears it is the reference's own (gpu.rs:294-336: 1e-4 of
the light's weight against its CPU model)."""
import numpy as np
import pytest

from continuum_amd import _lib
from continuum_amd.kernel import Flat, MaskConeKernel, Scene, receivers_array
from continuum_amd.scenes import (box_mesh, camera_uniform, meshes_to_arrays, rect_light, rock_field, room_scene,
                                  sphere_light)
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _assert_same(gpu, ref, what):
    g, r = np.asarray(gpu), np.asarray(ref)
    if g.dtype == np.float32:
        g, r = _bits(g), _bits(r)
    bad = np.nonzero(g != r)[0] if g.ndim == 1 else np.nonzero(np.any(g != r, axis=tuple(range(1, g.ndim))))[0]
    assert bad.size == 0, f"{what}: {bad.size} of {len(g)} differ, first at {bad[:5].tolist()}"


def _opts(**kw):
    o = O.opts_array(**{k: v for k, v in kw.items() if k in ("hierarchy", "prefilter", "wide_extent",
                                                            "light_cone_cap", "wide_cone_cap", "profile_stage")})
    p = _lib.default_options(**kw)
    return o, p


def _run_both(scene_meshes, receivers, lights, want_cov=False, device=None, **kw):
    tris9, counts = meshes_to_arrays(scene_meshes)
    flat = Flat.build_arrays(tris9, counts)
    scene = Scene(device, flat)
    o, p = _opts(**kw)
    g = scene.query(receivers, lights, p, want_cov)
    r = O.irradiance(flat, receivers, lights, o, want_cov)
    counters, _ = scene.counters()
    return g, r, counters, scene, flat


def test_room_primary_hits_and_cones_match(device):
    sc = room_scene()
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    scene = Scene(device, flat)
    o, p = _opts(hierarchy=2)
    hits, vis, full, touched = scene.render_view(sc.camera, sc.width, sc.height, sc.lights, p)
    rh, rrecv, rof = O.primary(flat, sc.camera, sc.width, sc.height, sc.lights)
    assert rof == 0
    _assert_same(hits, rh, "hits")
    rv, rf, rt, _, capped, ovf = O.irradiance(flat, rrecv, sc.lights, o)
    assert ovf == 0
    _assert_same(full, rf, "full masks")
    _assert_same(touched, rt, "touched masks")
    _assert_same(vis, rv, "visible")
    c, _ = scene.counters()
    assert c.tolist() == [int(capped.sum()), 0, 0]
    lit = rrecv[:, 7] >= 0
    assert lit.sum() > 0.9 * len(lit)  # the view sees the room, not the sky
    assert (rt[lit] != 0).any() and (rf[lit] != 0).any() and (vis[lit] > 0).any()


@pytest.mark.parametrize("hierarchy", [0, 1, 2])
def test_room_cone_queries_match_every_hierarchy(device, hierarchy):
    sc = room_scene()
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    _, rrecv, _ = O.primary(flat, sc.camera, sc.width, sc.height, sc.lights)
    sub = rrecv[::16] if hierarchy < 2 else rrecv
    (gv, gf, gt, gc), (rv, rf, rt, rc, capped, ovf), counters, _, _ = _run_both(
        sc.meshes, sub, sc.lights, want_cov=True, device=device, hierarchy=hierarchy)
    _assert_same(gf, rf, "full")
    _assert_same(gt, rt, "touched")
    _assert_same(gv, rv, "visible")
    _assert_same(gc, rc, "coverage")


CPU_SCENE = [box_mesh([0.30, 0.10, 0.55], [0.70, 0.35, 0.58]), box_mesh([0.62, 0.40, 0.30], [0.85, 0.60, 0.33]),
             box_mesh([0.15, 0.45, 0.20], [0.25, 0.55, 0.30]), box_mesh([0.40, 0.65, 0.45], [0.48, 0.90, 0.62])]
FLOOR_LIGHT = rect_light([0.4, 0.4, 1.0], [0.2, 0, 0], [0, 0.2, 0])


def _floor(size=48):
    k = np.arange(size * size)
    pts = [(((i % size) + 0.5) / size, ((i // size) + 0.5) / size, 0.0) for i in k]
    return [(p, (0.0, 0.0, 1.0)) for p in pts]


def test_cones_match_the_cpu(device):
    """crates/continuum-gltf/tests/gpu.rs:294-336 through the kernel mirror:
    within 1e-4 of the light's weight of the CPU model, and bit-exact with
    the oracle."""
    kernel = MaskConeKernel(device)
    recv = _floor()
    gpu, _ = kernel.irradiance(recv, FLOOR_LIGHT, CPU_SCENE, 1)
    r6 = np.array([[*p, *n] for p, n in recv], np.float32)
    cpu, total = O.cone3_visible(r6, FLOOR_LIGHT, CPU_SCENE)
    assert float(np.max(np.abs(cpu - gpu) / total)) <= 1e-4
    t9, c = meshes_to_arrays(CPU_SCENE)
    ref, *_ = O.irradiance(O.OracleFlat(t9, c), receivers_array(recv), FLOOR_LIGHT[None], O.opts_array(hierarchy=0))
    _assert_same(gpu, ref, "visible")
    cov = kernel.coverage(recv, FLOOR_LIGHT, CPU_SCENE)
    _, _, _, rcov, _, _ = O.irradiance(O.OracleFlat(t9, c), receivers_array(recv), FLOOR_LIGHT[None],
                                       O.opts_array(hierarchy=0), want_cov=True)
    _assert_same(cov, rcov, "coverage")


def test_rock_field_small_view_with_sun_and_prefilter(device):
    """The rock field's meshes, lights and prefilter at 160 x 90 (a sphere
    light, prefiltered triangle-hierarchy splats)."""
    sc = rock_field(width=160, height=90)
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    scene = Scene(device, flat)
    o, p = _opts(hierarchy=2, prefilter=sc.prefilter)
    cam = sc.camera
    hits, vis, full, touched = scene.render_view(cam, 160, 90, sc.lights, p)
    rh, rrecv, _ = O.primary(flat, cam, 160, 90, sc.lights)
    _assert_same(hits, rh, "hits")
    rv, rf, rt, _, capped, ovf = O.irradiance(flat, rrecv, sc.lights, o)
    assert ovf == 0
    _assert_same(full, rf, "full")
    _assert_same(touched, rt, "touched")
    _assert_same(vis, rv, "visible")
    assert (rf != 0).any()


def test_refit_matches_the_oracle_and_queries_follow(device):
    sc = room_scene()
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    scene = Scene(device, flat)
    moved = flat.tris.copy()
    rng = np.random.default_rng(5)
    # displace the two boxes (the last 24 triangles) and sag the floor
    moved[-24:, [0, 4, 8]] += np.float32(0.31)
    moved[-24:, [2, 6, 10]] += np.float32(0.2)
    floor = slice(0, counts[0])
    moved[floor, [2, 6, 10]] += rng.uniform(0, 0.05, (counts[0], 3)).astype(np.float32)
    scene.refit(moved)
    blas, cov = scene.read_blas()
    rblas, rcov = O.refit(flat, moved)
    _assert_same(blas, rblas, "refit nodes")
    _assert_same(cov.reshape(-1), rcov.reshape(-1), "refit coverage")
    # queries on the refit scene equal the oracle on the same refit arrays
    new = Flat(moved, flat.ranges, rblas, flat.blas_items, rcov, flat.nodes, flat.items)
    new9 = moved.reshape(-1, 3, 4)[:, :, :3].reshape(-1, 9)
    rebuilt = O.OracleFlat(new9, counts)
    new.nodes, new.items = rebuilt.nodes, rebuilt.items
    o, p = _opts(hierarchy=2)
    _, rrecv, _ = O.primary(new, sc.camera, sc.width, sc.height, sc.lights)
    g = scene.query(rrecv, sc.lights, p)
    r = O.irradiance(new, rrecv, sc.lights, o)
    _assert_same(g[1], r[1], "full after refit")
    _assert_same(g[0], r[0], "visible after refit")


def test_edge_cases(device):
    sc = room_scene()
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    scene = Scene(device, flat)
    o, p = _opts(hierarchy=2)
    light = sc.lights
    rng = np.random.default_rng(11)
    recv = np.zeros((203, 8), np.float32)  # not a multiple of 64
    recv[:, 0:3] = rng.uniform([0.1, 0.1, 0.01], [3.9, 3.9, 2.99], (203, 3))
    recv[:, 6] = 1.0
    recv[::7, 7] = -1.0                     # inactive receivers
    recv[3, 2] = 2.9985                     # just under the light's plane: a wide cone
    recv[4, 2] = 3.0005                     # above the one-sided light: behind it
    recv[5, 4:7] = [0, 0, -1]               # facing away
    recv[6, 0:3] = [2.0, 2.0, 2.9989]       # right below the light: wide, near its plane
    g = scene.query(recv, light, p, want_cov=True)
    r = O.irradiance(flat, recv, light, o, want_cov=True)
    for a, b, w in zip(g[:4], r[:4], ("visible", "full", "touched", "coverage")):
        _assert_same(a, b, w)
    assert g[0][::7].max() == 0 and g[1][::7].max() == 0
    assert g[0][4] == 0 and g[0][5] == 0
    # zero receivers and zero lights are no-ops
    e = scene.query(np.zeros((0, 8), np.float32), light, p)
    assert e[0].size == 0
    # two lights, one a sphere
    two = np.stack([light[0], sphere_light([2.0, 2.0, 2.5], 0.2)])
    g = scene.query(recv, two, p)
    r = O.irradiance(flat, recv, two, o)
    _assert_same(g[0], r[0], "two lights visible")
    _assert_same(g[1], r[1], "two lights full")


def test_caps_and_wide_cones_match(device):
    """Receivers centimetres from a large light's plane: clipped wide cones
    and capped walks (WGSL :1000-1059, :1296-1304). Capped walks may differ
    from a lone walk only in which nodes they reached; uncapped queries match
    bit for bit."""
    floor = np.concatenate([box_mesh([x, y, 0], [x + 0.2, y + 0.2, 0.1]) for x in np.arange(0, 4, 0.25)
                            for y in np.arange(0, 4, 0.25)])
    big = rect_light([-20, -20, 0.3], [40, 0, 0], [0, 40, 0])
    rng = np.random.default_rng(2)
    recv = np.zeros((512, 8), np.float32)
    recv[:, 0:2] = rng.uniform(0, 4, (512, 2))
    recv[:, 2] = rng.uniform(0.12, 0.2999, 512)
    recv[:, 6] = 1.0
    for caps in ((4096, 512), (64, 16)):
        (gv, gf, gt, _), (rv, rf, rt, _, capped, _), counters, _, _ = _run_both(
            [floor], recv, big[None], device=device, hierarchy=2, light_cone_cap=caps[0], wide_cone_cap=caps[1])
        ok = capped == 0
        _assert_same(gf[ok], rf[ok], "full (uncapped)")
        _assert_same(gv[ok], rv[ok], "visible (uncapped)")
        assert counters[0] >= 0


def test_repeated_view_launches_take_the_longest_items_first_with_equal_results(device):
    """The view launch records each item's time and orders the next launch
    of the same view longest first (and the sky's the same way): the
    ordered launches' results equal the first, natural-order launch's bit
    for bit, as every item writes only its own queries; a launch on another
    stream runs in natural order and agrees too."""
    import torch

    sc = rock_field(width=320, height=180)
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    scene = Scene(device, flat)
    p = _lib.default_options(hierarchy=2, prefilter=sc.prefilter)
    dev = torch.device("cuda:0")
    w, h, nl = sc.width, sc.height, len(sc.lights)
    cam = torch.from_numpy(sc.camera).to(dev)
    lights = torch.from_numpy(np.ascontiguousarray(sc.lights)).to(dev)
    hits = torch.zeros((w * h, 4), dtype=torch.int32, device=dev)
    recv = torch.zeros((w * h, 8), dtype=torch.float32, device=dev)
    scene.primary_device(cam, w, h, lights, p, hits, recv)
    runs = []
    other = torch.cuda.Stream()
    for k in range(4):
        vis = torch.zeros(w * h * nl, dtype=torch.float32, device=dev)
        full = torch.zeros(w * h * nl, dtype=torch.int64, device=dev)
        touched = torch.zeros(w * h * nl, dtype=torch.int64, device=dev)
        stream = other if k == 3 else torch.cuda.current_stream()
        if k == 3:
            other.wait_stream(torch.cuda.current_stream())
        scene.view_query_device(recv, w, h, lights, p, vis, full, touched, tile_w=16, stream=stream)
        torch.cuda.synchronize()
        runs.append((vis.cpu().numpy(), full.cpu().numpy(), touched.cpu().numpy()))
    for k in range(1, 4):
        for a, b, what in zip(runs[0], runs[k], ("visible", "full", "touched")):
            assert np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32)), \
                f"launch {k}: {what} differs"
    sky = []
    for k in range(3):
        out = torch.zeros(w * h, dtype=torch.float32, device=dev)
        scene.sky_device(recv, p, out)
        torch.cuda.synchronize()
        sky.append(out.cpu().numpy())
    assert np.array_equal(sky[0].view(np.uint32), sky[1].view(np.uint32))
    assert np.array_equal(sky[0].view(np.uint32), sky[2].view(np.uint32))


def test_a_captured_view_launch_replays_after_other_views(device):
    """A view launch captured into a graph runs in natural order and keeps
    no order buffer of the scene's: launches of a larger and a smaller view
    between the replays (which reallocate and rewrite the scene's order) do
    not change what a replay computes. Each replay follows eager work and a
    synchronization, and equals the eager launch bit for bit."""
    import torch

    sc = room_scene(width=96, height=64)
    tris9, counts = meshes_to_arrays(sc.meshes)
    scene = Scene(device, Flat.build_arrays(tris9, counts))
    p = _lib.default_options(hierarchy=2)
    dev = torch.device("cuda:0")
    lights = torch.from_numpy(np.ascontiguousarray(sc.lights)).to(dev)

    def view(w, h):
        # the room's camera (eye (0.4, 0.4, 1.6) toward (3.6, 3.6, 0.8), 60 degrees) at w x h
        cam = torch.from_numpy(camera_uniform((0.4, 0.4, 1.6), (3.6, 3.6, 0.8), 60.0, w, h)).to(dev)
        hits = torch.zeros((w * h, 4), dtype=torch.int32, device=dev)
        recv = torch.zeros((w * h, 8), dtype=torch.float32, device=dev)
        scene.primary_device(cam, w, h, lights, p, hits, recv)
        outs = (torch.zeros(w * h, dtype=torch.float32, device=dev), torch.zeros(w * h, dtype=torch.int64, device=dev),
                torch.zeros(w * h, dtype=torch.int64, device=dev))
        return recv, outs

    w, h = sc.width, sc.height
    recv, eager = view(w, h)
    for _ in range(2):  # the second launch runs ordered by the first one's times
        scene.view_query_device(recv, w, h, lights, p, *eager, tile_w=16)
    torch.cuda.synchronize()
    want = [t.cpu().numpy() for t in eager]
    _, graph_out = view(w, h)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g):
            scene.view_query_device(recv, w, h, lights, p, *graph_out, tile_w=16)
    torch.cuda.current_stream().wait_stream(side)
    big_recv, big = view(2 * w, 2 * h)
    small_recv, small = view(w // 2, h // 2)
    for r in range(3):
        for t in graph_out:
            t.zero_()
        vw, vh, vr, vo = (2 * w, 2 * h, big_recv, big) if r % 2 == 0 else (w // 2, h // 2, small_recv, small)
        scene.view_query_device(vr, vw, vh, lights, p, *vo, tile_w=16)
        scene.view_query_device(vr, vw, vh, lights, p, *vo, tile_w=16)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(graph_out, want):
            assert np.array_equal(np.ascontiguousarray(a.cpu().numpy()).view(np.uint32),
                                  np.ascontiguousarray(b).view(np.uint32)), r
