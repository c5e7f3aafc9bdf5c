"""GPU: the sky light (WGSL :1411-1510: six cube-face mask cones, meshes as
independent occluders, the receiver's horizon culling the walk) through the
C ABI against the oracle, bit for bit: every face's irradiance and full
bins and their sum, on the room (camera receivers), the rock field
(prefiltered triangle hierarchy, 1-bin splats) and the instanced view-cache
world (transforms: cones and horizons carried into each object's frame,
prefiltered object hierarchy), and with the maximum diagnostic and without
the horizon cull."""
import numpy as np
import pytest

from continuum_amd import _lib
from continuum_amd.kernel import Flat, MaskConeKernel, Scene, receivers_array
from continuum_amd.scenes import meshes_to_arrays, rock_field, room_scene, vcache_world
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _same(g, r, what):
    g, r = np.ascontiguousarray(g), np.ascontiguousarray(r)
    if g.dtype == np.float32:
        g, r = g.view(np.uint32), r.view(np.uint32)
    bad = np.nonzero(g.reshape(g.shape[0], -1) != r.reshape(r.shape[0], -1))
    assert bad[0].size == 0, f"{what}: {bad[0].size} differ, first at {list(zip(*bad))[:5]}"


def _check(flat, scene, recv, hierarchy=2, prefilter=0.0, object_cov=O.NONE, **sky):
    o = O.opts_array(hierarchy=hierarchy, prefilter=prefilter, object_cov=object_cov)
    p = _lib.default_options(hierarchy=hierarchy, prefilter=prefilter, object_cov=object_cov)
    so = O.sky_options(**sky)
    ps = _lib.default_sky_options(max_combine=int(sky.get("max_combine", False)),
                                  horizon_cull=int(sky.get("horizon_cull", True)),
                                  prefilter_scale=float(sky.get("prefilter_scale", 1.0)))
    g = scene.sky(recv, p, ps)
    r = O.sky(flat, recv, o, so)
    print(f"sky {len(recv)} receivers {sky}: mean {float(np.mean(r[0])):.4f}", flush=True)
    _same(g[1], r[1], "faces")
    _same(g[2], r[2], "full bins")
    _same(g[0], r[0], "sum")
    return g, r


@pytest.fixture(scope="module")
def room(device):
    sc = room_scene()
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    _, recv, of = O.primary(flat, sc.camera, sc.width, sc.height, sc.lights)
    assert of == 0
    return flat, Scene(device, flat), recv[::4]


def test_room_sky_matches(room):
    flat, scene, recv = room
    g, r = _check(flat, scene, recv)
    assert np.all(r[0] >= 0) and float(np.max(r[0])) < np.pi


@pytest.mark.parametrize("sky", [dict(max_combine=True), dict(horizon_cull=False)])
def test_room_sky_variants_match(room, sky):
    flat, scene, recv = room
    _check(flat, scene, recv[::4], **sky)


def test_rock_field_sky_matches(device):
    sc = rock_field(width=320, height=180)
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    _, recv, of = O.primary(flat, sc.camera, sc.width, sc.height, sc.lights)
    assert of == 0
    _check(flat, Scene(device, flat), recv[::6], prefilter=sc.prefilter)


def test_instanced_world_sky_matches(device):
    tris9, counts, pl, surf, lights, movable = vcache_world()
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True, movable=movable)
    rng = np.random.default_rng(5)
    n = 3000
    pts = np.stack([rng.uniform(-9, 9, n), rng.uniform(-9, 9, n), rng.uniform(1e-3, 1.2, n)], axis=1)
    nrm = rng.normal(size=(n, 3)) * [0.4, 0.4, 1.0]
    nrm[:, 2] = np.abs(nrm[:, 2])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    recv = receivers_array(np.concatenate([pts, nrm], axis=1))
    recv[::97, 7] = -1.0  # a few that trace nothing
    g, r = _check(f, Scene(device, f), recv, prefilter=1.0, object_cov=f.object_cov)
    assert np.all(r[0][::97] == 0.0)
    _check(f, Scene(device, f), recv[::3], prefilter=1.0, object_cov=f.object_cov, prefilter_scale=2.0)


def test_kernel_object_sky_irradiance(device):
    k = MaskConeKernel(device)
    sc = room_scene()
    recv = [((2.0, 2.0, 1.5), (0.0, 0.0, 1.0)), ((1.0, 1.0, 0.01), (0.0, 0.0, 1.0))]
    out = k.sky_irradiance(recv, sc.meshes)
    assert out.shape == (2,) and np.all(out >= 0) and out[0] < 1e-5 and np.all(out < np.pi)


def test_view_tiled_sky_equals_the_list_launch_and_the_oracle(device):
    """The sky over a view's receivers in pixel tiles (cnq_sky_view_irradiance_device)
    equals the list launch and the oracle bit for bit, for tiles of 8 x 8,
    16 x 4 and 64 x 1, on a view whose size is not a multiple of the tile."""
    import torch

    sc = rock_field(width=100, height=37)
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    scene = Scene(device, flat)
    p = _lib.default_options(hierarchy=2, prefilter=sc.prefilter)
    dev = torch.device("cuda:0")
    w, h = sc.width, sc.height
    n = w * h
    cam = torch.from_numpy(sc.camera).to(dev)
    lights = torch.from_numpy(np.ascontiguousarray(sc.lights)).to(dev)
    hits = torch.zeros((n, 4), dtype=torch.int32, device=dev)
    recv = torch.zeros((n, 8), dtype=torch.float32, device=dev)
    scene.primary_device(cam, w, h, lights, p, hits, recv)
    torch.cuda.synchronize()
    want = scene.sky(recv.cpu().numpy(), p)
    r = O.sky(flat, recv.cpu().numpy(), O.opts_array(hierarchy=2, prefilter=sc.prefilter))
    _same(want[0], r[0], "list launch sum")
    for tw in (8, 16, 64):
        out = torch.zeros(n, dtype=torch.float32, device=dev)
        faces = torch.zeros(6 * n, dtype=torch.float32, device=dev)
        full = torch.zeros(6 * n, dtype=torch.int64, device=dev)
        for _ in range(2):  # the second launch ordered by the first one's times
            scene.sky_view_device(recv, w, h, p, out, faces, full, tile_w=tw)
        torch.cuda.synchronize()
        _same(out.cpu().numpy(), want[0], f"sum, tile {tw}")
        _same(faces.cpu().numpy().reshape(6, n), want[1], f"faces, tile {tw}")
        _same(full.cpu().numpy().view(np.uint64).reshape(6, n), want[2], f"full, tile {tw}")
    assert (want[0] > 0).mean() > 0.5
