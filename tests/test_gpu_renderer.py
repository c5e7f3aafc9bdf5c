"""GPU: the `Renderer` facade (renderer.rs:196-660) over the view cache and
the plain per-pixel path: without the cache its colours are albedo times
the lights' radiance-weighted irradiance over pi, the irradiance the
oracle's per-pixel cones give bit for bit; with the cache the view settles
(`is_settled`), every page shown equals a fresh bake (`verify_cache`),
re-colouring a light scales its contribution with nothing baked, and a
moved instance or an edited light re-bakes what it reaches and stays
verified."""
import numpy as np
import pytest

from continuum_amd.renderer import Renderer, RendererOptions
from continuum_amd.scenes import camera_uniform, vcache_world
from oracle import oracle as O

pytestmark = pytest.mark.gpu

W, H = 96, 64


def _world():
    tris9, counts, pl, surf, lights, movable = vcache_world()
    lights = np.array(lights, np.float32, copy=True)
    lights[0, 12:15] = [3.0, 2.5, 2.0]
    lights[1, 12:15] = [1.0, 1.0, 4.0]
    cam = camera_uniform([-7.0, -7.0, 3.0], [2.0, 2.0, 0.0], 60.0, W, H)
    return tris9, counts, pl, surf, lights, movable, cam


def test_plain_colours_are_the_oracle_irradiance_weighted(device):
    tris9, counts, pl, surf, lights, movable, cam = _world()
    r = Renderer(device, (W, H), tris9, counts, pl, surf, lights, RendererOptions(cache=False))
    r.render(cam)
    colors = r.read_colors().reshape(-1, 3)
    hits, recv, _ = O.primary(r.flat, cam, W, H, lights)
    o = O.opts_array(hierarchy=2, prefilter=1.0, object_cov=r.flat.object_cov)
    vis, *_ = O.irradiance(r.flat, recv, lights, o)
    vis = vis.reshape(len(lights), -1).astype(np.float64)
    e = lights[:, 12:15].astype(np.float64).T @ vis
    surface = (hits[:, 0] != O.NONE) & (hits[:, 3] == 0)
    want = 0.8 * e[:, surface].T / np.pi
    assert surface.sum() > W * H // 2
    assert np.allclose(colors[surface], want, rtol=1e-5, atol=1e-7)


def test_cached_view_settles_verifies_and_follows_edits(device):
    tris9, counts, pl, surf, lights, movable, cam = _world()
    opts = RendererOptions(cache=True, movable=movable, pages=4096, base_texel=0.1)
    r = Renderer(device, (W, H), tris9, counts, pl, surf, lights, opts)
    for frame in range(60):
        r.render(cam)
        if r.is_settled(2):
            break
    print(f"settled after {frame + 1} frames: {r.stats()}", flush=True)
    assert r.is_settled(2)
    differ, compared = r.verify_cache()
    assert differ == 0 and compared > 0
    base = r.read_colors()
    # re-colouring light 1 (blue) by 2: nothing baked, its share doubles
    r.set_light_radiance(1, lights[1, 12:15] * 2.0)
    r.render(cam)
    st = r.stats()
    assert st["planned_cones"] == 0 and st["rebaked_cones"] == 0
    assert r.read_colors()[..., 2].sum() > base[..., 2].sum()
    # a movable instance moves: its pages and the shadows it casts re-bake
    m = movable[0]
    r.set_transform(m, pl[m]["rotation"], pl[m]["translation"] + np.float32([0.8, 0.4, 0.0]))
    r.render(cam)
    assert r.stats()["rebaked_cones"] > 0 or r.stats()["stale_pages"] > 0
    for _ in range(20):
        r.render(cam)
    assert r.verify_cache()[0] == 0
    # the rect light moves: its channel re-bakes, still verified
    moved = lights[0].copy()
    moved[0:3] += np.float32([0.7, -0.4, 0.0])
    r.set_light(0, moved)
    r.render(cam)
    assert r.verify_cache()[0] == 0
    assert np.isfinite(r.read_colors()).all()


def test_a_moving_light_goes_through_the_ledger(device):
    """RendererOptions::moving_lights: set_light on a moving light goes
    through the cache's error ledger, and the reference's verify_ledger
    finds every page shown within its bound and epsilon."""
    tris9, counts, pl, surf, lights, movable, cam = _world()
    opts = RendererOptions(cache=True, pages=4096, base_texel=0.1, moving_lights=[0])
    r = Renderer(device, (W, H), tris9, counts, pl, surf, lights, opts)
    for frame in range(60):
        r.render(cam)
        if r.is_settled(2):
            break
    assert r.is_settled(2)
    for k in range(1, 6):
        moved = lights[0].copy()
        moved[0:3] += np.float32([0.01 * k, -0.005 * k, 0.0])
        r.set_light(0, moved)
        r.render(cam)
    chk = r.verify_ledger(0)
    assert chk["pages"] > 0 and chk["violations"] == 0 and chk["over_epsilon"] == 0, chk
    assert np.isfinite(r.read_colors()).all()
