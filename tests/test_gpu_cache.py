"""GPU: the lighting cache through the C ABI against the oracle, bit for bit.

The baked lightmap equals the oracle's per-texel query; after each update
the blocks `mark_blocks` marked equal the oracle's, and every texel equals
both the oracle's incremental update and a fresh full-scene query of the
moved scene (the reference's invariant: the cache equals tracing the whole
scene, WGSL :2489; `verify_cache` compares f32 bit patterns,
instanced.rs:3721).
"""
import numpy as np
import pytest

from continuum_amd import _lib
from continuum_amd._lib import CnqError
from continuum_amd.kernel import Flat, LiveLightmap, Scene
from continuum_amd.lightmap import Lightmap
from continuum_amd.scenes import lit_room, meshes_to_arrays
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(device, boxes=2, size=1.0, texel=0.1):
    sc = lit_room(boxes=boxes, size=size)
    dyn = sc.notes["dynamic"]
    lm = Lightmap.from_meshes(sc.meshes, texel * size, lambda m: not dyn[m])
    tris9, counts = meshes_to_arrays(sc.meshes)
    scene = Scene(device, Flat.build_arrays(tris9, counts))
    return sc, dyn, lm, tris9, counts, scene


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def test_bake_equals_the_oracle(device):
    sc, dyn, lm, tris9, counts, scene = _setup(device)
    cache = LiveLightmap(scene, lm.texels, lm.blocks, sc.lights[0], dyn)
    got = cache.read_texels()
    want, *_ = O.irradiance(O.OracleFlat(tris9, counts), lm.texels, sc.lights[0][None], O.opts_array(hierarchy=2))
    assert np.array_equal(_bits(got), _bits(want))
    assert (want > 0).sum() > 1000


def test_updates_match_the_oracle_and_a_full_recompute(device):
    sc, dyn, lm, tris9, counts, scene = _setup(device)
    light = sc.lights[0]
    opts = O.opts_array(hierarchy=2)
    flat0 = O.OracleFlat(tris9, counts)
    planes = O.block_planes(lm.texels, lm.blocks, light)
    slots = O.block_slots(planes)
    sn, si = O.mesh_hierarchy(tris9, counts, keep=[not d for d in dyn])
    cov = O.bake_static(flat0, sn, si, lm.texels, slots, light, opts)
    baked, *_ = O.irradiance(flat0, lm.texels, light[None], opts)
    cache = LiveLightmap(scene, lm.texels, lm.blocks, light, dyn, irradiance=baked)
    off = [0] + [int(x) for x in np.cumsum(counts)]
    moved = tris9.copy()
    current = baked.copy()
    steps = [[(7, 0.45, 0.2)], [(8, -0.6, 0.3)], [(7, -0.2, -0.5), (8, 0.1, 0.9)]]
    for step, moves in enumerate(steps):
        changes, swept = [], []
        for m, dx, dy in moves:
            old = moved[off[m]:off[m + 1]].copy()
            new = old.copy()
            new[:, 0::3] += np.float32(dx)
            new[:, 1::3] += np.float32(dy)
            moved[off[m]:off[m + 1]] = new
            changes.append((m, new))
            o, n = old.reshape(-1, 3), new.reshape(-1, 3)
            swept.append([np.minimum(o.min(0), n.min(0)), np.maximum(o.max(0), n.max(0))])
        cache.update_meshes(changes)
        got, marked = cache.read()
        flat_m = O.OracleFlat(moved, counts)
        want_marked = O.mark_blocks(planes, swept)
        assert np.array_equal(marked, want_marked), step
        moving = []
        for m, d in enumerate(dyn):
            if d:
                p = moved[off[m]:off[m + 1]].reshape(-1, 3)
                lo, hi = p.min(0), p.max(0)
                pad = np.float32(1e-5) * np.float32(max(float(np.abs(lo).max()), float(np.abs(hi).max()), 1.0))
                moving.append((m, lo - pad, hi + pad))
        rb = O.rebake(flat_m, lm.texels, slots, cov, want_marked, light, moving, opts)
        for b in want_marked:
            current[64 * b:64 * b + 64] = rb[64 * b:64 * b + 64]
        assert np.array_equal(_bits(got), _bits(current)), step
        fresh, *_ = O.irradiance(flat_m, lm.texels, light[None], opts)
        assert np.array_equal(_bits(got), _bits(fresh)), step
        assert cache.rebaked_blocks() == len(want_marked)
    # the scene itself now holds the moved meshes: its own queries see them
    vis, _, _, _ = scene.query(lm.texels, light[None], _lib.default_options(hierarchy=2))
    assert np.array_equal(_bits(vis), _bits(current))


def test_update_rules(device):
    sc, dyn, lm, tris9, counts, scene = _setup(device)
    cache = LiveLightmap(scene, lm.texels, lm.blocks, sc.lights[0], dyn)
    before = cache.read_texels()
    off = [0] + [int(x) for x in np.cumsum(counts)]
    # an unchanged mesh moves nothing and re-bakes nothing
    cache.update_meshes([(7, tris9[off[7]:off[8]])])
    assert np.array_equal(_bits(cache.read_texels()), _bits(before))
    assert cache.rebaked_blocks() == 0
    # a static mesh holds texels and cannot move
    with pytest.raises(CnqError):
        cache.update_meshes([(6, tris9[off[6]:off[7]])])
    with pytest.raises(ValueError):
        cache.update_meshes([(7, tris9[off[7]:off[8] - 1])])
    assert np.array_equal(_bits(cache.read_texels()), _bits(before))
