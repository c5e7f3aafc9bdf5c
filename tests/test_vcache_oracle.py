"""CPU: the view cache's page passes as the oracle restates them
(crates/continuum-gpu/src/vcache.rs; mask_cones.wgsl vc_bake :3432,
vc_fill :3693, vc_stale :4233, vc_stale_texels :4385; hierarchy.rs:418
shadow_region).

Pins: the reference's own page-table tests (the hash's known value,
vcache.rs:4351-4355, and its deletion test, :4308-4349); shadow regions
and the hull test against brute force over the segments they bound; the
chart grid and a flat chart's texel places in closed form; an unoccluded
texel's ratio is exactly one; skipping a triangle equals removing it; and
the cache's invariant, as the reference's `verify_cache` checks it
(instanced.rs:3721): after objects move and the stale passes run, every
page still shown equals a page baked afresh in the moved scene, bit for
bit, whatever the append capacity.
"""
import numpy as np
import pytest

from continuum_amd.kernel import Flat
from continuum_amd.scenes import quad_pair, rect_light, vcache_world
from oracle import oracle as O

NONE = O.NONE


def test_page_hash_known_values():
    assert O.vc_hash([0, 0, 0]) == 0
    assert O.vc_hash([1, 0, 0]) != O.vc_hash([0, 1, 0])
    # the formula's arithmetic (wrapping u32) for one key, by hand
    k = [7, 0x01000003, 5 | (9 << 16)]
    h = ((k[0] * 0x9E3779B1) ^ (k[1] * 0x85EBCA77) ^ (k[2] * 0xC2B2AE3D)) & 0xFFFFFFFF
    h ^= h >> 15
    h = (h * 0x2C1B3C6D) & 0xFFFFFFFF
    h ^= h >> 12
    assert O.vc_hash(k) == h


def test_page_table_deletion_keeps_every_page_findable():
    moves = O.vc_table_check(1 << 10, 20_000)
    assert moves > 100, moves


def _segments_points(rng, box, n):
    lo, hi = np.asarray(box[:3]), np.asarray(box[3:])
    return lo + rng.random((n, 3)) * (hi - lo)


def test_shadow_region_holds_every_shadowing_point():
    """Points m + t (m - l), t >= 0, for m in the mover box and l in the
    light box, that fall in the scene box lie in the region; no region means
    no such point."""
    rng = np.random.default_rng(11)
    scene = np.array([-10, -10, -1, 10, 10, 8], np.float32)
    checked = 0
    for _ in range(300):
        c = rng.uniform(-6, 6, 3)
        e = rng.uniform(0.1, 1.5, 3)
        m = np.concatenate([c - e, c + e]).astype(np.float32)
        lc = rng.uniform(-8, 8, 3) + np.array([0, 0, 6])
        le = rng.uniform(0.05, 1.0, 3)
        lb = np.concatenate([lc - le, lc + le]).astype(np.float32)
        r = O.shadow_region(m, lb, scene)
        pm = _segments_points(rng, m, 200).astype(np.float64)
        pl = _segments_points(rng, lb, 200).astype(np.float64)
        t = rng.exponential(2.0, (200, 1))
        q = pm + t * (pm - pl)
        inside = np.all((q >= scene[:3]) & (q <= scene[3:]), axis=1)
        if r is None:
            assert not inside.any()
            continue
        q = q[inside]
        assert np.all(q >= r[:3] - 1e-4) and np.all(q <= r[3:] + 1e-4)
        checked += len(q)
    assert checked > 1000


def test_hull_meets_matches_sampling():
    """hull_meets(p, l, m): some box (1 - s) p + s l, s in [0, 1], meets m."""
    rng = np.random.default_rng(12)
    agree = 0
    for _ in range(400):
        boxes = []
        for _k in range(3):
            c = rng.uniform(-3, 3, 3)
            e = rng.uniform(0.05, 1.0, 3)
            boxes.append(np.concatenate([c - e, c + e]).astype(np.float32))
        p, l, m = boxes
        s = np.linspace(0.0, 1.0, 4001)[:, None]
        lo = (1 - s) * p[:3] + s * l[:3]
        hi = (1 - s) * p[3:] + s * l[3:]
        meets = np.any(np.all((lo <= m[3:]) & (hi >= m[:3]), axis=1))
        got = O.hull_meets(p, l, m)
        if meets:
            assert got
        else:
            # a miss by less than the sampling step may still be a hit
            s2 = np.linspace(0.0, 1.0, 400001)[:, None]
            lo2, hi2 = (1 - s2) * p[:3] + s2 * l[:3], (1 - s2) * p[3:] + s2 * l[3:]
            near = np.any(np.all((lo2 <= m[3:] + 1e-5) & (hi2 >= m[:3] - 1e-5), axis=1))
            assert got == near or near
        agree += 1
    assert agree == 400


def _world(**cfg):
    tris9, counts, pl, surf, lights, movable = vcache_world()
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True, movable=movable)
    opts = O.opts_array(hierarchy=2, prefilter=1.0, object_cov=f.object_cov)
    kw = dict(pages=1024, append_pages=1024, candidates=1024, base_texel=0.25)
    kw.update(cfg)
    vc = O.OracleViewCache(f, surf, lights, opts, **kw)
    return tris9, counts, pl, surf, lights, movable, f, opts, vc, kw


def _keys(vc, pl, mips=(0, 1)):
    first_chart = [0, 1, 7, 8]  # ground 1 chart, box 6, boulder 1
    keys = []
    for m, p in enumerate(pl):
        g = int(p["geometry"])
        for c in range(first_chart[g], first_chart[g + 1]):
            for mip in mips:
                gx, gy = vc.grid(m, c, mip)
                keys += [(m, c | (mip << 24), x | (y << 16)) for y in range((gy + 7) // 8) for x in range((gx + 7) // 8)]
    return keys


def test_grid_and_flat_chart_places():
    tris9, counts, pl, surf, lights, movable, f, opts, vc, kw = _world()
    # ground: chart 0, 20 m square, base texel 0.25 -> ceil(20 / 0.25 / 2^mip)
    for mip in range(4):
        assert vc.grid(0, 0, mip) == (int(np.ceil(80 / 2 ** mip)),) * 2
    # a box face scales with its object (chart 1 of object 1)
    sx, sy = surf.chart_size[1]
    s = float(pl[1]["scale"])
    for mip in range(3):
        want = [max(1, int(np.ceil(np.float32(np.float32(z) * np.float32(s)) * np.float32(4.0 / 2 ** mip))))
                for z in (sx, sy)]
        assert vc.grid(1, 1, mip) == tuple(want)
    # texel centres of the ground quad: corner + u du + v dv, lifted by the offset
    a, du, dv = np.array([-10.0, -10.0, 0.0]), np.array([20.0, 0, 0]), np.array([0, 20.0, 0])
    for tx, ty in [(0, 0), (3, 70), (79, 79), (40, 12)]:
        apex, normal, ok, tri = vc.surface(0, 0, 0, tx, ty)
        u, v = (tx + 0.5) / 80, (ty + 0.5) / 80
        p = a + u * du + v * dv
        off = max(1e-4, 2e-6 * np.linalg.norm(p))
        assert ok and tri in (0, 1)
        assert np.allclose(apex, p + np.array([0, 0, off]), atol=2e-5), (apex, p)
        assert np.allclose(normal, [0, 0, 1], atol=1e-6)


def test_open_and_covered_texels_snap_and_values_are_queries():
    tris9, counts, pl, surf, lights, movable, f, opts, vc, kw = _world()
    keys = _keys(vc, pl, mips=(0,))
    slots = vc.request(keys)
    assert np.all(slots != NONE)
    vc.frame()
    tex, sl = vc.read(keys)
    assert np.all(sl != NONE)
    # a texel whose cone meets nothing keeps its light whole (ratio exactly 1), one
    # whose cone is covered keeps none (exactly 0), whatever the estimators' rounding
    rng = np.random.default_rng(2)
    empty = covered = 0
    for i in rng.choice(len(keys), 60, replace=False):
        k = keys[i]
        m, c, mip = k[0], k[1] & 0xFFFFFF, k[1] >> 24
        n = vc.grid(m, c, mip)
        for t in rng.choice(81, 8, replace=False):
            tx, ty = (k[2] & 0xFFFF) * 8 + t % 9, (k[2] >> 16) * 8 + t // 9
            if tx >= n[0] or ty >= n[1]:
                continue
            apex, normal, ok, tri = vc.surface(m, c, mip, tx, ty)
            for l in range(len(lights)):
                open_ = float(O.lib().oracle_unshadowed_rect(O._p(apex), O._p(normal), O._p(lights[l])))
                if not ok or open_ <= 0:
                    continue
                _, full, touched = O.query_skip(f, opts, apex, normal, lights[l])
                if touched == 0:
                    assert tex[i, t, l] == 1.0
                    empty += 1
                if full == (1 << 64) - 1:
                    assert tex[i, t, l] == 0.0
                    covered += 1
    assert empty > 50
    # with light_ratio off, a texel keeps its query's irradiance
    vc2 = O.OracleViewCache(f, surf, lights, opts, **dict(kw, light_ratio=0))
    some = keys[:: max(1, len(keys) // 12)]
    vc2.request(some)
    vc2.frame()
    tex2, _ = vc2.read(some)
    rng = np.random.default_rng(3)
    for i, k in enumerate(some):
        m, c, mip = k[0], k[1] & 0xFFFFFF, k[1] >> 24
        n = vc2.grid(m, c, mip)
        for t in rng.choice(81, 6, replace=False):
            tx, ty = (k[2] & 0xFFFF) * 8 + t % 9, (k[2] >> 16) * 8 + t // 9
            if tx >= n[0] or ty >= n[1]:
                continue
            apex, normal, ok, tri = vc2.surface(m, c, mip, tx, ty)
            for l in range(len(lights)):
                open_ = float(O.lib().oracle_unshadowed_rect(O._p(apex), O._p(normal), O._p(lights[l])))
                if ok and open_ > 0:
                    want = O.query_skip(f, opts, apex, normal, lights[l])[0]
                    assert tex2[i, t, l] == np.float32(want)
    # the sky channel is zero everywhere
    assert np.all(tex[..., -1] == 0.0)


def test_skipping_a_triangle_equals_removing_it():
    light = rect_light([-0.5, -0.5, 4.0], [1.0, 0, 0], [0, 1.0, 0])
    occ = quad_pair([-0.6, -0.6, 2.0], [1.2, 0, 0], [0, 1.2, 0])  # two triangles over the point
    floor = quad_pair([-3, -3, 0.0], [6, 0, 0], [0, 6, 0])
    both = O.OracleFlat(np.concatenate([floor, occ]).reshape(-1, 9), np.array([2, 2]))
    one = O.OracleFlat(np.concatenate([floor, occ[1:]]).reshape(-1, 9), np.array([2, 1]))
    opts = O.opts_array(hierarchy=2)
    rng = np.random.default_rng(5)
    for _ in range(20):
        p = np.array([rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), 1e-3], np.float32)
        n = np.array([0, 0, 1], np.float32)
        skipped = O.query_skip(both, opts, p, n, light, skip_obj=1, skip_tri=2)
        removed = O.query_skip(one, opts, p, n, light)
        whole = O.query_skip(both, opts, p, n, light)
        assert np.float32(skipped[0]) == np.float32(removed[0]) and skipped[1:] == removed[1:]
        assert whole[0] <= skipped[0]


def _moved_boxes(tris9, counts, pl, moved_pl, objects):
    off = np.concatenate([[0], np.cumsum(counts)]).astype(int)
    out = []
    for m in objects:
        boxes = []
        for p in (pl[m], moved_pl[m]):
            g = int(p["geometry"])
            v = tris9[off[g]:off[g + 1]].reshape(-1, 3).astype(np.float64)
            w = v @ p["rotation"].reshape(3, 3).astype(np.float64).T * p["scale"] + p["translation"]
            boxes += [(w.min(0) - 1e-3).astype(np.float32), (w.max(0) + 1e-3).astype(np.float32)]
        out.append((m, *boxes))
    return out


def _scene_box(flat, moves):
    lo, hi = flat.nodes[0][0:3].view(np.float32), flat.nodes[0][4:7].view(np.float32)
    lo = np.minimum(lo, np.min([m[1] for m in moves], axis=0))
    hi = np.maximum(hi, np.max([m[2] for m in moves], axis=0))
    pad = np.float32(1e-3) * max(np.float32(1.0), np.abs(np.concatenate([lo, hi])).max())
    return np.concatenate([lo - pad, hi + pad]).astype(np.float32)


@pytest.mark.parametrize("touch,visible_frames,append_pages", [(True, 1, 1024), (False, 0, 1024), (True, 1, 24)])
def test_moves_keep_every_shown_page_current(touch, visible_frames, append_pages):
    tris9, counts, pl, surf, lights, movable, f, opts, vc, kw = _world(visible_frames=visible_frames,
                                                                      append_pages=append_pages)
    keys = _keys(vc, pl)
    vc.request(keys)
    vc.frame()
    before, _ = vc.read(keys)
    moved_pl = pl.copy()
    moved_pl[movable[0]]["translation"] += np.array([1.5, 0.5, 0.0], np.float32)
    moved_pl[movable[1]]["translation"] += np.array([-0.7, 1.0, 0.0], np.float32)
    g = Flat.build_instanced(tris9, counts, moved_pl, object_prefilter=True, movable=movable)
    moves = _moved_boxes(tris9, counts, pl, moved_pl, movable)
    vc.set_scene(g)
    vc.set_moved(moves, _scene_box(g, moves))
    if touch:
        vc.request(keys)
    st = vc.frame()
    after, slots = vc.read(keys)
    fresh = O.OracleViewCache(g, surf, lights, opts, **kw)
    fresh.request(keys)
    fresh.frame()
    want, _ = fresh.read(keys)
    shown = slots != NONE
    assert shown.any()
    differ = ~np.all(after[shown].view(np.uint32) == want[shown].view(np.uint32), axis=(1, 2))
    assert not differ.any(), f"{differ.sum()} shown pages differ from a fresh bake"
    changed = ~np.all(before.view(np.uint32) == want.view(np.uint32), axis=(1, 2))
    assert changed.sum() > 0  # the moves matter
    if touch and append_pages >= len(keys):
        # every page in view stays, re-baked where the moves reach it
        assert shown.all() and st["stale_pages"] > 0 and st["evicted"] == 0
    else:
        # pages the moves reach out of view, or past the append capacity, are dropped
        assert st["evicted"] > 0 and (~shown).sum() == st["evicted"]
        assert changed[~shown].sum() > 0 or append_pages < len(keys)


# ---- shading from the cache and planning (WGSL :5907-6130; vcache.rs:3704)

def _view(W=64, H=48):
    from continuum_amd.scenes import camera_uniform
    tris9, counts, pl, surf, lights, movable = vcache_world()
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True, movable=movable)
    opts = O.opts_array(hierarchy=2, prefilter=1.0, object_cov=f.object_cov)
    cam = camera_uniform([-7.0, -6.0, 2.5], [0.0, 0.0, 0.3], 60.0, W, H)
    hits, recv, _ = O.primary(f, cam, W, H, lights, hierarchy=2)
    return f, surf, lights, opts, cam, hits, recv


def test_shading_falls_back_to_unshadowed_light_then_converges():
    f, surf, lights, opts, cam, hits, recv = _view()
    W, H = 64, 48
    vc = O.OracleViewCache(f, surf, lights, opts, pages=4096, append_pages=256, candidates=1024, base_texel=0.02)
    surface = hits[:, 0] != NONE
    out, page, req, cnt = vc.shade(cam, W, H, hits)
    # no page yet: every surface pixel takes the unshadowed light at its point
    # (checked where shading normals are the face's: the ground and the boxes,
    # objects 0-4)
    assert cnt[1] == surface.sum() and np.all((page.reshape(-1)[surface] >> 31) == 1)
    flat = surface & (hits[:, 0] <= 4)
    assert flat.sum() > 100
    for i in np.nonzero(flat)[0][::37]:
        pt, nm = recv[i, 0:3].copy(), recv[i, 4:7].copy()  # kept alive across the calls
        for l in range(len(lights)):
            open_ = float(O.lib().oracle_unshadowed_rect(O._p(pt), O._p(nm), O._p(lights[l])))
            assert abs(out[l].reshape(-1)[i] - open_) <= 1e-5 * max(1.0, abs(open_)) + 1e-6
    # requests ask for the coarsest missing page, once each
    assert len(req) == cnt[0] > 0 and len({tuple(r[:3]) for r in req}) == len(req)
    assert np.all(req[:, 3] == 1)
    last = np.full(W * H, 99)
    budget = 4000
    for frame in range(12):
        st = vc.plan(req, budget)
        assert st["cones"] <= budget or st["planned"] == 1
        vc.frame()
        out, page, req, cnt = vc.shade(cam, W, H, hits)
        shown = page.reshape(-1)
        found = surface & ((shown >> 31) == 0)
        # coarse to fine: a pixel's page once found never gets coarser
        mip = np.where(found, shown & 0xFF, 99)
        assert np.all(mip[found] <= last[found])
        last = np.where(found, mip, last)
    assert cnt[1] == 0                       # every pixel has a page
    assert len(req) <= 4                     # little left to refine


def test_plan_evicts_the_least_recently_used():
    f, surf, lights, opts, cam, hits, recv = _view()
    W, H = 64, 48
    vc = O.OracleViewCache(f, surf, lights, opts, pages=40, append_pages=16, candidates=64, base_texel=0.02)
    for frame in range(8):
        out, page, req, cnt = vc.shade(cam, W, H, hits)
        st = vc.plan(req, 100000)
        vc.frame()
    s = vc.state()
    resident = (s["slot_info"][:, 3] != 0).sum()
    assert resident <= 40
    # pages in use this view stay; evictions (when any) took pages unused for
    # more than two frames
    out, page, req, cnt = vc.shade(cam, W, H, hits)
    used = s["slot_used"]
    assert np.all(used[s["slot_info"][:, 3] != 0] > 0)
