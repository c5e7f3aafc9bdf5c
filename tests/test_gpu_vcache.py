"""GPU: the view cache's page passes (`vc_update`, `vc_bake`, `vc_fill`,
`vc_stale`, `vc_stale_texels`, `vc_find`) through the C ABI against the
oracle, bit for bit: the page tables (the host mirror and the device copy),
every slot's info, stale mark and last use, and every texel of the pool,
after baking, after rigid moves of the movable objects, and through a churn
of requests and evictions. Where the append capacity runs out, which pages
the GPU's atomics append first differs from the oracle's order, so there
the invariant is checked instead: every page still shown equals a fresh
bake of the moved scene."""
import numpy as np
import pytest

from continuum_amd import _lib
from continuum_amd.kernel import Flat, Scene, ViewCache
from continuum_amd.scenes import vcache_world
from oracle import oracle as O

pytestmark = pytest.mark.gpu

FIRST_CHART = [0, 1, 7, 8]  # vcache_world: ground 1 chart, box 6, boulder 1


def _keys(vc, pl, mips=(0, 1)):
    keys = []
    for m, p in enumerate(pl):
        g = int(p["geometry"])
        for c in range(FIRST_CHART[g], FIRST_CHART[g + 1]):
            for mip in mips:
                keys += vc.pages_of(m, c, mip)
    return keys


def _setup(device, **cfg):
    tris9, counts, pl, surf, lights, movable = vcache_world()
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True, movable=movable)
    scene = Scene(device, f)
    kw = dict(pages=1024, append_pages=1024, candidates=1024, base_texel=0.25, light_ratio=1, visible_frames=1)
    kw.update(cfg)
    opts = _lib.default_options(hierarchy=2, prefilter=1.0, object_cov=f.object_cov)
    vc = ViewCache(scene, surf, lights, opts, **kw)
    oopts = O.opts_array(hierarchy=2, prefilter=1.0, object_cov=f.object_cov)
    ov = O.OracleViewCache(f, surf, lights, oopts, **kw)
    return dict(tris9=tris9, counts=counts, pl=pl, surf=surf, lights=lights, movable=movable, f=f, scene=scene,
                vc=vc, ov=ov, kw=kw, opts=opts, oopts=oopts)


def _same_state(vc, ov, label=""):
    a, b = vc.state(), ov.state()
    assert a["slot_extent"] == b["slot_extent"] and a["frame"] == b["frame"], label
    assert np.array_equal(a["table_host"], a["table_device"]), f"{label}: device table differs from the host's"
    assert np.array_equal(a["table_host"], b["table"]), f"{label}: page table differs from the oracle's"
    n = a["slot_extent"]
    for name in ("slot_info", "slot_stale", "slot_used"):
        bad = np.nonzero(np.any((a[name][:n] != b[name][:n]).reshape(n, -1), axis=1))[0]
        assert len(bad) == 0, f"{label}: {name} differs at slots {bad[:8]}"
    pa, pb = a["pool"][:n].view(np.uint32), b["pool"][:n].view(np.uint32)
    bad = np.nonzero(np.any((pa != pb).reshape(n, -1), axis=1))[0]
    if len(bad):
        s = bad[0]
        t = np.nonzero(np.any(pa[s] != pb[s], axis=1))[0]
        raise AssertionError(f"{label}: pool differs in {len(bad)} slots; slot {s} texels {t[:8]}: "
                             f"{a['pool'][s, t[0]]} vs {b['pool'][s, t[0]]}")


@pytest.mark.parametrize("light_ratio", [1, 0])
def test_bake_matches_the_oracle(device, light_ratio):
    w = _setup(device, light_ratio=light_ratio)
    vc, ov = w["vc"], w["ov"]
    keys = _keys(vc, w["pl"])
    assert [vc.grid(m, c, 0) for m in range(3) for c in range(3)] == [ov.grid(m, c, 0) for m in range(3) for c in range(3)]
    print(f"bake: {len(keys)} pages, light_ratio {light_ratio}", flush=True)
    sa, sb = vc.request(keys), ov.request(keys)
    assert np.array_equal(sa, sb) and np.all(sa != O.NONE)
    st_a, st_b = vc.frame(), ov.frame()
    assert st_a == st_b
    _same_state(vc, ov, "after the first bake")
    ta, la = vc.read(keys)
    tb, lb = ov.read(keys)
    assert np.array_equal(la, lb) and np.array_equal(ta.view(np.uint32), tb.view(np.uint32))
    # a second frame with every page in use bakes nothing and changes nothing
    vc.request(keys), ov.request(keys)
    st_a, st_b = vc.frame(), ov.frame()
    assert st_a == st_b and st_a["planned_jobs"] == 0
    _same_state(vc, ov, "after an idle frame")


def _move(w, shifts):
    """Moves the movers on the product's scene; returns the moved placements,
    the moves as the scene reports them, and the scene box."""
    pl = w["pl"].copy()
    for m, d in zip(w["movable"], shifts):
        pl[m]["translation"] = pl[m]["translation"] + np.asarray(d, np.float32)
        w["scene"].set_transform(m, pl[m]["rotation"], pl[m]["translation"])
    w["scene"].apply_moves()
    moves = w["scene"].last_moves()
    assert sorted(mv.object for mv in moves) == sorted(w["movable"])
    # each box holds its object's placed vertices, before and after
    off = np.concatenate([[0], np.cumsum(w["counts"])]).astype(int)
    for mv in moves:
        for p, lo, hi in ((w["pl"][mv.object], mv.old_lo, mv.old_hi), (pl[mv.object], mv.new_lo, mv.new_hi)):
            g = int(p["geometry"])
            v = w["tris9"][off[g]:off[g + 1]].reshape(-1, 3).astype(np.float64)
            x = v @ p["rotation"].reshape(3, 3).astype(np.float64).T * p["scale"] + p["translation"]
            assert np.all(x >= np.array(lo) - 1e-5) and np.all(x <= np.array(hi) + 1e-5)
    g = Flat.build_instanced(w["tris9"], w["counts"], pl, object_prefilter=True, movable=w["movable"])
    lo, hi = g.nodes[0][0:3].view(np.float32), g.nodes[0][4:7].view(np.float32)
    lo = np.minimum(lo, np.min([np.array(mv.old_lo, np.float32) for mv in moves], axis=0))
    hi = np.maximum(hi, np.max([np.array(mv.old_hi, np.float32) for mv in moves], axis=0))
    pad = np.float32(1e-3) * max(np.float32(1.0), np.abs(np.concatenate([lo, hi])).max())
    box = np.concatenate([lo - pad, hi + pad]).astype(np.float32)
    w["pl"] = pl
    return g, moves, box


def _oracle_moves(moves):
    return [(mv.object, np.array(mv.old_lo, np.float32), np.array(mv.old_hi, np.float32),
             np.array(mv.new_lo, np.float32), np.array(mv.new_hi, np.float32)) for mv in moves]


def _fresh(device, w, g, keys):
    fresh = ViewCache(Scene(device, g), w["surf"], w["lights"], w["opts"], **w["kw"])
    fresh.request(keys)
    fresh.frame()
    return fresh.read(keys)[0]


@pytest.mark.parametrize("touch", [True, False])
def test_moves_match_the_oracle_and_a_fresh_bake(device, touch):
    w = _setup(device, visible_frames=1 if touch else 0)
    vc, ov = w["vc"], w["ov"]
    keys = _keys(vc, w["pl"])
    vc.request(keys), ov.request(keys)
    vc.frame(), ov.frame()
    for step, shifts in enumerate([([1.5, 0.5, 0.0], [-0.7, 1.0, 0.0]), ([0.0, -2.0, 0.0], [1.2, 0.3, 0.0])]):
        print(f"move {step} (pages in view: {touch})", flush=True)
        g, moves, box = _move(w, shifts)
        vc.set_moved(moves, box)
        ov.set_scene(g)
        ov.set_moved(_oracle_moves(moves), box)
        if touch:
            vc.request(keys), ov.request(keys)
        st_a, st_b = vc.frame(), ov.frame()
        assert st_a == st_b, (st_a, st_b)
        assert st_a["candidates"] > 0 and (st_a["stale_pages"] > 0 if touch else st_a["evicted"] > 0)
        _same_state(vc, ov, f"after move {step}")
        got, slots = vc.read(keys)
        want = _fresh(device, w, g, keys)
        shown = slots != O.NONE
        assert shown.any()
        assert np.array_equal(got[shown].view(np.uint32), want[shown].view(np.uint32))
        # pages dropped are requested again, as a renderer's next frame would
        if not touch:
            vc.request(keys), ov.request(keys)
            assert vc.frame() == ov.frame()
            _same_state(vc, ov, f"after re-baking move {step}'s dropped pages")


def test_append_capacity_keeps_shown_pages_current(device):
    w = _setup(device, append_pages=12)
    vc = w["vc"]
    keys = _keys(vc, w["pl"])
    vc.request(keys)
    vc.frame()
    g, moves, box = _move(w, ([1.5, 0.5, 0.0], [-0.7, 1.0, 0.0]))
    vc.set_moved(moves, box)
    vc.request(keys)
    st = vc.frame()
    print(f"capacity: {st}", flush=True)
    assert st["overflow"] > 0 and st["evicted"] > 0
    got, slots = vc.read(keys)
    shown = slots != O.NONE
    assert (~shown).sum() == st["evicted"]
    want = _fresh(device, w, g, keys)
    assert np.array_equal(got[shown].view(np.uint32), want[shown].view(np.uint32))


def test_request_and_evict_churn_matches_the_oracle(device):
    w = _setup(device, pages=160)
    vc, ov = w["vc"], w["ov"]
    keys = _keys(vc, w["pl"], mips=(0, 1, 2))
    rng = np.random.default_rng(21)
    resident = set()
    for frame in range(6):
        ask = [keys[i] for i in rng.choice(len(keys), 90, replace=False)]
        sa, sb = vc.request(ask), ov.request(ask)
        assert np.array_equal(sa, sb), f"frame {frame}: slots differ"
        resident |= {k for k, s in zip(ask, sa) if s != O.NONE}
        assert vc.frame() == ov.frame()
        _same_state(vc, ov, f"churn frame {frame}")
        drop = [k for k in sorted(resident) if rng.random() < 0.4]
        vc.evict(drop), ov.evict(drop)
        resident -= set(drop)
    ta, la = vc.read(keys)
    tb, lb = ov.read(keys)
    assert np.array_equal(la, lb) and np.array_equal(ta.view(np.uint32), tb.view(np.uint32))


def test_shading_and_planning_match_the_oracle(device):
    """Shading from the cache (WGSL :5937 vc_shade, the direct light's part)
    and planning (vcache.rs:3704 plan) over ten frames of one view: per
    pixel and light the shaded irradiance bit for bit, the page shown, the
    requests (as a set), the pages planned and every slot and texel after
    each frame's bake."""
    import torch

    from continuum_amd.scenes import camera_uniform

    w = _setup(device, pages=4096, append_pages=256, candidates=1024, base_texel=0.02)
    vc, ov = w["vc"], w["ov"]
    W, H = 64, 48
    cam = camera_uniform([-7.0, -6.0, 2.5], [0.0, 0.0, 0.3], 60.0, W, H)
    hits, _, _, _ = w["scene"].render_view(cam, W, H, w["lights"], w["opts"])
    oh, _, _ = O.primary(w["f"], cam, W, H, w["lights"], hierarchy=2)
    assert np.array_equal(hits, oh)
    hits_d = torch.from_numpy(np.ascontiguousarray(hits).view(np.int32)).to("cuda")
    out_d = torch.zeros((len(w["lights"]), H, W), dtype=torch.float32, device="cuda")
    page_d = torch.zeros((H, W), dtype=torch.int32, device="cuda")
    for frame in range(10):
        print(f"shade frame {frame}", flush=True)
        counts = vc.shade_device(cam, W, H, hits_d, out_d, page_d)
        torch.cuda.synchronize()
        oout, opage, oreq, ocnt = ov.shade(cam, W, H, hits)
        got = out_d.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), oout.view(np.uint32)), f"frame {frame}: shading differs"
        assert np.array_equal(page_d.cpu().numpy().view(np.uint32), opage)
        assert counts == (int(ocnt[0]), int(ocnt[1]))
        st_a, st_b = vc.plan(4000), ov.plan(oreq, 4000)
        assert st_a == st_b, (frame, st_a, st_b)
        assert vc.frame() == ov.frame()
        _same_state(vc, ov, f"shade frame {frame}")
    assert counts[1] == 0  # every pixel found a page


def test_verify_finds_every_shown_page_current(device):
    """`verify_cache` (instanced.rs:3721-3729): every page shown baked afresh
    equals the pool bit for bit, after the first bake and after moves (the
    reference's invariant); it refuses to run while planned pages wait."""
    w = _setup(device)
    vc = w["vc"]
    keys = _keys(vc, w["pl"])
    vc.request(keys)
    with pytest.raises(_lib.CnqError):
        vc.verify()
    vc.frame()
    shown = int((vc.read(keys)[1] != O.NONE).sum())
    differ, compared = vc.verify()
    print(f"verify after the bake: {differ} of {compared}", flush=True)
    assert differ == 0 and compared == shown * 81 * vc.channels
    for shifts in ([1.5, 0.5, 0.0], [-0.7, 1.0, 0.0]), ([0.0, -2.0, 0.0], [1.2, 0.3, 0.0]):
        g, moves, box = _move(w, shifts)
        vc.set_moved(moves, box)
        vc.request(keys)
        vc.frame()
        differ, compared = vc.verify()
        print(f"verify after a move: {differ} of {compared}", flush=True)
        assert differ == 0 and compared > 0


def test_light_edits_match_the_oracle_and_a_fresh_bake(device):
    """A light edit (`Renderer::set_light`): the rect light moved and the
    sphere light grown, each edit re-baking that light's channel of every
    resident page; the state equals the oracle's, every page equals a fresh
    bake under the edited lights, and `verify` finds nothing stale. A move
    of the movers afterwards still matches the oracle (the shadow regions
    follow the edited light's box)."""
    w = _setup(device)
    vc, ov = w["vc"], w["ov"]
    keys = _keys(vc, w["pl"])
    vc.request(keys), ov.request(keys)
    vc.frame(), ov.frame()
    lights = np.array(w["lights"], np.float32, copy=True)
    rect = lights[0].copy()
    rect[0:3] += np.float32([1.0, -0.5, 0.3])
    sphere = lights[1].copy()
    sphere[4] *= np.float32(1.5)
    for idx, row in ((0, rect), (1, sphere)):
        st = vc.set_lights(row[None], idx)
        ov.set_lights(row[None], idx)
        lights[idx] = row
        print(f"light {idx} edited: {st}", flush=True)
        assert st["pages"] > 0 and st["cones"] > 0
        _same_state(vc, ov, f"after editing light {idx}")
    differ, compared = vc.verify()
    assert differ == 0 and compared > 0
    fresh = ViewCache(w["scene"], w["surf"], lights, w["opts"], **w["kw"])
    fresh.request(keys)
    fresh.frame()
    got, slots = vc.read(keys)
    want, _ = fresh.read(keys)
    shown = slots != O.NONE
    assert shown.any() and np.array_equal(got[shown].view(np.uint32), want[shown].view(np.uint32))
    g, moves, box = _move(w, ([1.5, 0.5, 0.0], [-0.7, 1.0, 0.0]))
    vc.set_moved(moves, box)
    ov.set_scene(g)
    ov.set_moved(_oracle_moves(moves), box)
    vc.request(keys), ov.request(keys)
    assert vc.frame() == ov.frame()
    _same_state(vc, ov, "after a move under the edited lights")
