"""CPU: moving lights through the view cache's error ledger, as the oracle
restates them (crates/continuum-gpu/src/mask_cones.wgsl :4457-5164:
`guard_cone`, `below_horizon`, `ledger_texel`, `ledger_reset`,
`light_shift`, the rect lights' lattice `lattice_frame` … `lattice_grew`,
and the `vc_ledger` pass; vcache.rs:1399-1430 and :2255-2380 on the host).

Pins:
- the arcsine, cosine and tangent (Cephes methods) against libm;
- `light_shift` as a bound: no point of the light moves through a larger
  angle than it says, seen from sampled points of the box, and it is tight
  for a point light moved square to the line of sight;
- the guard cone's angular radius against its definition (a sphere's
  arcsine, a rectangle's largest corner angle);
- the lattice's footprint areas summing to the footprint's shoelace area
  in bins, and an unoccluded footprint's ratio exactly 1;
- the reference's own tests of the ledger (crates/continuum-gltf/tests/
  gpu.rs:2084-2160 `a_drifting_sun_stays_within_the_ledger` and
  `a_slow_sun_holds_its_shadows_edges_within_the_bound`), restated on a
  small world: every 10th frame of a sun drifting by 0.002 or 1.2e-6
  radians of azimuth a frame, no page the cache may show is further from a
  fresh bake than its certified bound, nor than ε; the slow sun holds pages
  on their lattice, and some of those change (the bound is exercised);
- `verify_ledger` itself sees a stale cache (a light moved with no frame
  run: every changed page is a violation)."""
import math

import numpy as np
import pytest

from continuum_amd.kernel import Flat
from continuum_amd.scenes import sphere_light, sun_light, vcache_world
from oracle import oracle as O


def _f32(x):
    return float(np.float32(x))


def test_elementary_functions_against_libm():
    lib = O.lib()
    xs = np.linspace(-1.0, 1.0, 20001, dtype=np.float32)
    worst = max(abs(lib.oracle_asin(float(x)) - math.asin(float(x))) for x in xs)
    assert worst < 4e-7, worst
    xs = np.linspace(-3.0, 3.0, 20001, dtype=np.float32)
    worst = max(abs(lib.oracle_cos(float(x)) - math.cos(float(x))) for x in xs)
    assert worst < 3e-7, worst
    xs = np.linspace(-1.45, 1.45, 20001, dtype=np.float32)
    rel = max(abs(lib.oracle_tan(float(x)) - math.tan(float(x))) / max(1.0, abs(math.tan(float(x)))) for x in xs)
    assert rel < 4e-7, rel
    assert lib.oracle_asin(1.0) == _f32(math.pi / 2) and lib.oracle_asin(0.0) == 0.0
    assert lib.oracle_cos(0.0) == 1.0 and lib.oracle_tan(0.0) == 0.0


def _light_points(light, k=5):
    c, u, v = light[0:3].astype(np.float64), light[4:7].astype(np.float64), light[8:11].astype(np.float64)
    s = np.linspace(0.0, 1.0, k)
    return np.array([c + a * u + b * v for a in s for b in s])


def test_light_shift_bounds_every_points_move():
    """Seen from any point q of the box, light point p moving to p' turns
    by the angle between p - q and p' - q; light_shift bounds it."""
    rng = np.random.default_rng(21)
    checked = 0
    for _ in range(60):
        corner = rng.uniform(-3, 3, 3) + np.array([0, 0, 6.0])
        eu, ev = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        a = np.array([*corner, 0, *eu, 0, *ev, 0, 1, 1, 1, 0], np.float32)
        b = a.copy()
        b[0:3] += rng.normal(0, 0.05, 3).astype(np.float32)
        b[4:7] += rng.normal(0, 0.02, 3).astype(np.float32)
        lo = rng.uniform(-4, 4, 3)
        box = np.concatenate([lo, lo + rng.uniform(0.1, 1.0, 3)]).astype(np.float32)
        bound = O.light_shift(a, b, box)
        pa, pb = _light_points(a), _light_points(b)
        qs = box[:3] + rng.random((40, 3)) * (box[3:] - box[:3])
        for q in qs:
            da, db = pa - q, pb - q
            cosang = np.sum(da * db, 1) / np.linalg.norm(da, axis=1) / np.linalg.norm(db, axis=1)
            ang = np.arccos(np.clip(cosang, -1, 1)).max()
            assert ang <= bound * (1 + 1e-4) + 1e-7, (ang, bound)
            checked += 1
    assert checked == 2400


def test_light_shift_of_a_point_moved_square_to_the_view():
    """A sphere light of radius 0 moved by m square to the line from a point
    box at distance d turns by 2 asin(m / 2d) at most (the padded light
    boxes bring the box a little nearer); none when it does not move."""
    a = sphere_light([0.0, 0.0, 10.0], 0.0)
    b = sphere_light([0.3, 0.0, 10.0], 0.0)
    box = np.zeros(6, np.float32)
    s = O.light_shift(a, b, box)
    exact = 2 * math.asin(0.3 / 20.0)
    assert exact < s < exact * 1.01
    assert O.light_shift(a, a, box) == 0.0
    far = sphere_light([0.0, 0.0, 0.1], 0.0)
    assert O.light_shift(far, sphere_light([3.0, 0.0, 0.1], 0.0), box) == 4.0  # past twice the distance: unbounded


def test_guard_radius_is_the_lights_angular_radius():
    apex = np.array([0.3, -0.2, 0.0], np.float32)
    sph = sphere_light([1.0, 2.0, 5.0], 0.4)
    d = np.linalg.norm(np.array([1.0, 2.0, 5.0]) - apex)
    assert abs(O.guard_alpha(apex, sph) - math.asin(0.4 / d)) < 2e-6
    sun = sun_light([0, 0, 0], (0.3, -0.2, 0.9), 50.0)
    c = sun[0:3] + 0.5 * (sun[4:7] + sun[8:11])
    f = (c - apex) / np.linalg.norm(c - apex)
    corners = _light_points(sun, 2)
    want = max(math.acos(float(np.dot(f, (p - apex) / np.linalg.norm(p - apex)))) for p in corners)
    # acos in f32 near 1: a cosine rounded by an ulp or two moves the angle
    # by about ulp / sin(angle)
    assert abs(O.guard_alpha(apex, sun) - want) < 4 * 2.0**-24 / math.sin(want) + 1e-6


def _shoelace(q):
    x, y = q[:, 0], q[:, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


def test_lattice_areas_sum_to_the_footprint():
    """The footprint of the light now, on the lattice of its pose then,
    spreads over bins whose areas add up to its area in bins (the corners
    projected on the lattice's plane, in f64)."""
    rng = np.random.default_rng(22)
    done = 0
    for _ in range(50):
        then = sun_light([0, 0, 0], (rng.uniform(-0.5, 0.5), rng.uniform(-0.5, 0.5), 0.8), 30.0)
        now = then.copy()
        now[0:3] += rng.normal(0, 0.02, 3).astype(np.float32)
        apex = rng.uniform(-2, 2, 3).astype(np.float32)
        apex[2] = 0.0
        frame, kappa, ratio, flags, areas = O.lattice(apex, (0, 0, 1), now, then)
        if frame[12] == 0.0 or areas.sum() == 0.0:
            continue
        fwd, right, up = frame[0:3].astype(np.float64), frame[3:6].astype(np.float64), frame[6:9].astype(np.float64)
        beta = frame[9:11].astype(np.float64)
        pts = []
        for p in _light_points(now, 2)[[0, 1, 3, 2]]:
            d = p - apex
            pts.append([np.dot(right, d) / np.dot(fwd, d) / beta[0], np.dot(up, d) / np.dot(fwd, d) / beta[1]])
        want = _shoelace(np.array(pts))
        assert abs(float(areas.astype(np.float64).sum()) - want) < 2e-4 * max(want, 1.0), (areas.sum(), want)
        assert np.all(areas >= 0) and np.all(areas <= 1)
        done += 1
    assert done >= 40


def _moving_world():
    tris9, counts, pl, surf, lights, movable = vcache_world()
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True)
    opts = O.opts_array(hierarchy=2, prefilter=1.0, object_cov=f.object_cov)
    return tris9, counts, pl, surf, f, opts


def test_an_unoccluded_lattice_sees_all_its_light():
    tris9, counts, pl, surf, f, opts = _moving_world()
    then = sun_light([0, 0, 0], (0.5, -0.4, 0.75), 200.0)
    now = then.copy()
    now[0:3] += np.float32(0.01)
    frame, kappa, ratio, flags, areas = O.lattice([0.0, 0.0, 30.0], (0, 0, 1), now, then, f, opts, on=True)
    assert ratio == 1.0 and flags == (True, False) and kappa < 8.0


def sun_at(speed, frame):
    """gpu.rs:2110-2118: the sun drifting by `speed` radians of azimuth a frame."""
    a = speed * frame
    toward = (0.5 * math.cos(a) + 0.4 * math.sin(a), -0.4 * math.cos(a) + 0.5 * math.sin(a), 0.75)
    return sun_light([0, 0, 0], toward, 200.0, (2.4, 2.2, 1.9))


def world_keys(vc, pl, mips=(0,)):
    first_chart = [0, 1, 7, 8]  # ground 1 chart, box 6, boulder 1
    keys = []
    for m, p in enumerate(pl):
        g = int(p["geometry"])
        for c in range(first_chart[g], first_chart[g + 1]):
            for mip in mips:
                gx, gy = vc.grid(m, c, mip)
                keys += [(m, c | (mip << 24), x | (y << 16)) for y in range((gy + 7) // 8) for x in range((gx + 7) // 8)]
    return keys


def drifting(speed, frames=60, warm=3, check_every=10):
    """The reference's `drifting_sun` on the oracle: (pages held, pages
    changed at the checks, ledger texels re-baked)."""
    tris9, counts, pl, surf, f, opts = _moving_world()
    lights = np.stack([sun_at(speed, 0), sphere_light([6.0, -4.0, 5.0], 0.5)])
    vc = O.OracleViewCache(f, surf, lights, opts, pages=2048, append_pages=2048, candidates=16, base_texel=0.25)
    vc.set_moving([0])
    keys = world_keys(vc, pl)
    for _ in range(warm):
        vc.request(keys)
        vc.frame()
    held = changed = rebaked = 0
    for fr in range(1, frames + 1):
        vc.move_light(0, sun_at(speed, fr))
        vc.request(keys)
        vc.frame()
        st = vc.ledger_stats()
        held += st["held_lattice"]
        rebaked += st["texels"]
        assert st["checked"] == len(keys) and st["set_aside"] == 0
        if fr % check_every == 0:
            chk = vc.verify_ledger(0)
            assert chk["pages"] == len(keys), chk
            assert chk["violations"] == 0, f"a change beyond its bound: {chk}"
            assert chk["over_epsilon"] == 0, f"a change beyond epsilon: {chk}"
            changed += chk["changed"]
    return held, changed, rebaked


def test_a_drifting_sun_stays_within_the_ledger():
    held, changed, rebaked = drifting(0.002)
    assert rebaked > 0


def test_a_slow_sun_holds_its_shadows_edges_within_the_bound():
    held, changed, rebaked = drifting(1.2e-6)
    assert held > 0, "no page held its mixed texels"
    assert changed > 0, "no held page changed: the bound was not exercised"


def test_verify_ledger_sees_a_stale_cache():
    """With the sun moved and no frame run since, the cache still holds the
    old pose's ratios while a fresh bake sees the new one: pages change, and
    every change is a violation (no bound was certified for it)."""
    tris9, counts, pl, surf, f, opts = _moving_world()
    lights = np.stack([sun_at(0.05, 0), sphere_light([6.0, -4.0, 5.0], 0.5)])
    vc = O.OracleViewCache(f, surf, lights, opts, pages=2048, append_pages=2048, candidates=16, base_texel=0.25)
    vc.set_moving([0, 1])
    keys = world_keys(vc, pl)
    vc.request(keys)
    vc.frame()
    vc.move_light(0, sun_at(0.05, 10))
    chk = vc.verify_ledger(0)
    assert chk["changed"] > 0 and chk["violations"] == chk["changed"], chk
    vc.request(keys)
    vc.frame()
    chk = vc.verify_ledger(0)
    assert chk["violations"] == 0 and chk["over_epsilon"] == 0, chk
    with pytest.raises(ValueError):
        vc.set_moving([0])  # only before the first frame


def test_pages_out_of_view_are_set_aside_then_rebaked_when_wanted():
    """A page not read in the frame before a move that needs a re-bake is set
    aside (kept, shown as missing); once requested again the next frame's
    ledger re-bakes it and shows it (WGSL :5114-5116, :5161-5163)."""
    tris9, counts, pl, surf, f, opts = _moving_world()
    lights = np.stack([sun_at(0.05, 0), sphere_light([6.0, -4.0, 5.0], 0.5)])
    vc = O.OracleViewCache(f, surf, lights, opts, pages=2048, append_pages=2048, candidates=16, base_texel=0.25)
    vc.set_moving([0])
    keys = world_keys(vc, pl)
    vc.request(keys)
    vc.frame()
    vc.frame()  # nothing read: every page falls out of view
    vc.move_light(0, sun_at(0.05, 20))
    vc.frame()
    st = vc.ledger_stats()
    assert st["set_aside"] > 0 and st["checked"] == 0
    stale = vc.state()["slot_stale"]
    assert (stale == 2).sum() == st["set_aside"]
    vc.request(keys)
    vc.frame()
    st = vc.ledger_stats()
    assert st["pages"] > 0 and (vc.state()["slot_stale"] == 2).sum() == 0
    chk = vc.verify_ledger(0)
    assert chk["violations"] == 0 and chk["over_epsilon"] == 0, chk
