"""GPU: moving lights through the view cache's error ledger
(cnq_vcache_set_moving_lights / cnq_vcache_move_light; WGSL :4457-5164)
through the C ABI against the oracle, bit for bit after every frame: the
page tables, slot states and every texel of the pool, every word of every
ledger record, each record's certified bound, and the ledger's counters.
A rect sun drifts (its pages bake on their lattice) while a sphere light
moves (guard cones and slack); then the reference's own check, the
`verify_ledger` soundness test (gpu.rs:2084-2160), on the GPU's cache."""
import math

import numpy as np
import pytest

from continuum_amd import _lib
from continuum_amd.kernel import Flat, Scene, ViewCache
from continuum_amd.scenes import sphere_light, sun_light, vcache_world
from oracle import oracle as O
from tests.test_gpu_vcache import _keys, _same_state

pytestmark = pytest.mark.gpu


def _sun(speed, frame):
    a = speed * frame
    toward = (0.5 * math.cos(a) + 0.4 * math.sin(a), -0.4 * math.cos(a) + 0.5 * math.sin(a), 0.75)
    return sun_light([0, 0, 0], toward, 200.0, (2.4, 2.2, 1.9))


def _pair(device, speed, epsilon=0.01, **cfg):
    tris9, counts, pl, surf, lights, movable = vcache_world()
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True)
    scene = Scene(device, f)
    lights = np.stack([_sun(speed, 0), sphere_light([6.0, -4.0, 5.0], 0.5)])
    kw = dict(pages=2048, append_pages=2048, candidates=16, base_texel=0.25, light_ratio=1, visible_frames=1)
    kw.update(cfg)
    vc = ViewCache(scene, surf, lights, _lib.default_options(hierarchy=2, prefilter=1.0, object_cov=f.object_cov),
                   **kw)
    ov = O.OracleViewCache(f, surf, lights, O.opts_array(hierarchy=2, prefilter=1.0, object_cov=f.object_cov), **kw)
    vc.set_moving_lights([0, 1], epsilon)
    ov.set_moving([0, 1], epsilon)
    return vc, ov, pl, kw


def _same_ledger(vc, ov, pages, label):
    ra, ba = vc.ledger_state()
    rb, bb = ov.ledger_state(pages)
    bad = np.nonzero(np.any((ra != rb).reshape(pages, -1), axis=1))[0]
    if len(bad):
        s = bad[0]
        j, w = np.argwhere(ra[s] != rb[s])[0]
        raise AssertionError(f"{label}: ledger records differ in {len(bad)} slots; slot {s} light {j} word {w}: "
                             f"{ra[s, j, w]:#x} vs {rb[s, j, w]:#x}")
    assert np.array_equal(ba.view(np.uint32), bb.view(np.uint32)), f"{label}: certified bounds differ"


@pytest.mark.parametrize("speed", [0.002, 1.2e-6])
def test_moving_lights_match_the_oracle_frame_by_frame(device, speed):
    vc, ov, pl, kw = _pair(device, speed)
    keys = _keys(vc, pl, mips=(0,))
    print(f"ledger: {len(keys)} pages, sun speed {speed}", flush=True)
    for fr in range(3):
        assert np.array_equal(vc.request(keys), ov.request(keys))
        assert vc.frame() == ov.frame()
    _same_state(vc, ov, "after the first bakes")
    _same_ledger(vc, ov, kw["pages"], "after the first bakes")
    held = 0
    for fr in range(1, 13):
        sun = _sun(speed, fr)
        sph = sphere_light([6.0 + 0.02 * fr, -4.0, 5.0 - 0.01 * fr], 0.5)
        vc.move_light(0, sun), ov.move_light(0, sun)
        vc.move_light(1, sph), ov.move_light(1, sph)
        vc.request(keys), ov.request(keys)
        st_a, st_b = vc.frame(), ov.frame()
        la, lb = vc.ledger_stats(), ov.ledger_stats()
        assert la == lb, f"frame {fr}: {la} vs {lb}"
        assert st_a == st_b, f"frame {fr}"
        _same_state(vc, ov, f"frame {fr}")
        _same_ledger(vc, ov, kw["pages"], f"frame {fr}")
        held += la["held_lattice"]
        if fr % 4 == 0:
            for light in (0, 1):
                ca, cb = vc.verify_ledger(light), ov.verify_ledger(light)
                assert ca == cb, (ca, cb)
                assert ca["pages"] == len(keys) and ca["violations"] == 0 and ca["over_epsilon"] == 0, ca
    if speed < 1e-3:
        assert held > 0


def test_pages_out_of_view_are_set_aside_and_rebaked_when_wanted(device):
    vc, ov, pl, kw = _pair(device, 0.05)
    keys = _keys(vc, pl, mips=(0,))
    vc.request(keys), ov.request(keys)
    vc.frame(), ov.frame()
    vc.frame(), ov.frame()  # nothing read: every page falls out of view
    sun = _sun(0.05, 20)
    vc.move_light(0, sun), ov.move_light(0, sun)
    vc.frame(), ov.frame()
    la, lb = vc.ledger_stats(), ov.ledger_stats()
    assert la == lb and la["set_aside"] > 0
    _same_state(vc, ov, "set aside")
    vc.request(keys), ov.request(keys)
    vc.frame(), ov.frame()
    assert vc.ledger_stats() == ov.ledger_stats()
    _same_state(vc, ov, "wanted again")
    _same_ledger(vc, ov, kw["pages"], "wanted again")
    assert (vc.state()["slot_stale"] == 2).sum() == 0
    chk = vc.verify_ledger(0)
    assert chk["violations"] == 0 and chk["over_epsilon"] == 0, chk


def test_moving_light_arguments_are_checked(device):
    vc, ov, pl, kw = _pair(device, 0.002)
    with pytest.raises(RuntimeError):
        vc.set_moving_lights([0])  # once, before the first frame
    with pytest.raises(RuntimeError):
        vc.move_light(0, sphere_light([0.0, 0.0, 5.0], 0.5))  # a rect light stays a rect
    with pytest.raises(RuntimeError):
        vc.set_lights(_sun(0.002, 5)[None, :], first=0)  # a moving light moves by move_light


def test_pages_past_the_append_capacity_are_set_aside_and_the_rest_stay_sound(device):
    """With room for only a few re-baked pages a frame, which pages the
    GPU's atomics append first is not the oracle's order, so the invariant
    is checked instead (as for the movers' stale passes): pages that did not
    fit are set aside (kept, missing), and every page still shown is within
    its certified bound and epsilon of a fresh bake."""
    vc, ov, pl, kw = _pair(device, 0.05, append_pages=4)
    keys = _keys(vc, pl, mips=(0,))
    vc.request(keys)
    vc.frame()
    for fr in range(1, 4):
        vc.move_light(0, _sun(0.05, fr))
        vc.move_light(1, sphere_light([6.0 + 0.3 * fr, -4.0, 5.0], 0.5))
        vc.request(keys)
        vc.frame()
        st = vc.ledger_stats()
        assert st["pages"] <= 4 and st["set_aside"] > 0, st
        for light in (0, 1):
            chk = vc.verify_ledger(light)
            assert chk["violations"] == 0 and chk["over_epsilon"] == 0, (fr, light, chk)
    stale = vc.state()["slot_stale"]
    assert (stale == 2).sum() > 0 and (stale == 1).sum() == 0
