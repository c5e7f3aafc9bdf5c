"""The oracle's covering weights against their closed forms
(crates/continuum-gpu/src/mask_cones.wgsl :578-616 and :730-742): one sheet
of an instanced world between a receiver and a rect light, covering the
whole cone window, placed opaque (clear 0), alpha-blended (clear > 0),
transmissive (clear < 0: thin glass passes the factor times one less
Schlick's reflectance, F0 0.04, at the cone's centre direction) or a
transmissive volume (clear <= -2: two surfaces). Every bin is covered by
the sheet's weight alone, so each bin's coverage is that weight, and a
tilted sheet reflects more (passes less) than one facing the cone."""
import math

import numpy as np
import pytest

from continuum_amd.kernel import PLACEMENT_DTYPE
from continuum_amd.scenes import quad_pair, rect_light
from oracle import oracle as O

LIGHT = rect_light([-0.25, 0.25, 3.0], [0.5, 0, 0], [0, -0.5, 0], one_sided=True)
RECV = np.array([[0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.0, 0.0]], np.float32)


def _sheet_world(clear, tilt_deg=0.0):
    # 8 x 8 m, facing down (toward the receiver), 1.5 m up, tilted about x
    sheet = quad_pair([-4.0, -4.0, 0.0], [0.0, 8.0, 0.0], [8.0, 0.0, 0.0])
    tris9 = sheet.reshape(-1, 9)
    t = math.radians(tilt_deg)
    rot = np.array([[1, 0, 0], [0, math.cos(t), -math.sin(t)], [0, math.sin(t), math.cos(t)]], np.float32)
    pl = np.zeros(1, dtype=PLACEMENT_DTYPE)
    pl[0]["geometry"] = 0
    pl[0]["rotation"] = rot.reshape(9)
    pl[0]["scale"] = 1.0
    pl[0]["translation"] = [0.0, 0.0, 1.5]
    pl[0]["clear"] = clear
    return O.OracleInstancedFlat(tris9, np.array([2], np.uint32), pl, object_prefilter=False)


def _cov(clear, tilt_deg=0.0):
    f = _sheet_world(clear, tilt_deg)
    vis, full, touched, cov, capped, _ = O.irradiance(f, RECV, LIGHT[None], O.opts_array(hierarchy=2), want_cov=True)
    assert capped.sum() == 0
    return float(vis[0]), cov[0], int(touched[0])


def _unshadowed():
    """The light through a fully clear sheet (clear 1: weight 0, it covers nothing)."""
    vis, cov, touched = _cov(1.0)
    assert touched == 0 and not cov.any()
    return vis


@pytest.mark.parametrize("clear,weight", [
    (0.0, 1.0),                               # opaque
    (0.3, 0.7),                               # blended: its opacity
    (-0.5, 1.0 - 0.5 * 0.96),                 # thin glass at normal incidence: F = 0.04
    (-1.0, 1.0 - 0.96),
    (-2.5, 1.0 - 0.5 * 0.96 * 0.96),          # a volume: two surfaces
])
def test_a_sheet_covers_every_bin_by_its_weight(clear, weight):
    vis, cov, touched = _cov(clear)
    assert touched == (1 << 64) - 1
    assert np.allclose(cov, weight, rtol=0, atol=2e-6), (clear, float(cov.min()), float(cov.max()))
    # what passes is the unshadowed light times one less the weight
    assert abs(vis - _unshadowed() * (1.0 - weight)) <= 1e-5 * _unshadowed() + 1e-7


def test_tilted_glass_reflects_more():
    """At 60 degrees from the cone's axis, Schlick's reflectance is
    0.04 + 0.96 (1 - cos 60)^5 = 0.07: the sheet passes less and covers more
    than facing the cone (the window's centre direction is within a degree
    of the light's centre, so cos_t is cos 60 to about 1e-2)."""
    _, cov0, _ = _cov(-1.0, 0.0)
    _, cov60, _ = _cov(-1.0, 60.0)
    f60 = 0.04 + 0.96 * (1.0 - math.cos(math.radians(60.0))) ** 5
    assert cov60.min() > cov0.max()
    assert np.allclose(cov60, 1.0 - (1.0 - f60), atol=3e-3)
