"""Seeded multi-view workloads over the courtyard: the batched form's 64
views of uneven resolution and light count, and the request stream the
server is measured on."""
from __future__ import annotations

import math

import numpy as np

from .scenes import Pcg32, camera_uniform, rect_light

RESOLUTIONS = [(256, 256), (640, 360), (1280, 720), (1920, 1080)]


def random_rect_light(rng: Pcg32, area_lo: float = 0.25, area_hi: float = 8.0) -> np.ndarray:
    """A one-sided square light facing down somewhere in the courtyard's
    volume, 0.25-8 m^2, random colour."""
    area = area_lo + (area_hi - area_lo) * rng.next_f32()
    s = math.sqrt(area)
    cx = 2.0 + 26.0 * rng.next_f32()
    cy = 3.5 + 8.0 * rng.next_f32()
    z = 6.0 + 12.0 * rng.next_f32()
    col = tuple(2.0 + 8.0 * rng.next_f32() for _ in range(3))
    return rect_light([cx - s / 2, cy + s / 2, z], [s, 0, 0], [0, -s, 0], one_sided=True, radiance=col)


def random_pose(rng: Pcg32):
    """An eye inside the courtyard looking at a point across it."""
    eye = (2.0 + 26.0 * rng.next_f32(), 4.0 + 7.0 * rng.next_f32(), 1.5 + 14.0 * rng.next_f32())
    tgt = (2.0 + 26.0 * rng.next_f32(), 0.5 + 14.0 * rng.next_f32(), 0.5 + 12.0 * rng.next_f32())
    if math.dist(eye, tgt) < 3.0:
        tgt = (tgt[0], 15.0 - tgt[1], tgt[2])
    return eye, tgt


def courtyard_views(n_views: int = 64, seed: int = 0xC0_0004, max_lights: int = 16, resolutions=RESOLUTIONS):
    """(views, lights): views as (camera20, width, height, light_first,
    light_count), lights as LightGpu rows; each view owns 1-max_lights of
    them."""
    rng = Pcg32(seed, 17)
    views, lights = [], []
    for _ in range(n_views):
        w, h = resolutions[rng.next_u32() % len(resolutions)]
        nl = 1 + rng.next_u32() % max_lights
        eye, tgt = random_pose(rng)
        first = len(lights)
        for _ in range(nl):
            lights.append(random_rect_light(rng))
        views.append((camera_uniform(eye, tgt, 60.0, w, h), w, h, first, nl))
    return views, np.stack(lights)


def view_queries(views) -> int:
    return int(sum(w * h * nl for _, w, h, _, nl in views))
