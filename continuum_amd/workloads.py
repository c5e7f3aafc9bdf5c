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


def serving_requests(n: int, rate: float, lights: np.ndarray, seed: int = 0xC0_0007, view_share: float = 0.8,
                     max_view_lights: int = 32, resolutions=RESOLUTIONS, extent: float = 900.0):
    """A seeded request stream over the boulder field: Poisson arrivals at
    `rate` per second; a `view_share` of them views (a resolution from
    `resolutions`, an eye 1.5-20 m up anywhere in the field looking 20-200 m
    away, 1-`max_view_lights` distinct lights of the table), the rest light
    edits (move, resize or recolour one rectangle of the table). Returns
    [(arrival_s, "view", (camera20, w, h, light indices))
     | (arrival_s, "edit", (index, LightGpu))]."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n))
    table = np.array(lights, np.float32, copy=True)
    n_lights = len(table)
    out = []
    for k in range(n):
        if rng.uniform() < view_share:
            w, h = resolutions[int(rng.integers(len(resolutions)))]
            eye = np.array([rng.uniform(-extent, extent), rng.uniform(-extent, extent), rng.uniform(1.5, 20.0)])
            ang = rng.uniform(0, 2 * np.pi)
            dist = rng.uniform(20.0, 200.0)
            tgt = eye + np.array([np.cos(ang) * dist, np.sin(ang) * dist, -eye[2] + rng.uniform(0.0, 2.0)])
            nl = int(rng.integers(1, min(max_view_lights, n_lights) + 1))
            idx = sorted(rng.choice(n_lights, size=nl, replace=False).tolist())
            out.append((float(t[k]), "view", (camera_uniform(eye, tgt, 60.0, w, h), w, h, idx)))
        else:
            i = int(rng.integers(1, n_lights))
            row = table[i].copy()
            what = int(rng.integers(3))
            if what == 0:
                row[0:2] += rng.uniform(-5.0, 5.0, 2).astype(np.float32)
            elif what == 1:
                s = np.float32(rng.uniform(0.7, 1.4))
                row[4:7] *= s
                row[8:11] *= s
            else:
                row[12:15] = rng.uniform(2.0, 10.0, 3).astype(np.float32)
            table[i] = row
            out.append((float(t[k]), "edit", (i, row)))
    return out
