"""The reference's `Renderer` (crates/continuum-gpu/src/renderer.rs:196-660)
over this library: an instanced scene with its lights, each frame a camera
pass and the view lit through mask cones, either from the view cache kept
across frames (`RendererOptions.cache`, the reference's default) or with one
cone per pixel and light. The camera moves freely; movable instances move
(`set_transform`, from the next frame) and the cache re-bakes what their
moves can reach; a light edit (`set_light`) re-bakes that light's channel;
re-colouring a light (`set_light_radiance`) costs nothing.

Colours are linear RGB: albedo (one per geometry here; textures are not part
of this library) times the sum over lights of radiance times irradiance per
unit radiance, over pi (`per_pixel_hit_color`, mask_cones.wgsl:2279-2292);
a pixel that sees a light's front shows its radiance (no spot falloff), a
ray that misses shows the sky colour. The shading arithmetic is a few
elementwise tensor operations on the device; every cone, page and hit comes
from libcnq."""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from . import _lib
from .kernel import Device, Flat, Scene, ViewCache


@dataclass
class RendererOptions:
    """The parts of `RendererOptions` (renderer.rs:40-190) this library has."""
    cache: bool = True                    # shade through the view cache (else a cone per pixel and light)
    movable: Sequence[int] = ()           # instances that may move
    moving_lights: Sequence[int] = ()     # lights that move through the cache's error ledger
    ledger_epsilon: float = 0.01          # the ledger's threshold (`CacheOptions::ledger_epsilon`)
    bake_budget: int = 2_000_000          # cones of page bakes a frame (`CacheOptions::bake_budget`)
    pages: int = 1 << 16                  # cache slots
    base_texel: float = 0.05              # metres per texel at mip 0
    albedo: Optional[Sequence[float]] = None  # per geometry (default 0.8)
    sky_color: Sequence[float] = (0.0, 0.0, 0.0)  # what rays that miss see
    pixels_per_texel: float = 2.0
    refine_levels: int = 3
    request_tile: int = 4
    extra: dict = field(default_factory=dict)


class Renderer:
    """`Renderer::new(kernel, format, (w, h), scene, origin, lights, options)`:
    here `(tris9, counts, placements)` is the instanced scene (geometries and
    placements, as `Flat.build_instanced` takes them), `surface` its charts
    (continuum_amd.surfaces.Surface over the geometries), `lights` LightGpu
    rows whose radiance (floats 12-14) colours them. The origin is the
    placements' own frame (positions are f32 here)."""

    def __init__(self, device: Device, size: tuple, tris9: np.ndarray, counts: Sequence[int], placements,
                 surface, lights: np.ndarray, options: Optional[RendererOptions] = None):
        import torch

        self.options = options or RendererOptions()
        self.width, self.height = int(size[0]), int(size[1])
        self.flat = Flat.build_instanced(tris9, counts, placements, object_prefilter=True,
                                         movable=list(self.options.movable) or None)
        self.scene = Scene(device, self.flat)
        self.lights = np.array(lights, np.float32, copy=True).reshape(-1, 16)
        self.opts = _lib.default_options(hierarchy=2, prefilter=1.0, object_cov=self.flat.object_cov)
        self.dev = torch.device("cuda:0")
        px = self.width * self.height
        self._hits = torch.zeros((px, 4), dtype=torch.int32, device=self.dev)
        self._recv = torch.zeros((px, 8), dtype=torch.float32, device=self.dev)
        self._light_rows = torch.from_numpy(self.lights).to(self.dev)
        n_geom = len(counts)
        albedo = np.full(n_geom, 0.8, np.float32) if self.options.albedo is None else np.asarray(
            self.options.albedo, np.float32)
        geometry = np.array([int(p["geometry"]) for p in placements], np.int64)
        self._albedo = torch.from_numpy(albedo[geometry]).to(self.dev)  # per object
        self._colors = torch.zeros((px, 3), dtype=torch.float32, device=self.dev)
        self.cache = None
        if self.options.cache:
            self.cache = ViewCache(self.scene, surface, self.lights, self.opts, pages=self.options.pages,
                                   append_pages=1024, candidates=8192, base_texel=self.options.base_texel,
                                   light_ratio=1, visible_frames=2)
            if self.options.moving_lights:
                self.cache.set_moving_lights(list(self.options.moving_lights), self.options.ledger_epsilon)
            self._out = torch.zeros((len(self.lights), px), dtype=torch.float32, device=self.dev)
            self._page = torch.zeros(px, dtype=torch.int32, device=self.dev)
        else:
            self._vis = torch.zeros(len(self.lights) * px, dtype=torch.float32, device=self.dev)
            self._full = torch.zeros(len(self.lights) * px, dtype=torch.int64, device=self.dev)
            self._touched = torch.zeros(len(self.lights) * px, dtype=torch.int64, device=self.dev)
        self._moved = False
        self._log: list = []

    # ---- scene changes (from the next frame)

    def set_transform(self, instance: int, rotation, position) -> None:
        """renderer.rs:442: instance `instance` (one of the movable) to
        `rotation` (3 x 3, orthonormal) and `position`; its scale stays."""
        self.scene.set_transform(instance, rotation, position)
        self._moved = True

    def set_light(self, index: int, light) -> None:
        """renderer.rs:421: light `index` moved, resized or re-aimed (and
        re-coloured). A moving light (`moving_lights`) goes through the
        cache's error ledger from the next frame; another has its channel of
        every resident page re-baked now."""
        import torch

        row = np.ascontiguousarray(light, np.float32).reshape(16)
        self.lights[index] = row
        self._light_rows[index] = torch.from_numpy(row.copy()).to(self.dev)
        if self.cache is not None:
            if index in self.options.moving_lights:
                self.cache.move_light(index, row)
            else:
                self.cache.set_lights(row[None], index)

    def verify_ledger(self, light: int) -> dict:
        """renderer.rs:640 `verify_ledger`: moving light `light`'s pages shown,
        re-baked afresh, against their certified bounds."""
        return self.cache.verify_ledger(light)

    def set_light_radiance(self, index: int, radiance) -> None:
        """renderer.rs:548: re-colour light `index`; nothing is baked."""
        import torch

        rgb = np.asarray(radiance, np.float32).reshape(3)
        self.lights[index, 12:15] = rgb
        self._light_rows[index, 12:15] = torch.from_numpy(rgb.copy()).to(self.dev)

    # ---- frames

    def render(self, camera) -> None:
        """renderer.rs:468 (`render`/`compute`): one frame seen from `camera`
        (the 20-float camera uniform). Returns without waiting for the GPU
        except where the cache's planning reads its requests."""
        import torch

        cam = np.ascontiguousarray(camera, np.float32).reshape(20)
        w, h = self.width, self.height
        entry = {}
        if self._moved:
            self.scene.apply_moves()
            if self.cache is not None:
                self.cache.set_moved(self.scene.last_moves())
            self._moved = False
        cam_t = torch.from_numpy(cam).to(self.dev)
        self.scene.primary_device(cam_t, w, h, self._light_rows, self.opts, self._hits, self._recv)
        if self.cache is not None:
            params = self.cache.shade_params(cam, self.options.pixels_per_texel, self.options.refine_levels,
                                             self.options.request_tile)
            requests, missing = self.cache.shade_device(cam, w, h, self._hits, self._out, self._page, params)
            plan = self.cache.plan(self.options.bake_budget)
            baked = self.cache.frame()
            entry.update(requests=requests, pixels_without_page=missing, planned_pages=plan["planned"],
                         planned_cones=plan["cones"], evicted=plan["evicted"] + baked["evicted"],
                         rebaked_cones=baked["appended_jobs"], stale_pages=baked["stale_pages"])
            irr = self._out
        else:
            self.scene.view_query_device(self._recv, w, h, self._light_rows, self.opts, self._vis, self._full,
                                         self._touched)
            irr = self._vis.view(len(self.lights), -1)
        self._shade(irr)
        entry["frame"] = len(self._log)
        self._log.append(entry)

    def _shade(self, irr) -> None:
        """Colours from irradiance per unit radiance (lights x pixels)."""
        import torch

        hits = self._hits
        obj = hits[:, 0]
        surface = obj != -1
        radiance = self._light_rows[:, 12:15]                 # (lights, 3)
        e = radiance.t() @ irr                                # (3, pixels)
        albedo = torch.where(surface, self._albedo[obj.clamp(min=0).long()], torch.zeros_like(e[0]))
        colors = (albedo * e / np.float32(np.pi)).t()
        sky = torch.tensor(list(self.options.sky_color), dtype=torch.float32, device=self.dev)
        colors = torch.where(surface[:, None], colors, sky[None, :])
        seen = hits[:, 3]
        if bool((seen > 0).any()):
            lit = radiance[(seen - 1).clamp(min=0).long()]
            colors = torch.where((seen > 0)[:, None], lit, colors)
        self._colors.copy_(colors)

    # ---- reading back and diagnostics

    def wait(self) -> None:
        """renderer.rs:600."""
        import torch

        torch.cuda.synchronize()

    def read_colors(self) -> np.ndarray:
        """renderer.rs:606: the last frame's linear RGB colours, (h, w, 3),
        row by row from the top."""
        self.wait()
        return self._colors.cpu().numpy().reshape(self.height, self.width, 3)

    def stats(self) -> Optional[dict]:
        """renderer.rs:553: the last frame's cache report (None without the
        cache)."""
        if self.cache is None or not self._log:
            return None
        return dict(self._log[-1])

    def is_settled(self, frames: int) -> bool:
        """renderer.rs:565: the cache asked for nothing and baked nothing in
        the last `frames` frames."""
        if self.cache is None or len(self._log) < frames:
            return False
        return all(e["requests"] == 0 and e["planned_pages"] == 0 and e["rebaked_cones"] == 0
                   for e in self._log[-frames:])

    def verify_cache(self) -> tuple:
        """renderer.rs:631: every page the cache may show re-baked and
        compared bit for bit: (values that differ, values compared)."""
        if self.cache is None:
            return 0, 0
        return self.cache.verify()

    def gpu_bytes(self) -> int:
        """renderer.rs:619: the scene's resident bytes plus the cache's pool
        (81 texels x channels x 4 B a slot)."""
        total = self.scene.hbm_bytes
        if self.cache is not None:
            total += self.options.pages * 81 * (len(self.lights) + 1) * 4
        return total
