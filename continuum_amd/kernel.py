"""Host-side mirror of the reference's `MaskConeKernel`
(crates/continuum-gpu/src/lib.rs:65) over the C ABI of libcnq.so.

Same method names and argument meanings as the reference:
``irradiance`` (lib.rs:1706), ``irradiance_with`` (:1718), ``coverage``
(:3289), the setters ``set_prefilter``/``set_wide_cones``/
``set_nearest_first``/``set_profile_stage`` (:1501-1675), and the per-view
launch ``render_view`` (the cone part of ``render_instanced_lights``,
:1857). Receivers are (point, normal) pairs; meshes are lists of triangles.
Timings are the fastest of ``runs`` executions, submit to completion,
excluding uploads and readback, as the reference's ``run`` times them
(lib.rs:3497-3512).

Failures raise ``CnqError``; contract violations (bad shapes, too many
lights) raise ``ValueError``, where the reference panics.
"""
from __future__ import annotations

import ctypes as C
import time
from enum import IntEnum
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import NONE, CnqError, SceneDesc, check, default_options, lib
from .scenes import meshes_to_arrays, rect_light

MAX_LIGHTS = 16  # lib.rs:751; the device path has no such limit, the mirror keeps it


class Hierarchy(IntEnum):
    """hierarchy.rs:727-741"""

    NONE = 0
    MESHES = 1
    TRIANGLES = 2


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data if a.size else 0


# `cnq_placement` as a numpy record (60 bytes).
PLACEMENT_DTYPE = np.dtype([("geometry", "<u4"), ("rotation", "<f4", (9,)), ("scale", "<f4"),
                            ("translation", "<f4", (3,)), ("clear", "<f4")])


class Flat:
    """The reference's flattened scene (hierarchy.rs `flatten` + `mesh_hierarchy`),
    built by the library's host builder."""

    def __init__(self, tris, ranges, blas, blas_items, blas_cov, nodes, items, xforms=None, object_cov=NONE):
        self.tris = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 12)
        self.ranges = np.ascontiguousarray(ranges, dtype=np.uint32).reshape(-1, 4)
        self.blas = np.ascontiguousarray(blas, dtype=np.uint32).reshape(-1, 12)
        self.blas_items = np.ascontiguousarray(blas_items, dtype=np.uint32).reshape(-1)
        self.blas_cov = np.ascontiguousarray(blas_cov, dtype=np.float32).reshape(-1, 27)
        self.nodes = np.ascontiguousarray(nodes, dtype=np.uint32).reshape(-1, 12)
        self.items = np.ascontiguousarray(items, dtype=np.uint32).reshape(-1)
        self.xforms = (np.zeros((0, 16), np.float32) if xforms is None else
                       np.ascontiguousarray(xforms, dtype=np.float32).reshape(-1, 16))
        # First row of the object hierarchy's coverage in blas_cov (NONE: none).
        self.object_cov = int(object_cov)
        self.movers = None  # the movers' subtree layout, for instanced worlds with movers

    @classmethod
    def build(cls, meshes) -> "Flat":
        tris9, counts = meshes_to_arrays(meshes)
        return cls.build_arrays(tris9, counts)

    @classmethod
    def build_arrays(cls, tris9: np.ndarray, counts: np.ndarray) -> "Flat":
        L = lib()
        tris9 = np.ascontiguousarray(tris9, dtype=np.float32)
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        h = C.c_void_p()
        check(L.cnq_flat_build(_ptr(tris9), _ptr(counts), len(counts), C.byref(h)))
        try:
            sizes = np.zeros(6, dtype=np.uint64)
            check(L.cnq_flat_sizes(h, _ptr(sizes)))
            nt, nr, nb, nbi, nn, ni = (int(x) for x in sizes)
            out = dict(tris=np.zeros((nt, 12), np.float32), ranges=np.zeros((nr, 4), np.uint32),
                       blas=np.zeros((nb, 12), np.uint32), blas_items=np.zeros(nbi, np.uint32),
                       blas_cov=np.zeros((nb, 27), np.float32), nodes=np.zeros((nn, 12), np.uint32),
                       items=np.zeros(ni, np.uint32))
            check(L.cnq_flat_copy(h, *(_ptr(out[k]) for k in
                                       ("tris", "ranges", "blas", "blas_items", "blas_cov", "nodes", "items"))))
        finally:
            L.cnq_flat_destroy(h)
        return cls(**out)

    @classmethod
    def build_instanced(cls, tris9: np.ndarray, counts, placements, object_prefilter: bool = True,
                        movable: Optional[Sequence[int]] = None) -> "Flat":
        """The instanced world (`flatten_instanced`, hierarchy.rs:464): geometries
        as 9 floats per triangle, `counts[g]` each; placements as
        (geometry, rotation 3x3 row-major, scale, translation, clear) tuples
        or a `Placement` array; `movable` objects get their own subtree
        (`object_hierarchy_with_movers`, hierarchy.rs:331)."""
        L = lib()
        tris9 = np.ascontiguousarray(tris9, dtype=np.float32)
        counts = np.ascontiguousarray(counts, dtype=np.uint32)
        if isinstance(placements, np.ndarray):
            arr = np.ascontiguousarray(placements, dtype=PLACEMENT_DTYPE)
            ptr, n = _ptr(arr), len(arr)
        elif isinstance(placements, C.Array):
            arr = placements
            ptr, n = C.cast(arr, C.c_void_p), len(arr)
        else:
            arr = (_lib.Placement * len(placements))()
            for k, (g, rot, scale, t, clear) in enumerate(placements):
                arr[k].geometry = int(g)
                arr[k].rotation[:] = [float(x) for x in np.asarray(rot, np.float32).reshape(9)]
                arr[k].scale = float(scale)
                arr[k].translation[:] = [float(x) for x in np.asarray(t, np.float32).reshape(3)]
                arr[k].clear = float(clear)
            ptr, n = C.cast(arr, C.c_void_p), len(arr)
        h = C.c_void_p()
        mv = np.ascontiguousarray(list(movable or []), dtype=np.uint32)
        check(L.cnq_flat_build_instanced_movers(_ptr(tris9), _ptr(counts), len(counts), ptr, n, _ptr(mv), len(mv),
                                                int(object_prefilter), C.byref(h)))
        tree = None
        try:
            sizes = np.zeros(6, dtype=np.uint64)
            check(L.cnq_flat_sizes(h, _ptr(sizes)))
            nt, nr, nb, nbi, nn, ni = (int(x) for x in sizes)
            extra = np.zeros(3, dtype=np.uint64)
            check(L.cnq_flat_instanced_sizes(h, _ptr(extra)))
            nx, ncov, ocov = (int(x) for x in extra)
            out = dict(tris=np.zeros((nt, 12), np.float32), ranges=np.zeros((nr, 4), np.uint32),
                       blas=np.zeros((nb, 12), np.uint32), blas_items=np.zeros(nbi, np.uint32),
                       nodes=np.zeros((nn, 12), np.uint32), items=np.zeros(ni, np.uint32))
            check(L.cnq_flat_copy(h, _ptr(out["tris"]), _ptr(out["ranges"]), _ptr(out["blas"]),
                                  _ptr(out["blas_items"]), 0, _ptr(out["nodes"]), _ptr(out["items"])))
            xforms = np.zeros((nx, 16), np.float32)
            cov = np.zeros((ncov, 27), np.float32)
            check(L.cnq_flat_copy_instanced(h, _ptr(xforms), _ptr(cov)))
            if len(mv):
                t = _lib.MoverTree()
                check(L.cnq_flat_mover_tree(h, C.byref(t)))
                objs = np.ctypeslib.as_array((C.c_uint32 * t.n_objects).from_address(t.objects)).copy()
                tree = dict(fixed_root=np.array(t.fixed_root, np.uint32), first_node=t.first_node, nodes=t.nodes,
                            first_item=t.first_item, objects=objs)
        finally:
            L.cnq_flat_destroy(h)
        f = cls(blas_cov=cov, xforms=xforms, object_cov=ocov, **out)
        f.movers = tree
        if tree is not None:
            pl = placements if isinstance(placements, np.ndarray) else None
            f.mover_placements = None if pl is None else np.ascontiguousarray(pl[tree["objects"]], PLACEMENT_DTYPE)
            f.geometry_bounds = np.stack([np.concatenate([t.reshape(-1, 3).min(0), t.reshape(-1, 3).max(0)])
                                          for t in np.split(tris9, np.cumsum(counts)[:-1])]).astype(np.float32)
        return f

    def desc(self) -> SceneDesc:
        d = SceneDesc()
        d.tris, d.n_tris = _ptr(self.tris), len(self.tris)
        d.ranges, d.n_objects = _ptr(self.ranges), len(self.ranges)
        d.blas, d.n_blas = _ptr(self.blas), len(self.blas)
        d.blas_items, d.n_blas_items = _ptr(self.blas_items), len(self.blas_items)
        d.blas_cov = _ptr(self.blas_cov)
        d.nodes, d.n_nodes = _ptr(self.nodes), len(self.nodes)
        d.items, d.n_items = _ptr(self.items), len(self.items)
        d.xforms, d.n_xforms = _ptr(self.xforms), len(self.xforms)
        d.n_blas_cov = len(self.blas_cov)
        return d


class Device:
    def __init__(self, ordinal: int = 0):
        h = C.c_void_p()
        check(lib().cnq_device_create(ordinal, C.byref(h)))
        self.handle = h
        self.ordinal = ordinal
        # Bound now: module globals may already be cleared when a destructor
        # runs at interpreter exit.
        self._destroy = lib().cnq_device_destroy

    def __del__(self):
        if getattr(self, "handle", None):
            self._destroy(self.handle)
            self.handle = None


class Scene:
    """A flattened scene resident in HBM."""

    def __init__(self, device: Device, flat: Flat):
        self.device = device
        self.flat = flat
        h = C.c_void_p()
        d = flat.desc()
        check(lib().cnq_scene_create(device.handle, C.byref(d), C.byref(h)))
        self.handle = h
        self._destroy = lib().cnq_scene_destroy
        if getattr(flat, "movers", None) is not None and getattr(flat, "mover_placements", None) is not None:
            t = flat.movers
            mt = _lib.MoverTree()
            mt.fixed_root[:] = [int(x) for x in t["fixed_root"]]
            mt.first_node, mt.nodes, mt.first_item = t["first_node"], t["nodes"], t["first_item"]
            self._mover_objects = np.ascontiguousarray(t["objects"], np.uint32)
            mt.n_objects, mt.objects = len(self._mover_objects), _ptr(self._mover_objects)
            gb = np.ascontiguousarray(flat.geometry_bounds, np.float32)
            check(lib().cnq_scene_set_movers(self.handle, C.byref(mt), _ptr(gb), len(gb),
                                             _ptr(flat.mover_placements)))

    def set_transform(self, obj: int, rotation, translation) -> None:
        """Renderer::set_transform (renderer.rs:442) for a movable object."""
        r = np.ascontiguousarray(rotation, np.float32).reshape(9)
        t = np.ascontiguousarray(translation, np.float32).reshape(3)
        check(lib().cnq_scene_set_transform(self.handle, int(obj), _ptr(r), _ptr(t)))

    def apply_moves(self, stream=None) -> None:
        """Write moved objects' transforms and rebuild the movers' subtree
        (instanced.rs:1226-1410)."""
        check(lib().cnq_scene_apply_moves(self.handle, _stream_ptr(stream)))

    def set_pose(self, geometry: int, tris, stream=None) -> None:
        """InstancedRenderer::apply_poses (instanced.rs:1154) for one
        deforming geometry: new triangles (its count, its order), its
        hierarchy refit, its instances (all movable) marked moved; call
        apply_moves next."""
        t = np.ascontiguousarray(tris, np.float32).reshape(-1, 9)
        check(lib().cnq_scene_set_pose(self.handle, int(geometry), _ptr(t), len(t), _stream_ptr(stream)))

    def last_moves(self) -> list:
        """The last apply_moves's moves: (object, old box (lo, hi), new box),
        as `_lib.MovedBox` records for ViewCache.set_moved."""
        cap = len(getattr(self, "_mover_objects", ())) or 1
        out = (_lib.MovedBox * cap)()
        n = C.c_uint32()
        check(lib().cnq_scene_last_moves(self.handle, out, cap, C.byref(n)))
        return list(out[: n.value])

    def __del__(self):
        if getattr(self, "handle", None):
            self._destroy(self.handle)
            self.handle = None

    @property
    def hbm_bytes(self) -> int:
        return int(lib().cnq_scene_bytes(self.handle))

    def query(self, receivers: np.ndarray, lights: np.ndarray, opts: _lib.Options, want_cov: bool = False):
        """Host buffers in and out: (visible, full, touched, cov) per query
        (query l * n + i)."""
        receivers = np.ascontiguousarray(receivers, dtype=np.float32).reshape(-1, 8)
        lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 16)
        q = len(receivers) * len(lights)
        vis = np.zeros(q, np.float32)
        full = np.zeros(q, np.uint64)
        touched = np.zeros(q, np.uint64)
        cov = np.zeros((q, 64), np.float32) if want_cov else None
        check(lib().cnq_irradiance(self.handle, _ptr(receivers), len(receivers), _ptr(lights), len(lights),
                                   C.byref(opts), _ptr(vis), _ptr(full), _ptr(touched),
                                   _ptr(cov) if want_cov else 0))
        return vis, full, touched, cov

    def query_device(self, receivers, lights, opts, vis, full, touched, cov=None, stream=None) -> None:
        """Device tensors (torch) in and out; enqueued on `stream` (a
        torch.cuda.Stream, or the current stream)."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream()
        n = receivers.numel() // 8
        nl = lights.numel() // 16
        for t, m in ((vis, 1), (full, 1), (touched, 1)):
            if t.numel() < n * nl * m:
                raise ValueError("output tensor too small")
        check(lib().cnq_irradiance_device(self.handle, receivers.data_ptr(), n, lights.data_ptr(), nl,
                                          C.byref(opts), vis.data_ptr(), full.data_ptr(), touched.data_ptr(),
                                          cov.data_ptr() if cov is not None else 0, stream.cuda_stream))

    def sky(self, receivers: np.ndarray, opts: _lib.Options, sky: Optional[_lib.SkyOptions] = None):
        """The sky's irradiance per unit radiance at each receiver through six
        cube-face cones (WGSL :1411 sky_irradiance), host buffers: (total
        (n,), per face (6, n), full bins per face (6, n))."""
        receivers = np.ascontiguousarray(receivers, dtype=np.float32).reshape(-1, 8)
        n = len(receivers)
        out = np.zeros(n, np.float32)
        faces = np.zeros((6, n), np.float32)
        full = np.zeros((6, n), np.uint64)
        so = sky if sky is not None else _lib.default_sky_options()
        check(lib().cnq_sky_irradiance(self.handle, _ptr(receivers), n, C.byref(opts), C.byref(so), _ptr(out),
                                       _ptr(faces), _ptr(full)))
        return out, faces, full

    def sky_device(self, receivers, opts, out, faces=None, full=None, sky=None, stream=None) -> None:
        """The same on device tensors, enqueued on `stream` (the current
        stream by default): out (n,), faces (6 * n) and full (6 * n) optional."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream()
        n = receivers.numel() // 8
        if out.numel() < n or (faces is not None and faces.numel() < 6 * n) or (
                full is not None and full.numel() < 6 * n):
            raise ValueError("output tensor too small")
        so = sky if sky is not None else _lib.default_sky_options()
        check(lib().cnq_sky_irradiance_device(self.handle, receivers.data_ptr(), n, C.byref(opts), C.byref(so),
                                              out.data_ptr(), faces.data_ptr() if faces is not None else 0,
                                              full.data_ptr() if full is not None else 0, stream.cuda_stream))

    def sky_view_device(self, receivers, width, height, opts, out, faces=None, full=None, sky=None, tile_w=8,
                        stream=None) -> None:
        """The sky over a view's receivers (device tensors), each wave taking a
        tile_w x 64/tile_w patch of pixels; the layouts of sky_device."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream()
        n = width * height
        if receivers.numel() < 8 * n or out.numel() < n or (faces is not None and faces.numel() < 6 * n) or (
                full is not None and full.numel() < 6 * n):
            raise ValueError("tensor too small")
        so = sky if sky is not None else _lib.default_sky_options()
        check(lib().cnq_sky_view_irradiance_device(self.handle, receivers.data_ptr(), width, height, tile_w,
                                                   C.byref(opts), C.byref(so), out.data_ptr(),
                                                   faces.data_ptr() if faces is not None else 0,
                                                   full.data_ptr() if full is not None else 0, stream.cuda_stream))

    def view_query_device(self, receivers, width, height, lights, opts, vis, full, touched, tile_w=16,
                          stream=None) -> None:
        """Cone queries over a view's receivers (device tensors), each wave
        taking a tile_w x 64/tile_w patch of pixels."""
        import torch

        if stream is None:
            stream = torch.cuda.current_stream()
        nl = lights.numel() // 16
        if receivers.numel() < 8 * width * height or vis.numel() < width * height * nl:
            raise ValueError("tensor too small")
        check(lib().cnq_view_irradiance_device(self.handle, receivers.data_ptr(), width, height, lights.data_ptr(),
                                               nl, C.byref(opts), tile_w, vis.data_ptr(), full.data_ptr(),
                                               touched.data_ptr(), stream.cuda_stream))

    def primary_device(self, camera, width, height, lights, opts, hits, receivers, stream=None) -> None:
        import torch

        if stream is None:
            stream = torch.cuda.current_stream()
        nl = lights.numel() // 16
        check(lib().cnq_primary_device(self.handle, camera.data_ptr(), width, height, lights.data_ptr(), nl,
                                       C.byref(opts), hits.data_ptr(), receivers.data_ptr(), stream.cuda_stream))

    def render_view(self, camera: np.ndarray, width: int, height: int, lights: np.ndarray, opts: _lib.Options):
        camera = np.ascontiguousarray(camera, dtype=np.float32)
        lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 16)
        px = width * height
        q = px * len(lights)
        hits = np.zeros((px, 4), np.uint32)
        vis = np.zeros(q, np.float32)
        full = np.zeros(q, np.uint64)
        touched = np.zeros(q, np.uint64)
        check(lib().cnq_render_view(self.handle, _ptr(camera), width, height, _ptr(lights), len(lights),
                                    C.byref(opts), _ptr(hits), _ptr(vis), _ptr(full), _ptr(touched)))
        return hits, vis, full, touched

    def refit(self, tris12: np.ndarray) -> None:
        tris12 = np.ascontiguousarray(tris12, dtype=np.float32)
        if tris12.size != self.flat.tris.size:
            raise ValueError("refit needs the same triangles in the same order")
        check(lib().cnq_scene_refit(self.handle, _ptr(tris12)))

    def refit_device(self, tris12, stream=None) -> None:
        """Refit to a new pose given as a device tensor (12 f32 per triangle,
        the flattened order); returns once the object hierarchy is rebuilt."""
        if tris12.numel() != self.flat.tris.size:
            raise ValueError("refit needs the same triangles in the same order")
        check(lib().cnq_scene_refit_device(self.handle, tris12.data_ptr(), _stream_ptr(stream)))

    def read_blas(self):
        blas = np.zeros_like(self.flat.blas)
        cov = np.zeros_like(self.flat.blas_cov)
        check(lib().cnq_scene_read_blas(self.handle, _ptr(blas), _ptr(cov)))
        return blas, cov

    def counters(self, reset: bool = True):
        c = np.zeros(3, np.uint32)
        w = np.zeros(6, np.uint64)
        check(lib().cnq_scene_counters(self.handle, int(reset), _ptr(c), _ptr(w)))
        return c, w

    def set_count_work(self, on: bool) -> None:
        check(lib().cnq_scene_set_count_work(self.handle, int(on)))

    def set_normals(self, normals=None) -> None:
        """Corner shading normals, 9 floats per triangle in the flat's order
        (None: face normals): the camera pass's receivers then take the
        interpolated shading normal."""
        if normals is None:
            check(lib().cnq_scene_set_normals(self.handle, None, 0))
            return
        nrm = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 9)
        check(lib().cnq_scene_set_normals(self.handle, _ptr(nrm), len(nrm)))

    def set_query_flags(self, flags=None) -> None:
        """Per-query walk flags (a device uint8 tensor, or None for none):
        bit 0 capped, bit 1 deeper than 48 entries, bit 2 truncated. The
        tensor must stay alive while launches write it."""
        self._query_flags = flags
        if flags is None:
            check(lib().cnq_scene_set_query_flags(self.handle, None, 0))
        else:
            check(lib().cnq_scene_set_query_flags(self.handle, flags.data_ptr(), int(flags.numel())))


class Batch:
    """Many views of one scene, each with its own camera, size and lights, in
    one camera pass and one cone launch (optionally a replayed hipGraph)."""

    def __init__(self, scene: Scene, views: Sequence[tuple], lights: np.ndarray):
        """`views`: (camera20, width, height, light_first, light_count) each;
        `lights`: the batch's LightGpu rows."""
        self.scene = scene
        arr = (_lib.View * len(views))()
        self.shapes = []
        for k, (cam, w, h, lf, lc) in enumerate(views):
            arr[k].camera[:] = [float(x) for x in np.asarray(cam, np.float32).reshape(20)]
            arr[k].width, arr[k].height, arr[k].light_first, arr[k].light_count = int(w), int(h), int(lf), int(lc)
            self.shapes.append((int(w), int(h), int(lc)))
        self._lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 16)
        h_ = C.c_void_p()
        check(lib().cnq_batch_create(scene.handle, C.cast(arr, C.c_void_p), len(views), _ptr(self._lights),
                                     len(self._lights), C.byref(h_)))
        self.handle = h_
        self._destroy = lib().cnq_batch_destroy

    def __del__(self):
        if getattr(self, "handle", None):
            self._destroy(self.handle)
            self.handle = None

    @property
    def queries(self) -> int:
        return int(lib().cnq_batch_queries(self.handle))

    @staticmethod
    def tuning(tile_w: int = 16, persistent_blocks: int = 0, waves_per_simd: int = 2, ray_stack: int = 64) -> _lib.Tuning:
        return _lib.Tuning(tile_w, persistent_blocks, waves_per_simd, ray_stack)

    def run(self, opts: _lib.Options, tuning: _lib.Tuning, stream=None) -> None:
        check(lib().cnq_batch_run(self.handle, C.byref(opts), C.byref(tuning), _stream_ptr(stream)))

    def capture(self, opts: _lib.Options, tuning: _lib.Tuning) -> None:
        check(lib().cnq_batch_capture(self.handle, C.byref(opts), C.byref(tuning)))

    def replay(self, stream=None) -> None:
        check(lib().cnq_batch_replay(self.handle, _stream_ptr(stream)))

    def autotune(self, opts: _lib.Options, stream=None):
        best = _lib.Tuning()
        ms = C.c_float()
        check(lib().cnq_batch_autotune(self.handle, C.byref(opts), _stream_ptr(stream), C.byref(best), C.byref(ms)))
        return best, float(ms.value)

    def read(self, view: int):
        """(hits (px, 4), visible, full, touched) of one view, queries light-major."""
        w, h, nl = self.shapes[view]
        px, q = w * h, w * h * nl
        hits = np.zeros((px, 4), np.uint32)
        vis = np.zeros(q, np.float32)
        full = np.zeros(q, np.uint64)
        touched = np.zeros(q, np.uint64)
        check(lib().cnq_batch_read(self.handle, view, _ptr(hits), _ptr(vis), _ptr(full), _ptr(touched)))
        return hits, vis, full, touched

    def set_lights(self, lights: np.ndarray) -> None:
        self._lights = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 16)
        check(lib().cnq_batch_set_lights(self.handle, _ptr(self._lights), len(self._lights)))


class Server:
    """The served form over one resident scene: view requests admitted in
    arrival order (an express slot may take small views first), batched per
    launch, launches in flight on several streams; light edits apply when
    submitted. Keyword arguments set `cnq_server_config` fields (and
    `tile_w`/`persistent_blocks` of its tuning); see `cnq_server_*` in
    include/cnq.h."""

    def __init__(self, scene: Scene, lights: np.ndarray, opts: Optional[_lib.Options] = None, **config):
        self.scene = scene
        cfg = _lib.ServerConfig()
        lib().cnq_server_config_default(C.byref(cfg))
        for k, v in config.items():
            if k in ("tile_w", "persistent_blocks"):
                setattr(cfg.tuning, k, int(v))
            else:
                setattr(cfg, k, int(v))
        self.config = cfg
        self.opts = opts or default_options(hierarchy=2)
        lt = np.ascontiguousarray(lights, dtype=np.float32).reshape(-1, 16)
        h = C.c_void_p()
        check(lib().cnq_server_create(scene.handle, C.byref(cfg), C.byref(self.opts), _ptr(lt), len(lt), C.byref(h)))
        self.handle = h
        self._destroy = lib().cnq_server_destroy
        self._shapes = {}

    def close(self) -> None:
        if getattr(self, "handle", None):
            self._destroy(self.handle)
            self.handle = None

    def __del__(self):
        self.close()

    def submit_view(self, camera, width: int, height: int, lights: Sequence[int]) -> int:
        cam = np.ascontiguousarray(camera, dtype=np.float32).reshape(20)
        idx = np.ascontiguousarray(list(lights), dtype=np.uint32)
        rid = C.c_uint64()
        check(lib().cnq_server_submit_view(self.handle, _ptr(cam), int(width), int(height), _ptr(idx), len(idx),
                                           C.byref(rid)))
        self._shapes[rid.value] = (int(width), int(height), len(idx))
        return rid.value

    def submit_light(self, index: int, light: np.ndarray) -> int:
        lt = np.ascontiguousarray(light, dtype=np.float32).reshape(16)
        rid = C.c_uint64()
        check(lib().cnq_server_submit_light(self.handle, int(index), _ptr(lt), C.byref(rid)))
        return rid.value

    def wait(self, rid: int, timeout_ms: float = 60_000.0) -> bool:
        done = C.c_int()
        check(lib().cnq_server_wait(self.handle, int(rid), float(timeout_ms), C.byref(done)))
        return bool(done.value)

    def info(self, rid: int) -> _lib.RequestInfo:
        out = _lib.RequestInfo()
        check(lib().cnq_server_info(self.handle, int(rid), C.byref(out)))
        return out

    def read(self, rid: int):
        """(hits (px, 4), visible, full, touched) of a done view, queries light-major."""
        w, h, nl = self._shapes[rid]
        px, q = w * h, w * h * nl
        hits = np.zeros((px, 4), np.uint32)
        vis = np.zeros(q, np.float32)
        full = np.zeros(q, np.uint64)
        touched = np.zeros(q, np.uint64)
        check(lib().cnq_server_read(self.handle, int(rid), _ptr(hits), _ptr(vis), _ptr(full), _ptr(touched)))
        return hits, vis, full, touched

    def release(self, rid: int) -> None:
        check(lib().cnq_server_release(self.handle, int(rid)))
        self._shapes.pop(rid, None)

    def counts(self) -> dict:
        c = np.zeros(6, np.uint64)
        check(lib().cnq_server_counts(self.handle, _ptr(c)))
        return dict(zip(("views", "edits", "launches", "queries", "pixels_held", "queries_held"), (int(x) for x in c)))


def block_planes(texels: np.ndarray, blocks: np.ndarray, light: np.ndarray) -> np.ndarray:
    """Lightmap::block_planes (lightmap.rs:256-333): (n_blocks, 16, 4) planes
    (n, w; inside n.p <= w) per block of 64 texels toward a rect light."""
    texels = np.ascontiguousarray(texels, dtype=np.float32).reshape(-1, 64, 8)
    blocks = np.ascontiguousarray(blocks, dtype=np.float32).reshape(-1, 6)
    light = np.ascontiguousarray(light, dtype=np.float32).reshape(16)
    if len(texels) != len(blocks):
        raise ValueError("one box per block of 64 texels")
    out = np.zeros((len(blocks), 16, 4), np.float32)
    check(lib().cnq_block_planes(_ptr(texels), _ptr(blocks), len(blocks), _ptr(light), _ptr(out)))
    return out


class LiveLightmap:
    """A lightmap's direct irradiance from one rect light, kept current as
    dynamic meshes move: the lighting cache of the reference's
    `LiveRenderer` (live.rs:255-1090). Texels are 64 receivers per block
    (`Lightmap::receivers`) with each block's bounds (`Lightmap::blocks`);
    `irradiance` is the baked lightmap (`Lightmap::irradiance`), or None to
    bake it here (`MaskConeKernel::bake`, lib.rs:2961)."""

    def __init__(self, scene: Scene, texels: np.ndarray, blocks: np.ndarray, light: np.ndarray,
                 dynamic: Sequence[bool], irradiance: Optional[np.ndarray] = None,
                 opts: Optional[_lib.Options] = None):
        self.scene = scene
        self._texels = np.ascontiguousarray(texels, dtype=np.float32).reshape(-1, 8)
        self._blocks = np.ascontiguousarray(blocks, dtype=np.float32).reshape(-1, 6)
        if len(self._texels) != 64 * len(self._blocks):
            raise ValueError("64 texels per block")
        self._light = np.ascontiguousarray(light, dtype=np.float32).reshape(16)
        self._dynamic = np.ascontiguousarray([1 if d else 0 for d in dynamic], dtype=np.uint8)
        self.counts = [int(c) for c in scene.flat.ranges[:, 1]]
        if len(self._dynamic) != len(self.counts):
            raise ValueError("one dynamic flag per mesh")
        self.opts = opts or default_options(hierarchy=2)
        irr = None if irradiance is None else np.ascontiguousarray(irradiance, dtype=np.float32)
        h = C.c_void_p()
        check(lib().cnq_cache_create(scene.handle, _ptr(self._texels), _ptr(self._blocks), len(self._blocks),
                                     _ptr(self._light), _ptr(self._dynamic), 0 if irr is None else _ptr(irr),
                                     C.byref(self.opts), C.byref(h)))
        self.handle = h
        self._destroy = lib().cnq_cache_destroy

    def __del__(self):
        if getattr(self, "handle", None):
            self._destroy(self.handle)
            self.handle = None

    @property
    def blocks(self) -> int:
        """Total lightmap blocks (live.rs:1088)."""
        return len(self._blocks)

    def update_meshes(self, changes: Sequence[tuple], stream=None) -> None:
        """LiveRenderer::update_meshes (live.rs:795): (mesh, triangles) pairs,
        each mesh dynamic and keeping its triangle count."""
        if not changes:
            return
        ids = np.ascontiguousarray([int(m) for m, _ in changes], dtype=np.uint32)
        parts = []
        for m, tris in changes:
            t = np.ascontiguousarray(tris, dtype=np.float32).reshape(-1, 9)
            if len(t) != self.counts[int(m)]:
                raise ValueError(f"mesh {m} keeps its {self.counts[int(m)]} triangles")
            parts.append(t)
        tris = np.ascontiguousarray(np.concatenate(parts), dtype=np.float32)
        check(lib().cnq_cache_update(self.handle, _ptr(ids), len(ids), _ptr(tris), _stream_ptr(stream)))

    def read(self):
        """(texel irradiance, the blocks the last update re-baked, sorted)."""
        irr = np.zeros(64 * self.blocks, np.float32)
        marked = np.zeros(max(self.blocks, 1), np.uint32)
        n = C.c_uint32()
        check(lib().cnq_cache_read(self.handle, _ptr(irr), _ptr(marked), C.byref(n)))
        return irr, np.sort(marked[: n.value])

    def read_texels(self) -> np.ndarray:
        """LiveRenderer::read_texels (live.rs:1078)."""
        return self.read()[0]

    def rebaked_blocks(self) -> int:
        """LiveRenderer::rebaked_blocks (live.rs:1083)."""
        return len(self.read()[1])


class ViewCache:
    """The view cache's pages (the reference's `VirtualCache`, vcache.rs):
    9 x 9 texels per page of one chart of one object at one mip, baked with
    one mask cone per texel and light, kept current as objects move. Keys
    are (object, chart | mip << 24, page x | page y << 16).

    `surface` is a continuum_amd.surfaces.Surface over the scene's triangles
    (charts numbered globally); `lights` LightGpu rows (at most 16)."""

    def __init__(self, scene: Scene, surface, lights: np.ndarray, opts: Optional[_lib.Options] = None, **config):
        self.scene = scene
        self.opts = opts or default_options()
        cfg = _lib.VcacheConfig()
        lib().cnq_vcache_config_default(C.byref(cfg))
        for k, v in config.items():
            setattr(cfg, k, v)
        self.config = cfg
        self._tri_chart = np.ascontiguousarray(surface.tri_chart, np.uint32)
        self._tri_uv = np.ascontiguousarray(surface.tri_uv, np.float32).reshape(-1, 6)
        self._normals = np.ascontiguousarray(surface.normals, np.float32).reshape(-1, 9)
        self._chart_size = np.ascontiguousarray(surface.chart_size, np.float32).reshape(-1, 2)
        self._chart_grid = np.ascontiguousarray(surface.chart_grid, np.uint32).reshape(-1, 2)
        n_tris = len(scene.flat.tris)
        if not (len(self._tri_chart) == len(self._tri_uv) == len(self._normals) == n_tris):
            raise ValueError("one chart, 6 chart coordinates and 9 normal floats per scene triangle")
        sf = _lib.VcacheSurfaces(_ptr(self._tri_chart), _ptr(self._tri_uv), _ptr(self._normals),
                                 len(self._chart_size), _ptr(self._chart_size), _ptr(self._chart_grid))
        self.lights = np.ascontiguousarray(lights, np.float32).reshape(-1, 16)
        self.channels = len(self.lights) + 1
        h = C.c_void_p()
        check(lib().cnq_vcache_create(scene.handle, C.byref(sf), _ptr(self.lights), len(self.lights), C.byref(cfg),
                                      C.byref(self.opts), C.byref(h)))
        self.handle = h
        self._destroy = lib().cnq_vcache_destroy

    def __del__(self):
        if getattr(self, "handle", None):
            self._destroy(self.handle)
            self.handle = None

    @staticmethod
    def key(obj: int, chart: int, mip: int, px: int, py: int) -> tuple:
        return (int(obj), int(chart) | (int(mip) << 24), int(px) | (int(py) << 16))

    def grid(self, obj: int, chart: int, mip: int) -> tuple:
        out = np.zeros(2, np.uint32)
        check(lib().cnq_vcache_grid(self.handle, int(obj), int(chart), int(mip), _ptr(out)))
        return int(out[0]), int(out[1])

    def pages_of(self, obj: int, chart: int, mip: int) -> list:
        """Every page key of a chart at a mip."""
        gx, gy = self.grid(obj, chart, mip)
        return [self.key(obj, chart, mip, x, y) for y in range((gy + 7) // 8) for x in range((gx + 7) // 8)]

    def request(self, keys) -> np.ndarray:
        k = np.ascontiguousarray(keys, np.uint32).reshape(-1, 3)
        slots = np.zeros(len(k), np.uint32)
        check(lib().cnq_vcache_request(self.handle, _ptr(k), len(k), _ptr(slots)))
        return slots

    def evict(self, keys) -> None:
        k = np.ascontiguousarray(keys, np.uint32).reshape(-1, 3)
        check(lib().cnq_vcache_evict(self.handle, _ptr(k), len(k)))

    def set_moved(self, moved, scene_box=None) -> None:
        arr = (_lib.MovedBox * max(len(moved), 1))(*moved)
        sb = None if scene_box is None else np.ascontiguousarray(scene_box, np.float32).reshape(6)
        check(lib().cnq_vcache_set_moved(self.handle, arr, len(moved), 0 if sb is None else _ptr(sb)))

    def frame(self, stream=None) -> dict:
        st = np.zeros(10, np.uint64)
        check(lib().cnq_vcache_frame(self.handle, _stream_ptr(stream), _ptr(st)))
        names = ("planned_pages", "planned_jobs", "appended_pages", "appended_jobs", "candidates", "invalid",
                 "stale_pages", "overflow", "visible_tested", "evicted")
        return {n: int(v) for n, v in zip(names, st)}

    def read(self, keys):
        """(texels (n, 81, channels), slots: NONE where absent or stale)."""
        k = np.ascontiguousarray(keys, np.uint32).reshape(-1, 3)
        tex = np.zeros((len(k), 81, self.channels), np.float32)
        slots = np.zeros(len(k), np.uint32)
        if len(k):
            check(lib().cnq_vcache_read(self.handle, _ptr(k), len(k), _ptr(tex), _ptr(slots)))
        return tex, slots

    @staticmethod
    def shade_params(camera, pixels_per_texel: float = 2.0, refine_levels: int = 3, request_tile: int = 4):
        """The shading pass's parameters for a camera uniform (pixel angle
        2 tan(fov/2) / height, from its params)."""
        cam = np.asarray(camera, np.float32).reshape(20)
        angle = np.float32(np.float32(2.0) * cam[18]) / cam[17]
        return _lib.VcacheShadeParams(pixels_per_texel, refine_levels, request_tile, float(angle))

    def shade_device(self, camera, width: int, height: int, hits, out, out_page, params=None, stream=None):
        """Shading from the cache for a camera pass's hits (device tensors):
        out (lights, h, w) f32, out_page (h, w) u32. Returns (requests,
        pixels without a page)."""
        cam = np.ascontiguousarray(camera, np.float32).reshape(20)
        p = params or self.shade_params(cam)
        counts = np.zeros(2, np.uint32)
        if stream is None:
            import torch

            stream = torch.cuda.current_stream()
        check(lib().cnq_vcache_shade(self.handle, _ptr(cam), int(width), int(height), hits.data_ptr(), out.data_ptr(),
                                     out_page.data_ptr(), C.byref(p), _stream_ptr(stream), _ptr(counts)))
        return int(counts[0]), int(counts[1])

    def plan(self, budget: int) -> dict:
        st = np.zeros(4, np.uint64)
        check(lib().cnq_vcache_plan(self.handle, int(budget), _ptr(st)))
        return dict(requests=int(st[0]), planned=int(st[1]), cones=int(st[2]), evicted=int(st[3]))

    def set_lights(self, lights, first: int = 0, stream=None) -> dict:
        """Lights [first, first + n) replaced (moved, resized, re-aimed) and
        their channels of every resident page baked afresh."""
        rows = np.ascontiguousarray(lights, np.float32).reshape(-1, 16)
        st = np.zeros(2, np.uint64)
        check(lib().cnq_vcache_set_lights(self.handle, _ptr(rows), int(first), len(rows), _stream_ptr(stream),
                                          _ptr(st)))
        self.lights[first:first + len(rows)] = rows
        return dict(pages=int(st[0]), cones=int(st[1]))

    LEDGER_STATS = ("pages", "texels", "set_aside", "checked", "mixed", "held_whole", "jobs", "held_lattice",
                    "grew", "unbounded")

    def set_moving_lights(self, lights, epsilon: float = 0.01) -> None:
        """Lights (indices into the table) that move through the error ledger
        (the reference's `moving` list), with threshold `epsilon`; before
        the first frame."""
        idx = np.ascontiguousarray(lights, np.uint32).reshape(-1)
        check(lib().cnq_vcache_set_moving_lights(self.handle, _ptr(idx), len(idx), float(epsilon)))
        self.moving = [int(i) for i in idx]

    def move_light(self, index: int, light) -> None:
        """Moving light `index` is `light` from the next frame."""
        row = np.ascontiguousarray(light, np.float32).reshape(16)
        check(lib().cnq_vcache_move_light(self.handle, int(index), _ptr(row)))
        self.lights[int(index)] = row

    def ledger_stats(self) -> dict:
        out = np.zeros(10, np.uint64)
        check(lib().cnq_vcache_ledger_stats(self.handle, _ptr(out)))
        return {n: int(v) for n, v in zip(self.LEDGER_STATS, out)}

    def ledger_state(self):
        """(records (pages, moving, 48) u32, bounds (pages, moving) f32)."""
        pages, m = int(self.config.pages), len(self.moving)
        rec = np.zeros((pages, m, 48), np.uint32)
        bound = np.zeros((pages, m), np.float32)
        check(lib().cnq_vcache_ledger_state(self.handle, _ptr(rec), _ptr(bound)))
        return rec, bound

    def verify_ledger(self, light: int, stream=None) -> dict:
        """`verify_ledger` (instanced.rs:3741): every page shown baked afresh
        for moving light `light` against the cache."""
        out = np.zeros(6, np.float32)
        check(lib().cnq_vcache_verify_ledger(self.handle, int(light), _stream_ptr(stream), _ptr(out)))
        return dict(pages=int(out[0]), violations=int(out[1]), over_epsilon=int(out[2]), changed=int(out[3]),
                    largest_change=float(out[4]), largest_bound=float(out[5]))

    def verify(self, stream=None) -> tuple:
        """`verify_cache` (instanced.rs:3721): every page shown baked afresh
        and compared bit for bit; (values that differ, values compared)."""
        out = np.zeros(2, np.uint64)
        check(lib().cnq_vcache_verify(self.handle, _stream_ptr(stream), _ptr(out)))
        return int(out[0]), int(out[1])

    def state(self) -> dict:
        sizes = np.zeros(5, np.uint32)
        check(lib().cnq_vcache_state(self.handle, 0, 0, 0, 0, 0, 0, _ptr(sizes)))
        pages, ch, tlen = int(sizes[0]), int(sizes[1]), int(sizes[2])
        info = np.zeros((pages, 12), np.uint32)
        stale = np.zeros(pages, np.uint32)
        used = np.zeros(pages, np.uint32)
        pool = np.zeros((pages, 81, ch), np.float32)
        td = np.zeros((tlen, 4), np.uint32)
        th = np.zeros((tlen, 4), np.uint32)
        check(lib().cnq_vcache_state(self.handle, _ptr(info), _ptr(stale), _ptr(used), _ptr(pool), _ptr(td), _ptr(th),
                                     _ptr(sizes)))
        return dict(slot_info=info, slot_stale=stale, slot_used=used, pool=pool, table_device=td, table_host=th,
                    slot_extent=int(sizes[3]), frame=int(sizes[4]))


def _stream_ptr(stream) -> int:
    if stream is None:
        return 0
    return int(getattr(stream, "cuda_stream", stream))


def receivers_array(receivers: Sequence) -> np.ndarray:
    """ReceiverGpu rows (lib.rs:3353-3359) from (point, normal) pairs or an
    (n, 2, 3) / (n, 6) array."""
    a = np.asarray(receivers, dtype=np.float32).reshape(-1, 6)
    out = np.zeros((len(a), 8), np.float32)
    out[:, 0:3] = a[:, 0:3]
    out[:, 4:7] = a[:, 3:6]
    return out


class MaskConeKernel:
    """The reference's kernel object, on one MI355X (lib.rs:65)."""

    def __init__(self, device: Device):
        self.device = device
        self._prefilter = 0.0
        self._wide = True
        self._nearest_first = True
        self._profile_stage = 0
        self._sky_max = False
        self._sky_prefilter = 1.0
        self._sky_horizon_cull = True

    @classmethod
    def new(cls, ordinal: int = 0) -> Optional["MaskConeKernel"]:
        """None without a GPU, as lib.rs:1155."""
        try:
            return cls(Device(ordinal))
        except CnqError as e:
            if e.code == _lib.NO_DEVICE:
                return None
            raise

    # setters (lib.rs:1501-1675)
    def set_wide_cones(self, on: bool) -> None:
        self._wide = bool(on)

    def set_prefilter(self, bins: float) -> None:
        self._prefilter = max(float(bins), 0.0)

    def prefilter(self) -> float:
        return self._prefilter

    def set_nearest_first(self, on: bool) -> None:
        self._nearest_first = bool(on)

    def set_profile_stage(self, stage: int) -> None:
        self._profile_stage = int(stage)

    def options(self, hierarchy: Hierarchy = Hierarchy.TRIANGLES) -> _lib.Options:
        return default_options(
            hierarchy=int(hierarchy), prefilter=self._prefilter,
            wide_extent=8.0 if self._wide else 1e30,
            light_cone_cap=4096 if self._wide else NONE, wide_cone_cap=512 if self._wide else NONE,
            profile_stage=self._profile_stage, nearest_first=int(self._nearest_first))

    def scene(self, meshes) -> Scene:
        return Scene(self.device, Flat.build(meshes))

    @staticmethod
    def _light_row(light) -> np.ndarray:
        if isinstance(light, np.ndarray):
            return light.astype(np.float32).reshape(16)
        corner, edge_u, edge_v, one_sided = light
        return rect_light(corner, edge_u, edge_v, one_sided)

    def _run(self, receivers, light, meshes, runs, hierarchy, want_cov=False):
        scene = self.scene(meshes)
        recv = receivers_array(receivers)
        lrow = self._light_row(light)[None, :]
        opts = self.options(hierarchy)
        best = float("inf")
        out = None
        for _ in range(max(1, runs)):
            t0 = time.perf_counter()
            out = scene.query(recv, lrow, opts, want_cov)
            best = min(best, time.perf_counter() - t0)
        return out, best

    def irradiance(self, receivers, light, meshes, runs: int = 1):
        """Every receiver tests every mesh's triangles (lib.rs:1706)."""
        return self.irradiance_with(receivers, light, meshes, Hierarchy.NONE, runs)

    def irradiance_with(self, receivers, light, meshes, hierarchy=Hierarchy.TRIANGLES, runs: int = 1):
        """(irradiance per unit radiance per receiver, best seconds) (lib.rs:1718)."""
        if isinstance(hierarchy, bool):
            hierarchy = Hierarchy.TRIANGLES if hierarchy else Hierarchy.NONE
        (vis, _, _, _), best = self._run(receivers, light, meshes, runs, hierarchy)
        return vis, best

    def bake(self, lightmap, light, meshes) -> float:
        """One cone per lightmap texel toward `light` through the meshes'
        hierarchy, into `lightmap.irradiance`; returns the time (lib.rs:2961)."""
        scene = self.scene(meshes)
        best = float("inf")
        t0 = time.perf_counter()
        vis, _, _, _ = scene.query(lightmap.texels, self._light_row(light)[None, :], self.options(Hierarchy.TRIANGLES))
        best = time.perf_counter() - t0
        lightmap.irradiance = vis
        return best

    def set_sky_max(self, on: bool) -> None:
        """Sky cones combine meshes by the maximum (a diagnostic, lib.rs:103)."""
        self._sky_max = bool(on)

    def set_sky_prefilter(self, scale: float) -> None:
        """The sky cones' prefilter as a multiple of the lights' (lib.rs:104)."""
        self._sky_prefilter = float(scale)

    def set_sky_horizon_cull(self, on: bool) -> None:
        """Sky walks skip geometry wholly below the receiver's horizon (lib.rs:105)."""
        self._sky_horizon_cull = bool(on)

    def sky_options(self) -> _lib.SkyOptions:
        return _lib.default_sky_options(max_combine=int(self._sky_max), horizon_cull=int(self._sky_horizon_cull),
                                        prefilter_scale=self._sky_prefilter)

    def sky_irradiance(self, receivers, meshes, hierarchy=Hierarchy.TRIANGLES) -> np.ndarray:
        """Per receiver, the irradiance per unit radiance of a uniform sky
        through six cube-face mask cones (WGSL :1411 `sky_irradiance`, the
        sky term of `per_pixel_hit_color` :2282-2289)."""
        out, _, _ = self.scene(meshes).sky(receivers_array(receivers), self.options(hierarchy), self.sky_options())
        return out

    def coverage(self, receivers, light, meshes) -> np.ndarray:
        """Per receiver, the 64 bins' coverage (lib.rs:3289; every triangle tested)."""
        (_, _, _, cov), _ = self._run(receivers, light, meshes, 1, Hierarchy.NONE, want_cov=True)
        return cov

    def render_view(self, camera, width, height, lights, meshes, hierarchy=Hierarchy.TRIANGLES):
        lights = np.asarray(lights, dtype=np.float32).reshape(-1, 16)
        if len(lights) > MAX_LIGHTS:
            raise ValueError(f"at most {MAX_LIGHTS} lights (lib.rs:1888)")
        return self.scene(meshes).render_view(camera, width, height, lights, self.options(hierarchy))
