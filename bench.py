#!/usr/bin/env python3
"""Cone-visibility throughput on one MI355X.

A step is one plain per-view launch over the 1M-triangle rock field: the
camera pass (1280 x 720 shading points) and one cone query per shading point
and light (3 lights: a 0.53 degree sun disk and two area lights), on inputs
already resident in HBM. Prints one JSON line: queries/s over the timed
steps, p50/p99 view latency, the cone kernel's roofline, the oracle's
throughput on the host's cores for a bounded sample of the same queries, and
a bit-for-bit comparison of that sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "cone-visibility queries/s; p50/p99 view latency, batched/served vs plain"
PEAK_HBM_GBPS = 8000.0
PROFILE_DIR = ROOT / "profiles"


# The plain launch's tile: 16 x 4 shading points a wave (a sweep over 4, 8,
# 16, 32 and 64 on the rock field puts it first).
PLAIN_TILE_W = 16


def algorithmic_bytes(work, queries: int, receivers: int, lights: int) -> dict:
    """Bytes a launch must move, per unit of work the kernel's counters
    report: 32 B per receiver read, 64 B per light read, 20 B of results
    (irradiance f32 + two u64 masks) per query; 32 B per node tested (object
    and triangle hierarchies, compact nodes), 36 B per triangle tested, 16 B
    range + 4 B item per mesh entered, 108 B of coverage samples per splat."""
    obj_nodes, blas_nodes, tris_tested, _drawn, meshes, splats = (int(x) for x in work)
    per = {
        "receivers": 32 * queries,
        "lights": 64 * queries,
        "results": 20 * queries,
        "object_nodes": 32 * obj_nodes,
        "blas_nodes": 32 * blas_nodes,
        "triangles": 36 * tris_tested,
        "meshes": 20 * meshes,
        "splats": 108 * splats,
    }
    per["total"] = sum(per.values())
    return per


def read_traffic(workload: str):
    """HBM bytes per cone launch measured by rocprofv3 --pmc (FETCH_SIZE and
    WRITE_SIZE, calibrated on a copy of known size), from the committed
    summary, if one exists for this workload."""
    f = PROFILE_DIR / f"cone_items_{workload}_traffic.md"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def courtyard_forms(device, steps: int) -> dict:
    """The 3M-triangle courtyard: (1) the plain per-view frame with its
    per-frame refit (curtains displaced by a sine field of time), camera pass
    and cones for 1280x720 x 4 lights; (2) 64 views of uneven size and 1-16
    lights as 64 sequential plain launches; (3) the same views as one batch
    (autotuned on a subset, captured as a hipGraph and replayed)."""
    import numpy as np
    import torch

    from continuum_amd import _lib
    from continuum_amd.kernel import Batch, Flat, Scene
    from continuum_amd.scenes import courtyard, courtyard_frame, meshes_to_arrays
    from continuum_amd.workloads import courtyard_views, view_queries

    out = {}
    sc = courtyard()
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    scene = Scene(device, flat)
    opts = _lib.default_options(hierarchy=2, prefilter=sc.prefilter)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    W, H, nl = sc.width, sc.height, len(sc.lights)
    cam = torch.from_numpy(sc.camera).to(dev)
    lights = torch.from_numpy(np.ascontiguousarray(sc.lights)).to(dev)
    hits = torch.empty((W * H, 4), dtype=torch.int32, device=dev)
    recv = torch.empty((W * H, 8), dtype=torch.float32, device=dev)
    vis = torch.empty(W * H * nl, dtype=torch.float32, device=dev)
    full = torch.empty(W * H * nl, dtype=torch.int64, device=dev)
    touched = torch.empty(W * H * nl, dtype=torch.int64, device=dev)
    poses = [torch.from_numpy(courtyard_frame(sc, f, flat.tris, flat.ranges)).to(dev) for f in range(steps + 1)]
    torch.cuda.synchronize()

    def frame(f):
        scene.refit_device(poses[f], stream)
        scene.primary_device(cam, W, H, lights, opts, hits, recv, stream)
        scene.view_query_device(recv, W, H, lights, opts, vis, full, touched, tile_w=PLAIN_TILE_W, stream=stream)

    frame(0)
    torch.cuda.synchronize()
    # The refit alone, wall clock: leaves, internal nodes by height, mesh
    # bounds, the host's object hierarchy over them and its upload.
    refit_ms = []
    for f in range(1, min(steps, 10) + 1):
        t0 = time.perf_counter()
        scene.refit_device(poses[f], stream)
        torch.cuda.synchronize()
        refit_ms.append((time.perf_counter() - t0) * 1e3)
    lat = []
    t0 = time.perf_counter()
    for f in range(1, steps + 1):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        frame(f)
        e1.record(stream)
        lat.append((e0, e1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = np.array([a.elapsed_time(b) for a, b in lat])
    q = W * H * nl
    out["courtyard_plain"] = {
        "workload": "courtyard 3.05M triangles, 1280x720 x 4 lights, per-frame GPU refit of the curtains",
        "queries_per_s": steps * q / wall, "ms_per_frame": wall / steps * 1e3,
        "p50_ms": float(np.percentile(ms, 50)), "p99_ms": float(np.percentile(ms, 99)),
        "triangles": int(sc.n_tris), "refit_ms_p50": float(np.percentile(refit_ms, 50))}

    views, vlights = courtyard_views()
    total_q = view_queries(views)
    # 64 sequential plain launches (camera pass + cone pass per view).
    maxpx = max(w * h for _, w, h, _, _ in views)
    maxq = max(w * h * n for _, w, h, _, n in views)
    s_hits = torch.empty((maxpx, 4), dtype=torch.int32, device=dev)
    s_recv = torch.empty((maxpx, 8), dtype=torch.float32, device=dev)
    s_vis = torch.empty(maxq, dtype=torch.float32, device=dev)
    s_full = torch.empty(maxq, dtype=torch.int64, device=dev)
    s_touched = torch.empty(maxq, dtype=torch.int64, device=dev)
    v_cams = [torch.from_numpy(c).to(dev) for c, *_ in views]
    v_lights = [torch.from_numpy(np.ascontiguousarray(vlights[lf:lf + n])).to(dev) for _, _, _, lf, n in views]

    def sequential():
        for k, (_, w, h, lf, n) in enumerate(views):
            scene.primary_device(v_cams[k], w, h, v_lights[k], opts, s_hits, s_recv, stream)
            scene.view_query_device(s_recv, w, h, v_lights[k], opts, s_vis, s_full, s_touched, tile_w=PLAIN_TILE_W,
                                    stream=stream)

    sequential()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sequential()
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0

    probe = Batch(scene, views[::8], vlights)
    best, _ = probe.autotune(opts)
    del probe
    batch = Batch(scene, views, vlights)
    batch.capture(opts, best)
    batch.replay()
    torch.cuda.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        batch.replay()
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    # The sequential buffers still hold the last view: the batch's results
    # for it must be the same bits.
    _, lw, lh, _, ln = views[-1]
    lq = lw * lh * ln
    b_hits, b_vis, b_full, b_touched = batch.read(len(views) - 1)
    last_view_differ = {
        "hits": int((s_hits[:lw * lh].cpu().numpy().view(np.uint32) != b_hits).any(axis=1).sum()),
        "visible": int((s_vis[:lq].cpu().numpy().view(np.uint32) != b_vis.view(np.uint32)).sum()),
        "full": int((s_full[:lq].cpu().numpy().view(np.uint64) != b_full).sum()),
        "touched": int((s_touched[:lq].cpu().numpy().view(np.uint64) != b_touched).sum()),
        "queries": int(lq)}
    out["courtyard_64_views"] = {
        "workload": "64 views over the courtyard, sizes from {256^2, 640x360, 1280x720, 1920x1080}, 1-16 lights each",
        "queries": total_q,
        "sequential_plain": {"queries_per_s": total_q / t_seq, "s": t_seq},
        "batched_graph": {"queries_per_s": total_q / t_batch, "s": t_batch,
                          "tuning": {"tile_w": best.tile_w, "persistent_blocks": best.persistent_blocks}},
        "batched_over_sequential": t_seq / t_batch,
        "last_view_batched_vs_plain_differ": last_view_differ,
    }
    return out


# Served form: cone launches cut into pieces of at most 4M queries, and a
# third slot that takes views of at most 1M queries (a 1280x720 view with one
# light) ahead of larger ones.
SERVED_LAUNCH_QUERIES = 1 << 22
SERVED_EXPRESS_QUERIES = 1 << 20


def served_form(device, n_requests: int, seconds: float) -> dict:
    """The served form over the 1M-instance boulder field (10 boulder shapes,
    a 32-light table): a seeded Poisson stream of view requests (a random
    size from 256^2 to 1920x1080, 1-32 lights) and light edits goes through
    the continuous-batching server. Run 1 offers the stream faster than the
    GPU drains it (throughput); run 2 offers it at half that throughput
    (latency under load). Views of run 2 are also rendered one by one as
    plain per-view launches, for the plain rate on the same work."""
    import threading

    import numpy as np
    import torch

    from continuum_amd import _lib
    from continuum_amd.kernel import Flat, Scene, Server
    from continuum_amd.scenes import boulder_field
    from continuum_amd.workloads import serving_requests

    t_build = time.perf_counter()
    tris9, counts, pl, lights, _, _ = boulder_field()
    flat = Flat.build_instanced(tris9, counts, pl, object_prefilter=True)
    scene = Scene(device, flat)
    t_build = time.perf_counter() - t_build
    opts = _lib.default_options(hierarchy=2, prefilter=1.0, object_cov=flat.object_cov)

    def run(reqs):
        srv = Server(scene, lights, opts, pixel_capacity=40 << 20, query_capacity=400 << 20,
                     max_lights=len(lights), max_views_per_launch=16, slots=3,
                     launch_queries=SERVED_LAUNCH_QUERIES, express_queries=SERVED_EXPRESS_QUERIES)
        first_view = next(p for _, k, p in reqs if k == "view")
        warm = srv.submit_view(first_view[0], 64, 64, [0])
        srv.wait(warm)
        srv.release(warm)
        ids, done, kept = [], [], []
        lock = threading.Condition()

        def collector():
            k = 0
            while True:
                with lock:
                    while k >= len(ids):
                        lock.wait()
                    rid, kind, view = ids[k]
                if rid is None:
                    return
                k += 1
                if kind != "view":
                    continue
                srv.wait(rid)
                info = srv.info(rid)
                # The first view of at most 1M queries is kept, to compare
                # with its own plain launch afterwards.
                if not kept and info.width * info.height * info.n_lights <= 1 << 20:
                    kept.append((view, srv.read(rid)))
                done.append((info.latency_ms, info.queued_ms, info.width * info.height * info.n_lights,
                             time.perf_counter()))
                srv.release(rid)

        th = threading.Thread(target=collector)
        th.start()
        table = np.array(lights, np.float32, copy=True)  # the light table as each request is submitted
        t0 = time.perf_counter()
        for t_arr, kind, payload in reqs:
            delay = t0 + t_arr - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            rid = srv.submit_view(*payload) if kind == "view" else srv.submit_light(*payload)
            if kind == "view":
                cam_, w_, h_, idx_ = payload
                view = (cam_, w_, h_, table[list(idx_)].copy())
            else:
                table[payload[0]] = np.asarray(payload[1], np.float32).reshape(16)
                view = None
            with lock:
                ids.append((rid, kind, view))
                lock.notify()
        with lock:
            ids.append((None, None, None))
            lock.notify()
        th.join()
        wall = max(d[3] for d in done) - t0
        counts_ = srv.counts()
        srv.close()
        lat = np.array([d[0] for d in done])
        q = sum(d[2] for d in done)
        # Latency by the view's queries (pixels x lights): at most 1M (a
        # 1280x720 view with one light), at most 4M, and larger.
        sizes = np.array([d[2] for d in done])
        by_size = {}
        for name, lo, hi in (("le_1M_queries", 0, 1 << 20), ("le_4M_queries", (1 << 20) + 1, 1 << 22),
                             ("gt_4M_queries", (1 << 22) + 1, 1 << 62)):
            m = (sizes >= lo) & (sizes <= hi)
            if m.any():
                by_size[name] = {"views": int(m.sum()), "p50_ms": float(np.percentile(lat[m], 50)),
                                 "p99_ms": float(np.percentile(lat[m], 99))}
        return {"requests": len(reqs), "views": len(done), "edits": counts_["edits"], "launches": counts_["launches"],
                "queries": int(q), "s": wall, "queries_per_s": q / wall,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "queued_p50_ms": float(np.percentile([d[1] for d in done], 50)), "by_size": by_size,
                "_kept": kept}

    # Run 1: everything offered within the first 10 ms.
    burst = serving_requests(n_requests, n_requests / 0.01, lights)
    r1 = run(burst)
    views_per_s = r1["views"] / r1["s"]
    # A served view of the burst against its plain launch (the host-read
    # path), bit for bit.
    served_vs_plain = None
    if r1["_kept"]:
        (cam_, w_, h_, rows), got = r1["_kept"][0]
        want = scene.render_view(cam_, w_, h_, rows, opts)
        served_vs_plain = {"queries": int(w_ * h_ * len(rows)),
                           "differ": int(sum(int((np.ascontiguousarray(a).view(np.uint32).ravel()
                                                  != np.ascontiguousarray(b).view(np.uint32).ravel()).sum())
                                             for a, b in zip(got, want)))}
    r1.pop("_kept")
    # Run 2: Poisson at half the drained rate, for about `seconds`.
    rate = 0.5 * views_per_s / 0.8
    n2 = max(20, int(rate * seconds))
    steady = serving_requests(n2, rate, lights, seed=0xC0_0008)
    r2 = run(steady)
    r2.pop("_kept")

    # The same views of run 2 as plain launches, one after another.
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    views = [p for _, k, p in steady if k == "view"]
    table = np.array(lights, np.float32, copy=True)
    maxpx = max(w * h for _, w, h, _ in views)
    maxq = max(w * h * len(i) for _, w, h, i in views)
    hits = torch.empty((maxpx, 4), dtype=torch.int32, device=dev)
    recv = torch.empty((maxpx, 8), dtype=torch.float32, device=dev)
    vis = torch.empty(maxq, dtype=torch.float32, device=dev)
    full = torch.empty(maxq, dtype=torch.int64, device=dev)
    touched = torch.empty(maxq, dtype=torch.int64, device=dev)
    cams = [torch.from_numpy(np.asarray(c, np.float32)).to(dev) for c, _, _, _ in views]
    lts = [torch.from_numpy(np.ascontiguousarray(table[i])).to(dev) for _, _, _, i in views]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    q = 0
    for k, (_, w, h, i) in enumerate(views):
        scene.primary_device(cams[k], w, h, lts[k], opts, hits, recv, stream)
        scene.view_query_device(recv, w, h, lts[k], opts, vis, full, touched, tile_w=PLAIN_TILE_W, stream=stream)
        q += w * h * len(i)
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    return {
        "workload": "boulder field: 10 boulder shapes x 999,987 instances + ground (100,002 triangles, 1.05M object "
                    "nodes), 32-light table; views 256^2..1920x1080 x 1-32 lights, 20% light edits",
        "instances": int(len(pl)), "host_build_s": round(t_build, 2), "scene_hbm_bytes": scene.hbm_bytes,
        "server": {"slots": 3, "max_views_per_launch": 16, "launch_queries": SERVED_LAUNCH_QUERIES,
                   "express_queries": SERVED_EXPRESS_QUERIES},
        "saturated": r1, "steady": dict(r2, offered_views_per_s=0.8 * rate),
        "plain_same_views": {"queries_per_s": q / t_plain, "s": t_plain},
        "served_over_plain": r1["queries_per_s"] / (q / t_plain),
        "burst_view_vs_its_plain_launch": served_vs_plain,
    }


def view_cache_form(device, boulders: int = 512, moves: int = 5) -> dict:
    """The view cache over the boulder field: pages (9 x 9 texels, mip 0,
    base texel 5 cm) of the `boulders` boulders nearest the camera and of the
    ground within 24 m of it (mip 1), lit by the sun and the three rect
    lights nearest the camera. Frame 0 bakes every page (one cone per texel
    and light); then the 16 nearest boulders, movable, shift 25 cm a frame
    and the stale passes re-bake what their moves reach. Times are wall
    clock around each frame (host planning, uploads and the readback of the
    invalid list included)."""
    import numpy as np

    from continuum_amd import _lib
    from continuum_amd.kernel import Flat, Scene, ViewCache
    from continuum_amd.scenes import boulder_field, boulder_field_surface

    t_build = time.perf_counter()
    tris9, counts, pl, lights, eye, _ = boulder_field()
    surf = boulder_field_surface()
    movers = list(range(1, 17))
    flat = Flat.build_instanced(tris9, counts, pl, object_prefilter=True, movable=movers)
    scene = Scene(device, flat)
    t_build = time.perf_counter() - t_build
    centre = lights[1:, 0:3] + 0.5 * (lights[1:, 4:7] + lights[1:, 8:11])
    nearest = 1 + np.argsort(np.linalg.norm(centre[:, :2] - eye[:2], axis=1))[:3]
    vlights = np.ascontiguousarray(lights[[0, *nearest]])
    opts = _lib.default_options(hierarchy=2, prefilter=1.0, object_cov=flat.object_cov)
    vc = ViewCache(scene, surf, vlights, opts, pages=32768, append_pages=8192, candidates=32768, base_texel=0.05)
    keys = []
    for m in range(1, boulders + 1):
        keys += vc.pages_of(m, int(pl[m]["geometry"]), 0)
    gx, gy = vc.grid(0, 0, 1)
    texel = 2100.0 / gx
    for py in range(gy // 8 + 1):
        y = -1050.0 + (py * 8 + 4) * texel
        if abs(y - eye[1]) > 24.0:
            continue
        for px in range(gx // 8 + 1):
            x = -1050.0 + (px * 8 + 4) * texel
            if math.hypot(x - eye[0], y - eye[1]) <= 24.0:
                keys.append(vc.key(0, 0, 1, px, py))
    slots = vc.request(keys)
    t0 = time.perf_counter()
    st = vc.frame()
    bake_s = time.perf_counter() - t0
    # The same bake again, counting the walks' work (a second cache, so the
    # first frame's timing stays uncounted).
    scene.set_count_work(True)
    scene.counters(reset=True)
    vc2 = ViewCache(scene, surf, vlights, opts, pages=32768, append_pages=8192, candidates=32768, base_texel=0.05)
    vc2.request(keys)
    vc2.frame()
    capped, work = scene.counters(reset=True)
    scene.set_count_work(False)
    del vc2
    cones = max(st["planned_jobs"], 1)
    out = {
        "workload": f"boulder field view cache: {boulders} nearest boulders (mip 0) + ground within 24 m (mip 1), "
                    "base texel 5 cm, sun + 3 rect lights",
        "pages": len(keys), "resident": int((slots != _lib.NONE).sum()), "host_build_s": round(t_build, 2),
        "bake": {"cones": st["planned_jobs"], "s": bake_s, "cones_per_s": st["planned_jobs"] / bake_s,
                 "per_cone": {"object_nodes": int(work[0]) / cones, "blas_nodes": int(work[1]) / cones,
                              "triangles_tested": int(work[2]) / cones, "meshes_entered": int(work[4]) / cones,
                              "splats": int(work[5]) / cones, "capped": int(capped[0])}},
    }
    frames = []
    for k in range(moves):
        for m in movers:
            t = pl[m]["translation"].copy()
            t[0] += 0.25 * (k + 1)
            scene.set_transform(m, pl[m]["rotation"], t)
        scene.apply_moves()
        vc.set_moved(scene.last_moves())
        vc.request(keys)
        t0 = time.perf_counter()
        st = vc.frame()
        dt = time.perf_counter() - t0
        frames.append(dict(ms=dt * 1e3, candidates=st["candidates"], stale_pages=st["stale_pages"],
                           rebaked_cones=st["appended_jobs"], evicted=st["evicted"], overflow=st["overflow"]))
    out["moves"] = {"movers": len(movers), "shift_m": 0.25, "frames": frames,
                    "ms_mean": float(np.mean([f["ms"] for f in frames]))}
    # The reference's invariant after the moves: every page shown equals a
    # fresh bake of the moved scene, bit for bit (`verify_cache`).
    differ, compared = vc.verify()
    out["moves"]["verify"] = {"values_compared": compared, "values_differ": differ}
    # A light edit: the nearest rect light moved 0.5 m, its channel of every
    # resident page baked afresh; then verified again.
    row = np.array(vc.lights[1], np.float32, copy=True)
    row[0:2] += np.float32(0.5)
    t0 = time.perf_counter()
    st = vc.set_lights(row[None], 1)
    dt = time.perf_counter() - t0
    differ, compared = vc.verify()
    out["light_edit"] = {"pages": st["pages"], "cones": st["cones"], "ms": dt * 1e3,
                         "verify": {"values_compared": compared, "values_differ": differ}}
    out["view"] = view_from_cache(device, scene, surf, vlights, opts, eye)
    return out


def sky_form(scene, flat, sc, recv, h_recv, check: bool = True, stride: int = 64) -> dict:
    """A uniform sky over the rock field's shading points (its 1280x720
    camera pass): six cube-face cones per point (the reference's
    `sky_irradiance`, the sky term of its per-pixel shading), faces summed
    in order. Timed with HIP events on the launch stream; every `stride`th
    point checked against the oracle bit for bit."""
    import numpy as np
    import torch

    from continuum_amd import _lib

    dev = recv.device
    px = recv.shape[0]
    out = torch.zeros(px, dtype=torch.float32, device=dev)
    faces = torch.zeros(6 * px, dtype=torch.float32, device=dev)
    full = torch.zeros(6 * px, dtype=torch.int64, device=dev)
    opts = _lib.default_options(hierarchy=2, prefilter=sc.prefilter)
    stream = torch.cuda.current_stream()
    scene.sky_device(recv, opts, out, faces, full, stream=stream)
    torch.cuda.synchronize()
    ms = []
    for _ in range(5):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        scene.sky_device(recv, opts, out, faces, full, stream=stream)
        e1.record(stream)
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    t = float(np.median(ms))
    # the walks' work per face cone, from one more launch with the counters on
    scene.counters(reset=True)
    scene.set_count_work(True)
    scene.sky_device(recv, opts, out, faces, full, stream=stream)
    torch.cuda.synchronize()
    _, work = scene.counters(reset=True)
    scene.set_count_work(False)
    lit = int((h_recv[:, 7] >= 0).sum())
    res = {"workload": f"uniform sky over the rock field's {sc.width}x{sc.height} shading points, six cube-face "
                       "cones each (independent occluders, horizon cull, 1-bin prefilter)",
           "points": px, "traced_points": lit, "face_cones": 6 * px, "ms": t, "face_cones_per_s": 6 * px / t * 1e3,
           "mean_sky_irradiance": float(out.mean().item()),
           "per_face_cone": {"object_nodes": int(work[0]) / (6 * px), "blas_nodes": int(work[1]) / (6 * px),
                             "triangles_tested": int(work[2]) / (6 * px), "triangles_drawn": int(work[3]) / (6 * px),
                             "meshes_entered": int(work[4]) / (6 * px), "splats": int(work[5]) / (6 * px)}}
    if check:
        from oracle import oracle as O

        sample = h_recv[::stride]
        t0 = time.perf_counter()
        rv, rf, rfull = O.sky(flat, sample, O.opts_array(hierarchy=2, prefilter=sc.prefilter))
        t_cpu = time.perf_counter() - t0
        gv = out.cpu().numpy()[::stride]
        gf = faces.cpu().numpy().reshape(6, px)[:, ::stride]
        gfull = full.cpu().numpy().view(np.uint64).reshape(6, px)[:, ::stride]
        res["parity"] = {"sample_points": len(sample),
                         "sum_bits_differ": int((gv.view(np.uint32) != rv.view(np.uint32)).sum()),
                         "face_bits_differ": int((gf.view(np.uint32) != rf.view(np.uint32)).sum()),
                         "full_masks_differ": int((gfull != rfull).sum())}
        res["cpu_oracle_face_cones_per_s"] = 6 * len(sample) / t_cpu
    return res


def lightmap_form(device, size: float = 2.0, texel: float = 0.04, boxes: int = 8, frames: int = 10) -> dict:
    """The lightmap cache (the reference's `LiveRenderer` lighting: block
    planes, `mark_blocks`, `bake_static`, `rebake`): an 8 x 8 x 6 m room of
    large quads with a pillar and eight dynamic boxes under a ceiling rect,
    lightmap texels of 4 cm on the static surfaces (8 x 8 texel blocks).
    The cache is built (static coverage of every lit block and the first
    bake), then two boxes move 4 cm a frame and each update re-bakes the
    blocks their swept boxes may shadow. Times are wall clock around each
    call (host uploads included). At the end every texel is compared with a
    fresh query of the moved scene, bit for bit."""
    import numpy as np

    from continuum_amd import _lib
    from continuum_amd.kernel import Flat, LiveLightmap, Scene
    from continuum_amd.lightmap import Lightmap
    from continuum_amd.scenes import lit_room, meshes_to_arrays

    sc = lit_room(boxes=boxes, size=size)
    dyn = sc.notes["dynamic"]
    t0 = time.perf_counter()
    lm = Lightmap.from_meshes(sc.meshes, texel, lambda m: not dyn[m])
    t_host = time.perf_counter() - t0
    tris9, counts = meshes_to_arrays(sc.meshes)
    scene = Scene(device, Flat.build_arrays(tris9, counts))
    light = sc.lights[0]
    t0 = time.perf_counter()
    cache = LiveLightmap(scene, lm.texels, lm.blocks, light, dyn)
    t_build = time.perf_counter() - t0
    opts = _lib.default_options(hierarchy=2)
    t0 = time.perf_counter()
    scene.query(lm.texels, light[None], opts)
    t_full = time.perf_counter() - t0
    off = [0] + [int(x) for x in np.cumsum(counts)]
    moved = tris9.copy()
    movers = [i for i, d in enumerate(dyn) if d][:2]
    ms, rebaked = [], []
    for k in range(frames):
        changes = []
        for j, m in enumerate(movers):
            new = moved[off[m]:off[m + 1]].copy()
            new[:, 0::3] += np.float32(0.04 if j == 0 else -0.03)
            new[:, 1::3] += np.float32(0.02)
            moved[off[m]:off[m + 1]] = new
            changes.append((m, new))
        t0 = time.perf_counter()
        cache.update_meshes(changes)
        _, marked = cache.read()
        ms.append((time.perf_counter() - t0) * 1e3)
        rebaked.append(len(marked))
    got = cache.read_texels()
    fresh, _, _, _ = Scene(device, Flat.build_arrays(moved, counts)).query(lm.texels, light[None], opts)
    return {"workload": f"lit room {4 * size:g} x {4 * size:g} x {3 * size:g} m, {boxes} dynamic boxes, ceiling rect, "
                        f"{texel * 100:g} cm texels on the static surfaces; two boxes move each frame",
            "texels": int(len(lm.texels)), "blocks": int(len(lm.blocks)), "host_lightmap_s": round(t_host, 1),
            "build_and_first_bake_ms": t_build * 1e3, "full_query_ms": t_full * 1e3,
            "update_ms_mean": float(np.mean(ms[1:])), "update_ms": [round(x, 2) for x in ms],
            "blocks_rebaked_mean": float(np.mean(rebaked)),
            "blocks_rebaked_share": float(np.mean(rebaked)) / len(lm.blocks),
            "texels_differ_from_a_fresh_query": int((got.view(np.uint32) != fresh.view(np.uint32)).sum())}


def view_from_cache(device, scene, surf, lights, opts, eye, width: int = 1280, height: int = 720,
                    budget: int = 2_000_000, max_frames: int = 40) -> dict:
    """A 1280x720 view shaded from a fresh view cache (the reference's
    renderer with a cache): each frame shades from the pages resident, plans
    the pages its sample pixels asked for (at most `budget` cones) and bakes
    them. Sample pixels move through each request tile frame by frame, so
    finer pages keep arriving for a while after every pixel has one; pages
    off their chart's grid are asked for and never made. Reported: the frame
    from which every pixel reads a resident page, the frames, the shading
    pass at the end, and the plain per-view launch (camera pass and a cone
    per pixel and light) for the same view and lights."""
    import numpy as np
    import torch

    from continuum_amd.kernel import ViewCache
    from continuum_amd.scenes import camera_uniform

    dev = torch.device("cuda:0")
    cam = camera_uniform(eye, [eye[0] + 10.0, eye[1] + 10.0, 0.4], 60.0, width, height)
    px = width * height
    hits = torch.zeros((px, 4), dtype=torch.int32, device=dev)
    recv = torch.zeros((px, 8), dtype=torch.float32, device=dev)
    lt = torch.from_numpy(np.ascontiguousarray(lights)).to(dev)
    scene.primary_device(torch.from_numpy(cam).to(dev), width, height, lt, opts, hits, recv)
    torch.cuda.synchronize()
    vc = ViewCache(scene, surf, lights, opts, pages=1 << 17, append_pages=256, candidates=4096, base_texel=0.05)
    out = torch.zeros((len(lights), height, width), dtype=torch.float32, device=dev)
    page = torch.zeros((height, width), dtype=torch.int32, device=dev)
    frames, settled = [], None
    for k in range(max_frames):
        t0 = time.perf_counter()
        req, fell = vc.shade_device(cam, width, height, hits, out, page)
        t1 = time.perf_counter()
        st = vc.plan(budget)
        tp = time.perf_counter()
        vc.frame()
        t2 = time.perf_counter()
        frames.append(dict(frame=k, shade_ms=(t1 - t0) * 1e3, plan_ms=(tp - t1) * 1e3, bake_ms=(t2 - tp) * 1e3,
                           requests=req, pixels_without_page=fell,
                           cones=st["cones"], pages=st["planned"]))
        if fell == 0 and settled is None:
            settled = k
    shade = []
    for _ in range(5):
        t0 = time.perf_counter()
        vc.shade_device(cam, width, height, hits, out, page)
        shade.append((time.perf_counter() - t0) * 1e3)
    vis = torch.zeros(len(lights) * px, dtype=torch.float32, device=dev)
    full = torch.zeros(len(lights) * px, dtype=torch.int64, device=dev)
    touched = torch.zeros(len(lights) * px, dtype=torch.int64, device=dev)
    plain = []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        scene.primary_device(torch.from_numpy(cam).to(dev), width, height, lt, opts, hits, recv)
        scene.view_query_device(recv, width, height, lt, opts, vis, full, touched, tile_w=PLAIN_TILE_W)
        torch.cuda.synchronize()
        plain.append((time.perf_counter() - t0) * 1e3)
    return {"width": width, "height": height, "budget_cones_per_frame": budget, "every_pixel_from_frame": settled,
            "frames": frames[:8] + frames[8::8] + ([frames[-1]] if (len(frames) - 9) % 8 else []),
            "cones_baked": int(sum(f["cones"] for f in frames)),
            "bake_ms_total": float(sum(f["bake_ms"] for f in frames)),
            "resident_pages": vc.state()["slot_extent"], "shade_ms_settled": float(np.median(shade)),
            "plain_view_ms": float(np.median(plain))}


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=1280)
    ap.add_argument("--height", type=int, default=720)
    ap.add_argument("--tile-w", type=int, default=PLAIN_TILE_W, help="pixels across a wave's tile (tile_w x 64/tile_w)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the courtyard's plain/batched forms")
    ap.add_argument("--no-served", action="store_true", help="skip the served form over the boulder field")
    ap.add_argument("--no-vcache", action="store_true", help="skip the view cache over the boulder field")
    ap.add_argument("--no-sky", action="store_true", help="skip the sky light over the rock field")
    ap.add_argument("--no-lightmap", action="store_true", help="skip the lightmap cache over the lit room")
    ap.add_argument("--only-form", choices=["courtyard", "served", "view_cache", "lightmap"],
                    help="run one form alone and print its JSON (for profiling it); no headline line")
    ap.add_argument("--served-requests", type=int, default=64, help="requests in the saturated served run")
    ap.add_argument("--served-seconds", type=float, default=10.0, help="length of the steady served run")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU: this path runs on a single MI355X and has no multi-GPU form "
              f"(asked for {max(args.gpus, world)} GPUs).", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    from continuum_amd import _lib
    from continuum_amd.kernel import Device, Flat, Scene
    from continuum_amd.scenes import meshes_to_arrays, rock_field

    if args.only_form:
        device = Device(0)
        if args.only_form == "courtyard":
            out = courtyard_forms(device, steps=5)
        elif args.only_form == "served":
            out = served_form(device, n_requests=args.served_requests, seconds=args.served_seconds)
        elif args.only_form == "lightmap":
            out = lightmap_form(device)
        else:
            out = view_cache_form(device)
        print(json.dumps({"form": args.only_form, **out}))
        return 0

    t0 = time.perf_counter()
    sc = rock_field(width=args.width, height=args.height)
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    t_build = time.perf_counter() - t0
    dev = torch.device("cuda:0")
    device = Device(0)
    scene = Scene(device, flat)
    opts = _lib.default_options(hierarchy=2, prefilter=sc.prefilter)

    W, H = sc.width, sc.height
    px = W * H
    nl = len(sc.lights)
    Q = px * nl
    cam = torch.from_numpy(sc.camera).to(dev)
    lights = torch.from_numpy(np.ascontiguousarray(sc.lights)).to(dev)
    hits = torch.empty((px, 4), dtype=torch.int32, device=dev)
    recv = torch.empty((px, 8), dtype=torch.float32, device=dev)
    vis = torch.empty(Q, dtype=torch.float32, device=dev)
    full = torch.empty(Q, dtype=torch.int64, device=dev)
    touched = torch.empty(Q, dtype=torch.int64, device=dev)

    # A copy of known size (256 MiB read, 256 MiB written), for calibrating
    # HBM byte counters when this command runs under rocprofv3 --pmc.
    ca = torch.ones(64 << 20, dtype=torch.float32, device=dev)
    cb = torch.empty_like(ca)
    torch.neg(ca, out=cb)
    del ca, cb

    stream = torch.cuda.current_stream()

    def step(e0=None, e1=None, e2=None):
        if e0 is not None:
            e0.record(stream)
        scene.primary_device(cam, W, H, lights, opts, hits, recv, stream)
        if e1 is not None:
            e1.record(stream)
        scene.view_query_device(recv, W, H, lights, opts, vis, full, touched, tile_w=args.tile_w, stream=stream)
        if e2 is not None:
            e2.record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    scene.counters(reset=True)

    K = args.steps
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(K)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for k in range(K):
        step(*ev[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start

    lat = np.array([ev[k][0].elapsed_time(ev[k][2]) for k in range(K)])
    cone_ms = np.array([ev[k][1].elapsed_time(ev[k][2]) for k in range(K)])
    prim_ms = np.array([ev[k][0].elapsed_time(ev[k][1]) for k in range(K)])
    counters, _ = scene.counters(reset=True)

    # One more step with the work counters on (untimed), for the roofline's bytes.
    scene.set_count_work(True)
    step()
    torch.cuda.synchronize()
    _, work = scene.counters(reset=True)
    scene.set_count_work(False)

    h_recv = recv.cpu().numpy()
    lit = int((h_recv[:, 7] >= 0).sum())
    traced_q = lit * nl
    algo = algorithmic_bytes(work, traced_q, lit, nl)
    cone_avg_s = float(cone_ms.mean()) / 1e3
    achieved = algo["total"] / cone_avg_s / 1e9
    traffic = read_traffic(sc.name)

    result = {
        "metric": METRIC,
        "value": K * Q / wall,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": wall / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded rock field; geometry generated, no assets)",
        "config": {
            "workload": "rock field, plain per-view launch (camera pass + one cone per pixel and light)",
            "triangles": int(sc.n_tris),
            "objects": len(sc.meshes),
            "shading_points": px,
            "lights": nl,
            "queries_per_step": Q,
            "traced_queries_per_step": traced_q,
            "hierarchy": "triangles",
            "prefilter_bins": sc.prefilter,
            "scene_hbm_bytes": scene.hbm_bytes,
            "host_build_s": round(t_build, 3),
        },
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "camera_pass_mean": float(prim_ms.mean()), "cone_pass_mean": float(cone_ms.mean())},
        "roofline": {
            "bound": "hbm",
            "kernel": "cone_items_kernel",
            "achieved": achieved,
            "peak": PEAK_HBM_GBPS,
            "unit": "GB/s",
            "frac": achieved / PEAK_HBM_GBPS,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": algo["total"],
            "kernel_ms": float(cone_ms.mean()),
            "work": {"object_nodes": int(work[0]), "blas_nodes": int(work[1]), "triangles_tested": int(work[2]),
                     "triangles_drawn": int(work[3]), "meshes_entered": int(work[4]), "splats": int(work[5])},
        },
        "diagnostics": {"capped_walks": int(counters[0]), "deeper_than_48": int(counters[1]),
                        "truncated": int(counters[2])},
    }

    if not args.no_cpu_baseline:
        from oracle import oracle as O

        threads = O.threads_default()
        o_opts = O.opts_array(hierarchy=2, prefilter=sc.prefilter)
        h_vis = vis.cpu().numpy().reshape(nl, px)
        h_full = full.cpu().numpy().view(np.uint64).reshape(nl, px)
        h_touched = touched.cpu().numpy().view(np.uint64).reshape(nl, px)
        # calibrate on 1/1024 of the pixels, then a sample of about cpu_seconds
        probe = h_recv[::1024]
        t = time.perf_counter()
        O.irradiance(flat, probe, sc.lights, o_opts, threads=threads)
        rate = len(probe) * nl / max(time.perf_counter() - t, 1e-6)
        stride = max(1, int(np.ceil(Q / max(rate * args.cpu_seconds, 1.0))))
        sample = h_recv[::stride]
        t = time.perf_counter()
        rv, rf, rt, _, capped, _ = O.irradiance(flat, sample, sc.lights, o_opts, threads=threads)
        t_cpu = time.perf_counter() - t
        sq = len(sample) * nl
        gv = h_vis[:, ::stride].reshape(-1)
        gf = h_full[:, ::stride].reshape(-1)
        gt = h_touched[:, ::stride].reshape(-1)
        ok = capped == 0
        result["cpu_baseline"] = {
            "value": sq / t_cpu,
            "unit": "queries/s",
            "cores": threads,
            "kind": "port",
            "sample": f"every {stride}th shading point of the {W}x{H} view x {nl} lights = {sq} queries "
                      f"({t_cpu:.1f} s on {threads} threads; cones only, receivers from the GPU camera pass)",
        }
        # camera pass: every pixel against the oracle
        rh, _, _ = O.primary(flat, sc.camera, W, H, sc.lights, threads=threads)
        result["parity"] = {
            "sample_queries": sq,
            "visible_bits_differ": int((gv.view(np.uint32)[ok] != rv.view(np.uint32)[ok]).sum()),
            "full_masks_differ": int((gf[ok] != rf[ok]).sum()),
            "touched_masks_differ": int((gt[ok] != rt[ok]).sum()),
            "capped_skipped": int((~ok).sum()),
            "camera_hits_differ": int((hits.cpu().numpy().view(np.uint32) != rh).any(axis=1).sum()),
        }

    if not args.no_sky:
        result.setdefault("forms", {})["sky"] = sky_form(scene, flat, sc, recv, h_recv, check=not args.no_cpu_baseline)
    if not args.no_forms:
        del scene
        result.setdefault("forms", {}).update(courtyard_forms(device, steps=5))
    if not args.no_served:
        result.setdefault("forms", {})["served"] = served_form(device, n_requests=args.served_requests,
                                                               seconds=args.served_seconds)
    if not args.no_vcache:
        result.setdefault("forms", {})["view_cache"] = view_cache_form(device)
    if not args.no_lightmap:
        result.setdefault("forms", {})["lightmap"] = lightmap_form(device)

    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
