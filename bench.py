#!/usr/bin/env python3
"""Cone-visibility throughput on one MI355X.

A step is one frame of the 3M-triangle courtyard as the plain per-view
launch renders it: the curtains (1.2M triangles) move by a seeded sine field
and their triangle hierarchies are refit on the GPU, then the camera pass
(1280 x 720 shading points) and one cone query per shading point and light
(4 lights: a 0.53 degree sun disk and three rectangles), on inputs already
resident in HBM. Prints one JSON line: queries/s over the timed steps,
p50/p99 frame latency, the cone kernel's roofline, the oracle's throughput
on the host's cores for a bounded sample of the same queries (and one
thread's ns per query), and a bit-for-bit comparison of that sample. Under
`forms`: the courtyard's 64-view batch against 64 sequential plain
launches, the rock field's plain frame and sky, the served boulder field,
the view cache and the lightmap cache.
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


def note(msg: str) -> None:
    """A progress line on stderr as each leg ends (stdout keeps the one JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)
PEAK_HBM_GBPS = 8000.0
PROFILE_DIR = ROOT / "profiles"


# The plain launch's tile: 16 x 4 shading points a wave on the rock field
# (a sweep over 4, 8, 16, 32 and 64 puts it first there), 8 x 8 on the
# courtyard (62.2 ms a frame against 66.8 ms at 16 x 4, 72.2 at 4 x 16 and
# 86.5 at 32 x 2).
PLAIN_TILE_W = 16
COURTYARD_TILE_W = 8


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


def posed_oracle_flat(flat, counts, pose):
    """The oracle's arrays after a refit to `pose`: its refit of every
    triangle hierarchy and its object hierarchy over the posed meshes."""
    import types

    from oracle import oracle as O

    rblas, rcov = O.refit(flat, pose)
    nodes, items = O.mesh_hierarchy(pose.reshape(-1, 3, 4)[:, :, :3].reshape(-1, 9), counts)
    return types.SimpleNamespace(tris=pose, ranges=flat.ranges, blas=rblas, blas_items=flat.blas_items,
                                 blas_cov=rcov, nodes=nodes, items=items, xforms=flat.xforms)


def plain_frames(scene, sc, steps: int, warmup: int, tile_w: int, poses=None) -> dict:
    """The plain per-view launch, `steps` frames after `warmup`: (the
    refit to frame f's pose, when `poses` is given,) the camera pass and one
    cone query per pixel and light, all on torch's current stream. HIP
    events bracket each pass of each timed frame; the wall clock brackets
    the frames (the refit reads the meshes' bounds back to the host once a
    frame, so frames do not overlap). Before the warmup, the first launch of
    the view (natural item order, no times yet) is timed on its own. After
    the timed frames: one frame with the walk counters on (the roofline's
    bytes), and one with per-query flags (the capped walks)."""
    import numpy as np
    import torch

    from continuum_amd import _lib

    dev = torch.device("cuda:0")
    opts = _lib.default_options(hierarchy=2, prefilter=sc.prefilter)
    W, H, nl = sc.width, sc.height, len(sc.lights)
    px, Q = W * H, W * H * nl
    cam = torch.from_numpy(sc.camera).to(dev)
    lights = torch.from_numpy(np.ascontiguousarray(sc.lights)).to(dev)
    hits = torch.empty((px, 4), dtype=torch.int32, device=dev)
    recv = torch.empty((px, 8), dtype=torch.float32, device=dev)
    vis = torch.empty(Q, dtype=torch.float32, device=dev)
    full = torch.empty(Q, dtype=torch.int64, device=dev)
    touched = torch.empty(Q, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()

    def frame(f, ev=None):
        if ev is not None:
            ev[0].record(stream)
        if poses is not None:
            scene.refit_device(poses[f % len(poses)], stream)
        if ev is not None:
            ev[1].record(stream)
        scene.primary_device(cam, W, H, lights, opts, hits, recv, stream)
        if ev is not None:
            ev[2].record(stream)
        scene.view_query_device(recv, W, H, lights, opts, vis, full, touched, tile_w=tile_w, stream=stream)
        if ev is not None:
            ev[3].record(stream)

    first = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    frame(0, first)
    torch.cuda.synchronize()
    first_cone_ms = first[2].elapsed_time(first[3])
    for k in range(warmup):
        frame(1 + k)
    torch.cuda.synchronize()
    scene.counters(reset=True)
    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(steps):
        frame(1 + warmup + k, ev[k])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    lat = np.array([e[0].elapsed_time(e[3]) for e in ev])
    refit_ms = np.array([e[0].elapsed_time(e[1]) for e in ev])
    prim_ms = np.array([e[1].elapsed_time(e[2]) for e in ev])
    cone_ms = np.array([e[2].elapsed_time(e[3]) for e in ev])
    counters, _ = scene.counters(reset=True)
    last = 1 + warmup + steps - 1
    # untimed: the same (last) frame with the walk counters on, then with flags
    scene.set_count_work(True)
    frame(last)
    torch.cuda.synchronize()
    _, work = scene.counters(reset=True)
    scene.set_count_work(False)
    flags = torch.zeros(Q, dtype=torch.uint8, device=dev)
    scene.set_query_flags(flags)
    frame(last)
    torch.cuda.synchronize()
    scene.set_query_flags(None)
    h_flags = flags.cpu().numpy()
    h_recv = recv.cpu().numpy()
    lit = int((h_recv[:, 7] >= 0).sum())
    return {
        "W": W, "H": H, "nl": nl, "Q": Q, "lit": lit, "wall": wall, "steps": steps, "last_frame": last,
        "lat": lat, "refit_ms": refit_ms, "prim_ms": prim_ms, "cone_ms": cone_ms, "first_cone_ms": first_cone_ms,
        "counters": counters, "work": work, "flags": h_flags, "h_recv": h_recv,
        "h_hits": hits.cpu().numpy(), "h_vis": vis.cpu().numpy(), "h_full": full.cpu().numpy().view(np.uint64),
        "h_touched": touched.cpu().numpy().view(np.uint64), "recv": recv,
    }


def roofline_of(m, workload: str) -> dict:
    """The cone kernel's roofline from a plain_frames measurement: the
    algorithmic bytes of one launch (the counted walk's work, priced per
    unit) over its mean time in the timed frames."""
    traced_q = m["lit"] * m["nl"]
    algo = algorithmic_bytes(m["work"], traced_q, m["lit"], m["nl"])
    cone_s = float(m["cone_ms"].mean()) / 1e3
    achieved = algo["total"] / cone_s / 1e9
    w = m["work"]
    return {"bound": "hbm", "kernel": "cone_items_kernel", "achieved": achieved, "peak": PEAK_HBM_GBPS,
            "unit": "GB/s", "frac": achieved / PEAK_HBM_GBPS, "traffic": read_traffic(workload),
            "algorithmic_bytes_per_launch": algo["total"], "kernel_ms": float(m["cone_ms"].mean()),
            "work": {"object_nodes": int(w[0]), "blas_nodes": int(w[1]), "triangles_tested": int(w[2]),
                     "triangles_drawn": int(w[3]), "meshes_entered": int(w[4]), "splats": int(w[5])},
            "per_query": {k: int(x) / max(traced_q, 1) for k, x in
                          zip(("object_nodes", "blas_nodes", "triangles_tested", "triangles_drawn",
                               "meshes_entered", "splats"), w)}}


def oracle_check(m, oflat, sc, cpu_seconds: float, single_seconds: float = 0.0, camera: bool = True) -> dict:
    """The oracle on the host's cores over a bounded sample of the measured
    frame's queries (every stride-th shading point, all lights; receivers
    from the GPU camera pass), its rate, and the sample compared bit for bit
    with the GPU's results (queries capped on either side excluded and
    counted). With `single_seconds`, also one thread's ns per query on a
    smaller sample. With `camera`, every pixel's camera hit is compared."""
    import numpy as np

    from oracle import oracle as O

    threads = O.threads_default()
    o_opts = O.opts_array(hierarchy=2, prefilter=sc.prefilter)
    W, H, nl, Q = m["W"], m["H"], m["nl"], m["Q"]
    px = W * H
    h_recv = m["h_recv"]
    probe = h_recv[::997]
    t = time.perf_counter()
    O.irradiance(oflat, probe, sc.lights, o_opts, threads=threads)
    rate = len(probe) * nl / max(time.perf_counter() - t, 1e-6)
    stride = max(1, int(np.ceil(Q / max(rate * cpu_seconds, 1.0))))
    pix = np.arange(0, px, stride)
    t = time.perf_counter()
    rv, rf, rt, _, capped, _ = O.irradiance(oflat, h_recv[pix], sc.lights, o_opts, threads=threads)
    t_cpu = time.perf_counter() - t
    sq = len(pix) * nl
    q = (np.arange(nl)[:, None] * px + pix[None, :]).reshape(-1)
    gv, gf, gt = m["h_vis"][q], m["h_full"][q], m["h_touched"][q]
    skip = (capped != 0) | ((m["flags"][q] & 1) != 0)
    ok = ~skip
    out = {"cpu_baseline": {
        "value": sq / t_cpu, "unit": "queries/s", "cores": threads, "kind": "port",
        "sample": f"every {stride}th shading point of the {W}x{H} view x {nl} lights = {sq} queries "
                  f"({t_cpu:.1f} s on {threads} threads; cones only, receivers from the GPU camera pass)"}}
    if single_seconds > 0:
        s_stride = max(1, int(np.ceil(Q / max(rate / threads * single_seconds, 1.0))))
        s_pix = np.arange(0, px, s_stride)
        t = time.perf_counter()
        O.irradiance(oflat, h_recv[s_pix], sc.lights, o_opts, threads=1)
        t1 = time.perf_counter() - t
        out["cpu_baseline"]["single_thread_ns_per_query"] = t1 / (len(s_pix) * nl) * 1e9
        out["cpu_baseline"]["single_thread_sample"] = f"every {s_stride}th shading point x {nl} lights"
    par = {"sample_queries": sq,
           "visible_bits_differ": int((gv.view(np.uint32)[ok] != rv.view(np.uint32)[ok]).sum()),
           "full_masks_differ": int((gf[ok] != rf[ok]).sum()),
           "touched_masks_differ": int((gt[ok] != rt[ok]).sum()),
           "capped_skipped": int(skip.sum()),
           "capped_gpu_only": int(((m["flags"][q] & 1) != 0).sum() - ((capped != 0) & ((m["flags"][q] & 1) != 0)).sum()),
           "capped_oracle_only": int(((capped != 0) & ((m["flags"][q] & 1) == 0)).sum())}
    if camera:
        rh, _, _ = O.primary(oflat, sc.camera, W, H, sc.lights, threads=threads)
        par["camera_hits_differ"] = int((m["h_hits"].view(np.uint32) != rh).any(axis=1).sum())
    out["parity"] = par
    return out


def courtyard_batch_form(device, scene, opts) -> dict:
    """64 views of the courtyard (sizes from {256^2, 640x360, 1280x720,
    1920x1080}, 1-16 lights each): as 64 sequential plain launches (camera
    pass + cone pass per view), and as one batch (autotuned on every eighth
    view, captured as a hipGraph and replayed). The batch's last view is
    compared with its plain launch bit for bit."""
    import numpy as np
    import torch

    from continuum_amd.kernel import Batch
    from continuum_amd.workloads import courtyard_views, view_queries

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    views, vlights = courtyard_views()
    total_q = view_queries(views)
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
    # The batch's last view against its own plain launch, bit for bit: both
    # once more (untimed) with per-query flags, the batch uncaptured (the
    # graph holds no flags). Walks capped on either side are excluded: a
    # capped walk keeps the nodes its packet reached, and the autotuned tile
    # can group other pixels into a wave than the plain launch's 16 x 4.
    _, lw, lh, lf, ln = views[-1]
    lq = lw * lh * ln
    p_flags = torch.zeros(lq, dtype=torch.uint8, device=dev)
    scene.set_query_flags(p_flags)
    scene.primary_device(v_cams[-1], lw, lh, v_lights[-1], opts, s_hits, s_recv, stream)
    scene.view_query_device(s_recv, lw, lh, v_lights[-1], opts, s_vis, s_full, s_touched, tile_w=PLAIN_TILE_W,
                            stream=stream)
    b_flags = torch.zeros(total_q, dtype=torch.uint8, device=dev)
    scene.set_query_flags(b_flags)
    batch.run(opts, best)
    torch.cuda.synchronize()
    scene.set_query_flags(None)
    b_hits, b_vis, b_full, b_touched = batch.read(len(views) - 1)
    capped = ((p_flags.cpu().numpy() | b_flags.cpu().numpy()[total_q - lq:]) & 1) != 0
    ok = ~capped
    last_view_differ = {
        "hits": int((s_hits[:lw * lh].cpu().numpy().view(np.uint32) != b_hits).any(axis=1).sum()),
        "visible": int((s_vis[:lq].cpu().numpy().view(np.uint32) != b_vis.view(np.uint32))[ok].sum()),
        "full": int((s_full[:lq].cpu().numpy().view(np.uint64) != b_full)[ok].sum()),
        "touched": int((s_touched[:lq].cpu().numpy().view(np.uint64) != b_touched)[ok].sum()),
        "capped_excluded": int(capped.sum()),
        "queries": int(lq)}
    return {
        "workload": "64 views over the courtyard, sizes from {256^2, 640x360, 1280x720, 1920x1080}, 1-16 lights each",
        "queries": total_q,
        "sequential_plain": {"queries_per_s": total_q / t_seq, "s": t_seq},
        "batched_graph": {"queries_per_s": total_q / t_batch, "s": t_batch,
                          "tuning": {"tile_w": best.tile_w, "persistent_blocks": best.persistent_blocks,
                                     "waves_per_simd": best.waves_per_simd, "ray_stack": best.ray_stack}},
        "batched_over_sequential": t_seq / t_batch,
        "last_view_batched_vs_plain_differ": last_view_differ,
    }


# Served form: cone launches cut into pieces of at most 4M queries, and a
# third slot that takes views of at most 1M queries (a 1280x720 view with one
# light) ahead of larger ones.
SERVED_LAUNCH_QUERIES = 1 << 22
SERVED_EXPRESS_QUERIES = 1 << 20


def serve_stream(scene, lights, opts, reqs) -> dict:
    """Offers `reqs` (serving_requests' list) to a fresh server over `scene` at
    their arrival times and collects each view as it finishes: throughput,
    latency (overall and by view size), queueing; keeps the first view of at
    most 1M queries with its results, for a check against its plain launch."""
    import threading

    import numpy as np

    from continuum_amd.kernel import Server

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

    # Run 1: everything offered within the first 10 ms.
    burst = serving_requests(n_requests, n_requests / 0.01, lights)
    r1 = serve_stream(scene, lights, opts, burst)
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
    r2 = serve_stream(scene, lights, opts, steady)
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


def served_sweep_form(device, rates, seconds: float) -> dict:
    """The served boulder field under Poisson streams at increasing offered
    view rates (views are 80% of the requests, light edits the rest; each
    stream about `seconds` long, with its own seed): per rate the views
    done per second, queries/s, p50/p99 latency, and the drain (last
    completion after the last arrival). A rate is stable when the server
    keeps up with the stream: the drain is at most 1 s, or 5% of the
    stream's span."""
    import numpy as np

    from continuum_amd import _lib
    from continuum_amd.kernel import Flat, Scene
    from continuum_amd.scenes import boulder_field
    from continuum_amd.workloads import serving_requests

    tris9, counts, pl, lights, _, _ = boulder_field()
    flat = Flat.build_instanced(tris9, counts, pl, object_prefilter=True)
    scene = Scene(device, flat)
    opts = _lib.default_options(hierarchy=2, prefilter=1.0, object_cov=flat.object_cov)
    out = []
    for k, rate in enumerate(rates):
        req_rate = rate / 0.8
        n = max(20, int(req_rate * seconds))
        reqs = serving_requests(n, req_rate, lights, seed=0xC0_0100 + k)
        last_arrival = reqs[-1][0]
        r = serve_stream(scene, lights, opts, reqs)
        r.pop("_kept")
        done_rate = r["views"] / max(r["s"], 1e-9)
        drain_s = max(r["s"] - last_arrival, 0.0)
        row = {"offered_views_per_s": rate, "views": r["views"], "views_done_per_s": done_rate,
               "queries_per_s": r["queries_per_s"], "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
               "queued_p50_ms": r["queued_p50_ms"], "drain_s": drain_s, "by_size": r["by_size"]}
        row["stable"] = bool(drain_s <= max(1.0, 0.05 * last_arrival))
        out.append(row)
        note(f"served at {rate:g} views/s: {done_rate:.2f} done/s, p50 {r['p50_ms']:.0f} ms, "
             f"p99 {r['p99_ms']:.0f} ms, drain {drain_s:.1f} s")
    stable = [r["offered_views_per_s"] for r in out if r["stable"]]
    mean_q = 0.5 * (256 * 256 + 1920 * 1080) * 16.5  # a C5 view's mean queries: pixels x lights
    return {"workload": "boulder field, Poisson view/edit stream (80% views of 256^2..1920x1080 x 1-32 lights)",
            "rates": out, "highest_stable_views_per_s": max(stable) if stable else None,
            "target": {"requests_per_s": 2000, "views_per_s": 1600,
                       "queries_per_s_needed": 1600 * mean_q}}


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
    out["moving_lights"] = moving_lights_leg(scene, surf, vlights, opts, keys)
    out["view"] = view_from_cache(device, scene, surf, vlights, opts, eye)
    return out


def moving_lights_leg(scene, surf, vlights, opts, keys, frames: int = 6) -> dict:
    """The same pages in caches whose sun (a distant square) and nearest
    rect light move through the error ledger, both on their lattices, at two
    speeds: the sun's corner turning 1.2e-6 rad a frame about the vertical
    axis (a real sun at 60 frames/s) with the rect light moving 1 mm a
    frame, and 1e-4 rad with 2 cm. After the first bake, each frame the
    ledger re-bakes only what its bounds say may be off by more than
    epsilon (mixed texels, or whole pages past their slack, whose fresh
    texels trace guard cones). Then the reference's `verify_ledger` check
    for both lights."""
    import numpy as np

    from continuum_amd.kernel import ViewCache

    def run(turn: float, step: float) -> dict:
        vc = ViewCache(scene, surf, vlights, opts, pages=32768, append_pages=32768, candidates=16, base_texel=0.05)
        vc.set_moving_lights([0, 1], 0.01)
        vc.request(keys)
        vc.frame()
        sun0 = np.array(vc.lights[0], np.float32, copy=True)
        rect0 = np.array(vc.lights[1], np.float32, copy=True)
        rows = []
        for k in range(1, frames + 1):
            rect = rect0.copy()
            rect[0] += np.float32(step * k)
            c, s_ = math.cos(turn * k), math.sin(turn * k)
            sun = sun0.copy()
            sun[0] = np.float32(c * float(sun0[0]) - s_ * float(sun0[1]))
            sun[1] = np.float32(s_ * float(sun0[0]) + c * float(sun0[1]))
            vc.move_light(1, rect)
            vc.move_light(0, sun)
            vc.request(keys)
            t0 = time.perf_counter()
            vc.frame()
            dt = time.perf_counter() - t0
            rows.append(dict(ms=dt * 1e3, **vc.ledger_stats()))
        checks = {str(light): vc.verify_ledger(light) for light in (0, 1)}
        del vc
        return {"sun_turn_rad_per_frame": turn, "rect_step_m_per_frame": step, "frames": rows,
                "ms_mean": float(np.mean([r["ms"] for r in rows])),
                "texels_rebaked_mean": float(np.mean([r["texels"] for r in rows])), "verify_ledger": checks}

    return {"workload": "the same pages; sun panel and nearest rect light moving (epsilon 0.01)",
            "slow": run(1.2e-6, 0.001), "fast": run(1e-4, 0.02)}


def sky_form(scene, flat, sc, recv, h_recv, check: bool = True, stride: int = 64) -> dict:
    """A uniform sky over the rock field's shading points (its 1280x720
    camera pass): six cube-face cones per point (the reference's
    `sky_irradiance`, the sky term of its per-pixel shading), faces summed
    in order, each wave taking an 8 x 8 patch of pixels (the list launch,
    64 consecutive points a wave, is timed beside it). Timed with HIP events
    on the launch stream; every `stride`th point checked against the oracle
    bit for bit."""
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
    W, H = sc.width, sc.height

    def timed(launch, reps=5):
        launch()
        torch.cuda.synchronize()
        ms = []
        for _ in range(reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            launch()
            e1.record(stream)
            torch.cuda.synchronize()
            ms.append(e0.elapsed_time(e1))
        return float(np.median(ms))

    t_list = timed(lambda: scene.sky_device(recv, opts, out, faces, full, stream=stream), reps=2)
    t = timed(lambda: scene.sky_view_device(recv, W, H, opts, out, faces, full, tile_w=8, stream=stream))
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
           "list_launch_ms": t_list,
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
    ap.add_argument("--tile-w", type=int, default=0,
                    help="pixels across a wave's tile (tile_w x 64/tile_w); 0: 8 on the courtyard, 16 on the rock field")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true", help="skip the courtyard's 64-view batch")
    ap.add_argument("--no-rock-field", action="store_true", help="skip the rock field's plain frame (and the sky)")
    ap.add_argument("--no-served", action="store_true", help="skip the served form over the boulder field")
    ap.add_argument("--no-vcache", action="store_true", help="skip the view cache over the boulder field")
    ap.add_argument("--no-sky", action="store_true", help="skip the sky light over the rock field")
    ap.add_argument("--no-lightmap", action="store_true", help="skip the lightmap cache over the lit room")
    ap.add_argument("--only-form", choices=["served", "served_sweep", "view_cache", "lightmap"],
                    help="run one form alone and print its JSON (for profiling it); no headline line")
    ap.add_argument("--served-requests", type=int, default=64, help="requests in the saturated served run")
    ap.add_argument("--served-seconds", type=float, default=10.0, help="length of the steady served run")
    ap.add_argument("--served-rates", default="0.5,1,2,3,4", help="offered views/s of the served sweep")
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
    from continuum_amd.scenes import courtyard, courtyard_frame, meshes_to_arrays, rock_field

    device = Device(0)
    if args.only_form:
        if args.only_form == "served":
            out = served_form(device, n_requests=args.served_requests, seconds=args.served_seconds)
        elif args.only_form == "served_sweep":
            out = served_sweep_form(device, [float(x) for x in args.served_rates.split(",")], args.served_seconds)
        elif args.only_form == "lightmap":
            out = lightmap_form(device)
        else:
            out = view_cache_form(device)
        print(json.dumps({"form": args.only_form, **out}))
        return 0

    dev = torch.device("cuda:0")
    # A copy of known size (256 MiB read, 256 MiB written), for calibrating
    # HBM byte counters when this command runs under rocprofv3 --pmc.
    ca = torch.ones(64 << 20, dtype=torch.float32, device=dev)
    cb = torch.empty_like(ca)
    torch.neg(ca, out=cb)
    del ca, cb

    # ---- The headline: the courtyard frame (3.05M triangles; per frame the
    # curtains move and their hierarchies are refit on the GPU, then the
    # camera pass and 1280x720 x 4 lights of cone queries).
    t0 = time.perf_counter()
    sc = courtyard(width=args.width, height=args.height)
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    t_build = time.perf_counter() - t0
    scene = Scene(device, flat)
    n_frames = 1 + args.warmup + args.steps
    h_poses = [courtyard_frame(sc, f, flat.tris, flat.ranges) for f in range(n_frames)]
    poses = [torch.from_numpy(p).to(dev) for p in h_poses]
    torch.cuda.synchronize()
    K = args.steps
    note(f"courtyard built ({sc.n_tris} triangles, {t_build:.1f} s on the host)")
    m = plain_frames(scene, sc, K, args.warmup, args.tile_w or COURTYARD_TILE_W, poses=poses)
    note(f"courtyard frames: {m['wall'] / K * 1e3:.2f} ms a frame, cone pass {m['cone_ms'].mean():.2f} ms")
    counters = m["counters"]
    flags = m["flags"]
    result = {
        "metric": METRIC,
        "value": K * m["Q"] / m["wall"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": args.warmup,
        "ms_per_step": m["wall"] / K * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded courtyard; geometry generated, no assets)",
        "config": {
            "workload": "courtyard frame: per-frame GPU refit of the displaced curtains, camera pass, one cone "
                        "query per shading point and light (plain per-view launch)",
            "triangles": int(sc.n_tris),
            "curtain_triangles": int(sum(len(sc.meshes[i]) for i in sc.notes["curtains"])),
            "objects": len(sc.meshes),
            "shading_points": m["W"] * m["H"],
            "lights": m["nl"],
            "queries_per_step": m["Q"],
            "traced_queries_per_step": m["lit"] * m["nl"],
            "hierarchy": "triangles",
            "prefilter_bins": sc.prefilter,
            "scene_hbm_bytes": scene.hbm_bytes,
            "host_build_s": round(t_build, 3),
        },
        "tile": f"{args.tile_w or COURTYARD_TILE_W} x {64 // (args.tile_w or COURTYARD_TILE_W)} shading points a wave",
        "latency_ms": {"p50": float(np.percentile(m["lat"], 50)), "p99": float(np.percentile(m["lat"], 99)),
                       "refit_mean": float(m["refit_ms"].mean()), "camera_pass_mean": float(m["prim_ms"].mean()),
                       "cone_pass_mean": float(m["cone_ms"].mean()),
                       "cone_pass_first_launch_natural_order": m["first_cone_ms"]},
        "roofline": roofline_of(m, "courtyard"),
        "diagnostics": {"capped_walks_timed_frames": int(counters[0]),
                        "capped_share": int(counters[0]) / (m["Q"] * m["steps"]),
                        "deeper_than_48": int(counters[1]), "truncated": int(counters[2]),
                        "flagged_capped_last_frame": int(((flags & 1) != 0).sum())},
    }
    if not args.no_cpu_baseline:
        oflat = posed_oracle_flat(flat, counts, h_poses[m["last_frame"]])
        chk = oracle_check(m, oflat, sc, args.cpu_seconds, single_seconds=4.0)
        result["cpu_baseline"] = chk["cpu_baseline"]
        result["parity"] = chk["parity"]
        note(f"oracle sample: {chk['cpu_baseline']['value']:.0f} queries/s, parity {chk['parity']}")
    del m, poses, h_poses
    torch.cuda.synchronize()
    forms = result.setdefault("forms", {})
    if not args.no_forms:
        forms["courtyard_64_views"] = courtyard_batch_form(device, scene,
                                                           _lib.default_options(hierarchy=2, prefilter=sc.prefilter))
        b = forms["courtyard_64_views"]
        b["batched_over_plain_frame_rate"] = b["batched_graph"]["queries_per_s"] / result["value"]
        note(f"courtyard 64 views: batched/sequential {b['batched_over_sequential']:.2f}")
    del scene, flat

    # ---- The rock field's plain frame (1.03M triangles, 3 lights), and the
    # sky over its shading points.
    if not args.no_rock_field:
        t0 = time.perf_counter()
        rf = rock_field(width=args.width, height=args.height)
        rt9, rc = meshes_to_arrays(rf.meshes)
        rflat = Flat.build_arrays(rt9, rc)
        t_rb = time.perf_counter() - t0
        rscene = Scene(device, rflat)
        rm = plain_frames(rscene, rf, K, args.warmup, args.tile_w or PLAIN_TILE_W)
        leg = {"workload": "rock field 1.03M triangles, 1280x720 x 3 lights (sun disk + two rects), plain per-view "
                           "launch (camera pass + cones)",
               "queries_per_s": K * rm["Q"] / rm["wall"], "ms_per_frame": rm["wall"] / K * 1e3,
               "p50_ms": float(np.percentile(rm["lat"], 50)), "p99_ms": float(np.percentile(rm["lat"], 99)),
               "camera_pass_mean": float(rm["prim_ms"].mean()), "cone_pass_mean": float(rm["cone_ms"].mean()),
               "cone_pass_first_launch_natural_order": rm["first_cone_ms"],
               "capped_walks": int(rm["counters"][0]), "host_build_s": round(t_rb, 3),
               "roofline": roofline_of(rm, "rock_field")}
        if not args.no_cpu_baseline:
            leg.update(oracle_check(rm, rflat, rf, cpu_seconds=4.0))
        forms["rock_field_plain"] = leg
        note(f"rock field: {leg['queries_per_s'] / 1e6:.1f} M queries/s")
        if not args.no_sky:
            forms["sky"] = sky_form(rscene, rflat, rf, rm["recv"], rm["h_recv"], check=not args.no_cpu_baseline)
            note(f"sky: {forms['sky']['ms']:.1f} ms")
        del rm, rscene, rflat
    if not args.no_served:
        forms["served"] = served_form(device, n_requests=args.served_requests, seconds=args.served_seconds)
        note("served form done")
    if not args.no_vcache:
        forms["view_cache"] = view_cache_form(device)
        note("view cache done")
    if not args.no_lightmap:
        forms["lightmap"] = lightmap_form(device)

    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
