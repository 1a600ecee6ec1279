"""Benchmark of the MI355X SortedAlpha path.

Default workload (one GPU): plain SortedAlpha, 300,000 splats SH degree 3,
one 1920x1080 view (synthetic scene from a fixed seed; see synth.py).
A step is one full frame: depth key + cull + projection + SH colour, depth
sort, tile binning, raster, with the scene resident in HBM.

Prints one JSON line (frames/s) with a roofline entry for the dominant
kernel and the CPU oracle timed on a bounded sample. Beside the plain
figure the line carries the batched forms, each measured against the plain
one-view-per-launch path on the same GPU and the same views:
  batched_1080p   8 cameras over the same 300k SH3 scene at 1920x1080, one
                  batch per step
  batched_64v     64 views over 8 resident scenes (50k..1M splats, SH3) at
                  320x240..1920x1080, one batch per step
and the served form (16 resident scenes totalling 10M splats, SH3):
  served_poisson  open loop, Poisson arrivals at 5,000 requests/s; latency
                  from each request's scheduled arrival to its RGBA8 image in
                  pinned host memory
  served_closed64 64 clients each submitting its next request when the last
                  one returns
"""

from __future__ import annotations

import argparse
import json
import math
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
for sub in ("gsplat-rs_amd",):
    sys.path.insert(0, str(ROOT / sub))

import numpy as np  # noqa: E402

import gsplat_abi as A  # noqa: E402
import synth  # noqa: E402

MI355X_HBM_GBPS = 8000.0
MI355X_FP32_TFLOPS = 157.3


def nearest_rank(sorted_vals, p):
    """tools/bench-runner/src/artifact.rs:433-436."""
    rank = max(math.ceil(p * len(sorted_vals)), 1)
    return sorted_vals[min(rank - 1, len(sorted_vals) - 1)]


def cpu_baseline(scene, view, cam):
    """Oracle (CPU restatement of the reference path) on the host cores."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_py as O

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    osc = O.Scene(scene.positions, scene.opacity, scene.scale_xyz, scene.rotation_xyzw, scene.color_dc,
                  scene.sh_degree, scene.sh_rest)
    frames, t0 = 0, time.perf_counter()
    while True:
        O.render(osc, cam.as_dict(), view.width, view.height, threads=threads)
        frames += 1
        if time.perf_counter() - t0 > 10.0 or frames >= 8:
            break
    el = time.perf_counter() - t0
    st = O.last_timings()
    return {"value": round(frames / el, 4), "unit": "frames/s", "cores": threads, "kind": "port",
            "sample": f"{frames} full frames of the same workload (preprocess 1 thread, sort 1 thread, raster "
                      f"{threads} threads); last frame ms: preprocess {st[0]:.1f}, sort {st[1]:.1f}, "
                      f"project {st[2]:.1f}, raster {st[3]:.1f}"}


def time_batches(batch, views, steps, warmup):
    """Frames (views) per second of `steps` back-to-back batch renders of `views`."""
    for _ in range(warmup):
        batch.render(views)
    batch.wait()
    t0 = time.perf_counter()
    for _ in range(steps):
        batch.render(views)
    batch.wait()
    return len(views) * steps / (time.perf_counter() - t0)


def time_one_per_launch(batch, views, passes, warmup):
    """Frames per second rendering the same views one per launch (the plain path)."""
    for _ in range(warmup):
        for v in views:
            batch.render([v])
    batch.wait()
    t0 = time.perf_counter()
    for _ in range(passes):
        for v in views:
            batch.render([v])
    batch.wait()
    return len(views) * passes / (time.perf_counter() - t0)


def bench_batched_1080p(scene, steps, warmup):
    rng = np.random.default_rng(synth.SEED_BASE + 1 + 500)
    dev = A.Scene(scene)
    views = [(dev, A.look_at(synth.sphere_point(rng, 4.0)), 1920, 1080) for _ in range(8)]
    b = A.Batch()
    batched = time_batches(b, views, steps, warmup)
    st = b.stage_times()
    plain = time_one_per_launch(A.Batch(), views, max(steps // 2, 1), warmup)
    return {"views": len(views), "frames_per_s": round(batched, 3), "plain_frames_per_s": round(plain, 3),
            "speedup_vs_plain": round(batched / plain, 3), "ms_per_batch": round(1e3 * len(views) / batched, 4),
            "stage_ms": {"project": round(st.project_ms, 4), "sort": round(st.sort_ms, 4),
                         "bin": round(st.bin_ms, 4), "raster": round(st.raster_ms, 4)},
            "tile_pairs": int(st.tile_pairs)}


def bench_batched_64v(steps, warmup):
    scenes, vs = synth.batched()
    dev = [A.Scene(s) for s in scenes]
    del scenes
    views = [(dev[v.scene], A.look_at(v.position, v.target), v.width, v.height) for v in vs]
    b = A.Batch()
    batched = time_batches(b, views, steps, warmup)
    st = b.stage_times()
    plain = time_one_per_launch(A.Batch(), views, max(steps // 4, 1), max(warmup // 4, 1))
    pixels = sum(w * h for (_, _, w, h) in views)
    return {"views": len(views), "scenes": len(dev), "frames_per_s": round(batched, 3),
            "plain_frames_per_s": round(plain, 3), "speedup_vs_plain": round(batched / plain, 3),
            "mpix_per_s": round(batched / len(views) * pixels / 1e6, 1),
            "ms_per_batch": round(1e3 * len(views) / batched, 4),
            "stage_ms": {"project": round(st.project_ms, 4), "sort": round(st.sort_ms, 4),
                         "bin": round(st.bin_ms, 4), "raster": round(st.raster_ms, 4)},
            "splat_slots": int(st.splat_count), "visible": int(st.visible_count), "tile_pairs": int(st.tile_pairs)}


def pct(vals):
    vals = sorted(vals)
    return {"p50": round(nearest_rank(vals, 0.5), 3), "p99": round(nearest_rank(vals, 0.99), 3),
            "max": round(vals[-1], 3)}


def bench_served(requests, rate, closed_seconds):
    import queue
    import threading

    scenes, reqs = synth.served(requests=requests, rate=rate)
    dev = [A.Scene(s) for s in scenes]
    del scenes
    srv = A.Server(max_batch_views=64)
    ids = [srv.add_scene(d) for d in dev]
    cams = [(ids[v.scene], A.look_at(v.position, v.target), v.width, v.height) for (_, v) in reqs]
    pool = queue.Queue()
    for _ in range(128):
        pool.put(A.PinnedImage(1920, 1080))
    # warm the pipelines' buffers on a few batches
    warm = [srv.submit(*cams[i % len(cams)]) for i in range(256)]
    for t in warm:
        srv.wait(t)

    # Open loop: one thread submits at the scheduled arrival times, one waits in order.
    tickets = queue.Queue()
    lat, server_ms, batch_views = [], [], []
    t0 = time.perf_counter() + 0.05

    def submitter():
        for (arrival, _), (sid, cam, w, h) in zip(reqs, cams):
            delay = t0 + arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            buf = pool.get()
            tickets.put((t0 + arrival, srv.submit(sid, cam, w, h, buf), buf))
        tickets.put(None)

    def waiter():
        while True:
            item = tickets.get()
            if item is None:
                return
            due, t, buf = item
            r = srv.wait(t)
            lat.append((time.perf_counter() - due) * 1e3)
            server_ms.append(r.total_ms)
            batch_views.append(r.batch_views)
            pool.put(buf)

    th = [threading.Thread(target=submitter), threading.Thread(target=waiter)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    span = max(time.perf_counter() - t0, 1e-9)
    open_loop = {"requests": len(reqs), "offered_per_s": rate, "completed_per_s": round(len(lat) / span, 1),
                 "latency_ms": pct(lat), "server_ms": pct(server_ms),
                 "mean_batch_views": round(float(np.mean(batch_views)), 2)}

    # Closed loop: 64 clients.
    done, clat = [0], []
    lock = threading.Lock()
    stop_at = time.perf_counter() + closed_seconds

    def client(k):
        buf = A.PinnedImage(1920, 1080)
        i = k
        while time.perf_counter() < stop_at:
            sid, cam, w, h = cams[i % len(cams)]
            ts = time.perf_counter()
            srv.wait(srv.submit(sid, cam, w, h, buf))
            with lock:
                clat.append((time.perf_counter() - ts) * 1e3)
                done[0] += 1
            i += 64
        buf.close()

    tc = time.perf_counter()
    th = [threading.Thread(target=client, args=(k,)) for k in range(64)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    closed = {"clients": 64, "completed_per_s": round(done[0] / (time.perf_counter() - tc), 1),
              "latency_ms": pct(clat)}
    st = srv.stats()
    srv.close()
    return {"scenes": len(dev), "splats": int(sum(d.count for d in dev)),
            "resident_scene_bytes": int(st.resident_scene_bytes), "memory_budget_bytes": int(st.memory_budget_bytes),
            "poisson": open_loop, "closed64": closed, "batches": int(st.batches),
            "rejected": int(st.rejected)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-only", action="store_true", help="skip the batched and served sections")
    ap.add_argument("--served-requests", type=int, default=2000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        sys.exit(2)
    if not A.device_available():
        print("no gfx950 device visible", file=sys.stderr)
        sys.exit(1)

    scene, view = synth.plain_1080p()
    cam = A.look_at(view.position, view.target)
    ctx = A.Context(view.width, view.height)
    ctx.load(scene)
    ctx.set_camera(cam)
    for _ in range(args.warmup):
        ctx.render()
    ctx.wait()

    # Timed region: K back-to-back frames.
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.render()
    ctx.wait()
    elapsed = time.perf_counter() - t0
    fps = args.steps / elapsed

    # Per-frame latency and stage device times (separate pass, one frame at a time).
    lat, stages = [], []
    for _ in range(min(args.steps, 60)):
        t = time.perf_counter()
        ctx.render()
        ctx.wait()
        lat.append((time.perf_counter() - t) * 1e3)
        s = ctx.stage_times()
        stages.append((s.project_ms, s.sort_ms, s.bin_ms, s.raster_ms, s.total_ms))
    lat.sort()
    st = ctx.stage_times()
    med = np.median(np.array(stages), axis=0)
    names = ["project", "sort", "bin", "raster"]
    dom = int(np.argmax(med[:4]))

    n, v = st.splat_count, st.visible_count
    k = 3 * ((scene.sh_degree + 1) ** 2 - 1)
    # Projection bytes: positions for all, covariance + alpha + cut for visible, colour terms and
    # record/rect writes for visible, keys/values for all.
    proj_bytes = 12 * n + 8 * n + (32 + 12 + 4 * k + 56) * v
    roof = {"project": ("hbm", proj_bytes / (med[0] * 1e-3) / 1e9, MI355X_HBM_GBPS, "GB/s", proj_bytes)}
    kernel, (bound, achieved, peak, unit, traffic) = names[0], roof["project"]
    line = {
        "metric": "frames/s @300k splats SH3 1080p (plain SortedAlpha)",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded 3-component Gaussian mixture, SH3)",
        "config": {"workload": "plain-1080p-300k-sh3", "splats": n, "width": view.width, "height": view.height,
                   "visible": v, "tile_pairs": int(st.tile_pairs)},
        "latency_ms": {"p50": round(nearest_rank(lat, 0.5), 4), "p99": round(nearest_rank(lat, 0.99), 4)},
        "stage_ms": {nm: round(float(x), 4) for nm, x in zip(names, med[:4])},
        "dominant_stage": names[dom],
        "roofline": {"kernel": "k_project", "bound": bound, "achieved": round(achieved, 2), "peak": peak,
                     "unit": unit, "frac": round(achieved / peak, 4), "traffic": None},
    }
    ctx.close()
    if not args.plain_only:
        line["batched_1080p"] = bench_batched_1080p(scene, max(args.steps // 8, 5), max(args.warmup // 4, 2))
        print(f"batched_1080p {line['batched_1080p']}", file=sys.stderr, flush=True)
        line["batched_64v"] = bench_batched_64v(max(args.steps // 20, 5), max(args.warmup // 4, 2))
        print(f"batched_64v {line['batched_64v']}", file=sys.stderr, flush=True)
        line["served"] = bench_served(args.served_requests, 5000.0, 2.0)
        print(f"served {line['served']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(scene, view, cam)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
