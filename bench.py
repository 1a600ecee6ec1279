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
and two forms of the 1080p scene:
  session         a context on a slow orbit, sorting every frame and with
                  session order reuse (sort interval 2), with the stale-order
                  image gate of the reused frames
  packed          the scene in the packed geometry path (20-byte records and
                  the i8 SH sidecar): frames/s, stage times, device bytes and
                  the packed-vs-direct image gate
and the served form again with its 16 scenes in the packed geometry path
(served_packed).
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


def host_cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, view, cam):
    """Oracle (CPU restatement of the reference path) on the host cores: per frame the preprocess and
    the sort on one thread (as the reference's), the projection of the sorted instances and the blend
    on `threads`; the per-splat alpha and covariance are computed once, as the reference does at load."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import oracle_py as O

    threads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    osc = O.Scene(scene.positions, scene.opacity, scene.scale_xyz, scene.rotation_xyzw, scene.color_dc,
                  scene.sh_degree, scene.sh_rest)
    prepared = O.precompute(osc)
    frames, t0 = 0, time.perf_counter()
    while True:
        O.render_prepared(osc, prepared, cam.as_dict(), view.width, view.height, threads=threads)
        frames += 1
        if time.perf_counter() - t0 > 10.0 or frames >= 8:
            break
    el = time.perf_counter() - t0
    st = O.last_timings()
    return {"value": round(frames / el, 4), "unit": "frames/s", "cores": threads, "kind": "port",
            "cpu": host_cpu_model(),
            "sample": f"{frames} full frames of the same workload (load-time alpha/covariance once; per frame "
                      f"preprocess 1 thread, sort 1 thread, projection and blend {threads} threads); last frame "
                      f"ms: preprocess {st[0]:.1f}, sort {st[1]:.1f}, project {st[2]:.1f}, raster {st[3]:.1f}"}


def bench_plain_config(name, scene, view, frames, warmup):
    """Back-to-back frames of one more plain configuration on a context (its
    own scene and camera), with the stage device times of stage-timed frames
    and the frame's sizes."""
    ctx = A.Context(view.width, view.height)
    ctx.load(scene)
    ctx.set_camera(A.look_at(view.position, view.target))
    ctx.set_stage_timing(False)
    for _ in range(warmup):
        ctx.render()
    ctx.wait()
    t0 = time.perf_counter()
    for _ in range(frames):
        ctx.render()
    ctx.wait()
    fps = frames / (time.perf_counter() - t0)
    ctx.set_stage_timing(True)
    stages = []
    for _ in range(5):
        ctx.render()
        ctx.wait()
        s = ctx.stage_times()
        stages.append((s.project_ms, s.sort_ms, s.bin_ms, s.raster_ms))
    st = ctx.stage_times()
    ctx.close()
    med = np.median(np.array(stages), axis=0)
    return {"workload": name, "splats": scene.count, "sh_degree": scene.sh_degree, "width": view.width,
            "height": view.height, "frames_per_s": round(fps, 3), "ms_per_frame": round(1e3 / fps, 4),
            "mpix_per_s": round(fps * view.width * view.height / 1e6, 1), "visible": int(st.visible_count),
            "tile_pairs": int(st.tile_pairs),
            "stage_ms": {nm: round(float(x), 4) for nm, x in zip(["project", "sort", "bin", "raster"], med)}}


def write_plain_artifact(ctx, scene, view, directory, iterations=120, warmup=10):
    """The plain workload as a gsplat-benchmark/v1 run: the reference runner's
    protocol (warmup 10, 120 measured frames, one render call + wait each) on
    the context, with its stage times as the frame stats."""
    import bench_artifact as BA
    started = BA.utc_now()
    run_id = "plain-1080p-300k-sh3-" + started.replace(":", "").replace("-", "").replace(".", "")
    ctx.set_stage_timing(True)
    frames, m0, m1 = BA.measure_context(ctx, iterations, warmup, run_id)
    arrays = [scene.positions, scene.opacity, scene.scale_xyz, scene.rotation_xyzw, scene.color_dc]
    if scene.sh_rest is not None:
        arrays.append(scene.sh_rest)
    blob = b"".join(np.ascontiguousarray(a, dtype=np.float32).tobytes() for a in arrays)
    trace = (f"gsplat-camera-trace/v1\nmode=static\ncamera=look_at position={list(view.position)} "
             f"target={list(view.target)}\n").encode()
    build, build_un = BA.build_info(ROOT)
    env, env_un = BA.environment_info(adapter="AMD Instinct MI355X (gfx950)")
    timed = frames[0]["preprocess_ms"] is not None
    unavailable = build_un + env_un + ["frames[*].sort_refreshed"]
    if not timed:
        unavailable += [f"frames[*].{k}" for k in ("preprocess_ms", "sort_ms", "geometry_submit_ms")]
    manifest = {
        "schema": BA.SCHEMA, "record_type": "manifest", "run_id": run_id,
        "identity": {"series_id": "mi355x-plain-1080p-300k-sh3", "started_at_utc": started,
                     "ended_at_utc": BA.utc_now(), "measurement_started_at_utc": m0,
                     "measurement_ended_at_utc": m1},
        "build": build,
        "dataset": {"id": "synthetic-plain-1080p-300k-sh3", "sha256": BA.sha256_bytes(blob), "bytes": len(blob),
                    "splat_count": scene.count, "sh_degree": scene.sh_degree},
        "trace": {"id": "static-look-at-v1", "sha256": BA.sha256_bytes(trace)},
        "renderer": {"implementation": "gsplat-rs_amd", "path": "sorted_alpha_tiles", "backend": "hip-gfx950",
                     "sort_policy": "gpu_every_frame"},
        "display": {"width": view.width, "height": view.height, "dpr": 1.0, "refresh_hz": 60.0,
                    "frame_budget_ms": 1000.0 / 60.0, "refresh_hz_source": "configured",
                    "frame_budget_source": "configured"},
        "environment": env,
        "timing": {"frame_wall_source": "host wall clock from the render call to gsplat_context_wait returning "
                                        "(device completion); stage times are device events"},
        "unavailable_fields": unavailable,
    }
    BA.write_run(directory, manifest, frames, warmup)
    BA.check_run(directory)
    print(f"artifact written: {directory}", file=sys.stderr, flush=True)


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


def stage_frame(b, views):
    """Stage device times of one more batch frame, rendered with stage events."""
    b.set_stage_timing(True)
    b.render(views)
    st = b.stage_times()
    b.set_stage_timing(False)
    return st


def bench_batched_1080p(scene, steps, warmup):
    rng = np.random.default_rng(synth.SEED_BASE + 1 + 500)
    dev = A.Scene(scene)
    views = [(dev, A.look_at(synth.sphere_point(rng, 4.0)), 1920, 1080) for _ in range(8)]
    b = A.Batch()
    batched = time_batches(b, views, steps, warmup)
    st = stage_frame(b, views)
    pb = A.Batch()
    plain = time_one_per_launch(pb, views, max(steps // 2, 1), warmup)
    b.close()
    pb.close()
    dev.close()
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
    st = stage_frame(b, views)
    pb = A.Batch()
    plain = time_one_per_launch(pb, views, max(steps // 4, 1), max(warmup // 4, 1))
    b.close()
    pb.close()
    for d in dev:
        d.close()
    pixels = sum(w * h for (_, _, w, h) in views)
    return {"views": len(views), "scenes": len(dev), "frames_per_s": round(batched, 3),
            "plain_frames_per_s": round(plain, 3), "speedup_vs_plain": round(batched / plain, 3),
            "mpix_per_s": round(batched / len(views) * pixels / 1e6, 1),
            "ms_per_batch": round(1e3 * len(views) / batched, 4),
            "stage_ms": {"project": round(st.project_ms, 4), "sort": round(st.sort_ms, 4),
                         "bin": round(st.bin_ms, 4), "raster": round(st.raster_ms, 4)},
            "splat_slots": int(st.splat_count), "visible": int(st.visible_count), "tile_pairs": int(st.tile_pairs)}


BATCH_WORKLOADS = {
    "batched-64v": "64 views over 8 scenes (50k-1M splats SH3) at 320x240-1920x1080",
    "batched-8x1080p": "8 cameras over the 300k SH3 scene at 1920x1080",
    "served-24": "the first 24 served requests (16 scenes, 10M splats SH3) as one batch, binned as the server bins",
}


def batch_views(name):
    """Views of a batch workload of BATCH_WORKLOADS, scenes uploaded (the objects stay alive with the views)."""
    if name == "batched-64v":
        scenes, vs = synth.batched()
        dev = [A.Scene(s) for s in scenes]
        return [(dev[v.scene], A.look_at(v.position, v.target), v.width, v.height) for v in vs]
    if name == "batched-8x1080p":
        scene, _ = synth.plain_1080p()
        rng = np.random.default_rng(synth.SEED_BASE + 1 + 500)
        dev = A.Scene(scene)
        return [(dev, A.look_at(synth.sphere_point(rng, 4.0)), 1920, 1080) for _ in range(8)]
    if name == "served-24":
        scenes, reqs = synth.served(requests=24)
        dev = [A.Scene(s) for s in scenes]
        return [(dev[v.scene], A.look_at(v.position, v.target), v.width, v.height) for (_, v) in reqs]
    raise SystemExit(f"unknown batch workload {name!r}; one of {sorted(BATCH_WORKLOADS)}")


def profile_batch(name, steps, warmup):
    """`steps` back-to-back renders of one batch workload (for rocprofv3), and one summary line."""
    views = batch_views(name)
    if name.startswith("served"):
        os.environ["GSPLAT_BIN_MAX_CTILES"] = "512"  # the server's binning cap (read when a pipeline is made)
    b = A.Batch()
    fps = time_batches(b, views, steps, warmup)
    print(json.dumps({"workload": name, "views": len(views), "batches": warmup + steps,
                      "ms_per_batch": round(1e3 * len(views) / fps, 4)}), flush=True)


def summarize_batch_profiles(name, trace_dir, fetch_dir, write_dir, out_dir, steps, warmup):
    """rocprofv3 CSVs of `bench.py --profile-batch NAME` -> profiles/kernels_NAME.md: per kernel and
    per stage, device time and HBM bytes per batch (FETCH_SIZE and WRITE_SIZE scaled by the copy
    calibration factors of the plain summary)."""
    import csv
    import collections

    out_dir = pathlib.Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)
    scale = json.loads(TRAFFIC_FILE.read_text())
    ff, wf = float(scale["fetch_scale"]), float(scale["write_scale"])

    def rows(d, suffix):
        return list(csv.DictReader(open(next(pathlib.Path(d).glob(f"*{suffix}")))))

    def batches(rs):
        return max(sum(_kernel_short(r["Kernel_Name"]).startswith("k_project") for r in rs), 1)

    trace = rows(trace_dir, "kernel_trace.csv")
    nb = batches(trace)
    k_us, k_calls, s_us = collections.defaultdict(float), collections.Counter(), collections.defaultdict(float)
    for stage, r in _stage_sequence(trace):
        k = _kernel_short(r["Kernel_Name"])
        d = (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
        k_us[k] += d
        k_calls[k] += 1
        if stage in STAGES:
            s_us[stage] += d

    def counter(d):
        rs = rows(d, "counter_collection.csv")
        per_k, per_s = collections.defaultdict(float), collections.defaultdict(float)
        for stage, r in _stage_sequence(rs):
            v = float(r["Counter_Value"]) * 1024.0
            per_k[_kernel_short(r["Kernel_Name"])] += v
            if stage in STAGES:
                per_s[stage] += v
        return per_k, per_s, batches(rs)

    fk, fs, nf = counter(fetch_dir)
    wk, ws, nw = counter(write_dir)
    total = sum(k_us.values())
    lines = [f"# Kernel trace and HBM traffic: {name} ({BATCH_WORKLOADS[name]})", "",
             "Commands (from /tmp, TMPDIR=/tmp):", "",
             f"    rocprofv3 --kernel-trace --stats --output-format csv -d TRACE -o b -- \\",
             f"        python3 bench.py --profile-batch {name} --steps {steps} --warmup {warmup}",
             f"    rocprofv3 --pmc FETCH_SIZE --output-format csv -d FETCH -o b -- python3 bench.py --profile-batch {name} "
             f"--steps {steps} --warmup {warmup}",
             "    rocprofv3 --pmc WRITE_SIZE (same, into WRITE)", "",
             f"Summarised with `python3 bench.py --summarize-batch-profiles {name} TRACE FETCH WRITE --steps {steps} "
             f"--warmup {warmup}`. {nb} batches traced; figures are per batch. FETCH_SIZE is scaled by {ff:.3f} and "
             f"WRITE_SIZE by {wf:.3f} (the copy calibration of pmc_traffic_plain-1080p-300k-sh3.json).", "",
             "| kernel | calls per batch | avg us | us per batch | % | fetch MB | write MB |",
             "|---|---:|---:|---:|---:|---:|---:|"]
    for k in sorted(k_us, key=lambda k: -k_us[k]):
        lines.append(f"| {k} | {k_calls[k] / nb:.2f} | {k_us[k] / k_calls[k]:.2f} | {k_us[k] / nb:.1f} | "
                     f"{100 * k_us[k] / total:.1f} | {fk[k] / nf * ff / 1e6:.1f} | {wk[k] / nw * wf / 1e6:.1f} |")
    lines += ["", "| stage | us per batch | fetch MB | write MB |", "|---|---:|---:|---:|"]
    for stage in STAGES:
        lines.append(f"| {STAGE_KERNEL[stage]} | {s_us[stage] / nb:.1f} | {fs[stage] / nf * ff / 1e6:.1f} | "
                     f"{ws[stage] / nw * wf / 1e6:.1f} |")
    path = out_dir / f"kernels_{name}.md"
    path.write_text("\n".join(lines) + "\n")
    print(f"wrote {path}", file=sys.stderr)


def summarize_counters(pmc_dir, out_dir, counters, lds=False):
    """rocprofv3 --pmc CSV of the plain workload -> profiles/sq_counters_plain-1080p-300k-sh3.md (or,
    with `lds`, lds_valu_plain-1080p-300k-sh3.md plus its .json): the per-dispatch average of each
    counter for every kernel of the frame. The LDS pass also derives, per kernel, LDS utilisation
    (SQ_LDS_IDX_ACTIVE over CU cycles: the share of cycles the CU's LDS pipe serves indexed
    operations) and VALU busy (SQ_ACTIVE_INST_VALU over CU cycles), with the kernel's cycles taken as
    GRBM_GUI_ACTIVE per XCD (rocprofv3 sums the 8 XCDs' GRBM counters) and 256 CUs."""
    import csv
    import collections

    path = next(pathlib.Path(pmc_dir).glob("*counter_collection.csv"))
    acc = collections.defaultdict(lambda: collections.defaultdict(list))
    for r in csv.DictReader(open(path)):
        acc[_kernel_short(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    names = sorted({c for d in acc.values() for c in d})
    stem = "lds_valu_plain-1080p-300k-sh3" if lds else "sq_counters_plain-1080p-300k-sh3"
    flag = "--summarize-lds-counters" if lds else "--summarize-counters"
    title = "LDS and VALU utilisation" if lds else "SQ counters"
    derived = ["LdsUtil", "VALUBusy"] if lds else []
    lines = [f"# {title}: plain-1080p-300k-sh3 (300,000 splats SH3, 1920x1080)", "",
             "Command (from /tmp, TMPDIR=/tmp):", "",
             f"    rocprofv3 --pmc {counters} --output-format csv -d DIR -o plain -- \\",
             "        python3 bench.py --plain-only --steps 8 --warmup 2 --no-cpu-baseline",
             "", f"Summarised with `python3 bench.py {flag} DIR`. Per-dispatch averages."
             + (" LdsUtil = SQ_LDS_IDX_ACTIVE / (256 CUs x GRBM_GUI_ACTIVE / 8 XCDs); VALUBusy = "
                "SQ_ACTIVE_INST_VALU / (256 x GRBM_GUI_ACTIVE / 8)." if lds else ""), "",
             "| kernel | dispatches | " + " | ".join(names + derived) + " |",
             "|---|---:|" + "---:|" * (len(names) + len(derived))]
    summary = {}
    for k in sorted(acc, key=lambda k: -sum(acc[k].get("SQ_WAVE_CYCLES", acc[k].get("GRBM_GUI_ACTIVE", [0])))):
        d = acc[k]
        n = max(len(v) for v in d.values())
        m = {c: sum(d[c]) / len(d[c]) for c in names if d.get(c)}
        cells = [f"{m[c]:.0f}" if c in m else "" for c in names]
        if lds and m.get("GRBM_GUI_ACTIVE"):
            cyc = 256.0 * m["GRBM_GUI_ACTIVE"] / 8.0
            util = {"lds_util": m.get("SQ_LDS_IDX_ACTIVE", 0.0) / cyc,
                    "valu_busy": m.get("SQ_ACTIVE_INST_VALU", 0.0) / cyc}
            summary[k] = {key: round(v, 4) for key, v in util.items()}
            cells += [f"{100 * util['lds_util']:.1f}%", f"{100 * util['valu_busy']:.1f}%"]
        elif lds:
            cells += ["", ""]
        lines.append(f"| {k} | {n} | " + " | ".join(cells) + " |")
    out = pathlib.Path(out_dir)
    out.mkdir(parents=True, exist_ok=True)
    (out / f"{stem}.md").write_text("\n".join(lines) + "\n")
    if lds:
        (out / f"{stem}.json").write_text(json.dumps({"counters": counters, "kernels": summary}, indent=1) + "\n")


SQ_COUNTERS = ("SQ_WAVES SQ_INSTS_VALU SQ_INSTS_SALU SQ_INSTS_LDS SQ_ACTIVE_INST_VALU SQ_WAIT_INST_LDS "
               "SQ_WAVE_CYCLES SQ_BUSY_CYCLES")
LDS_COUNTERS = ("SQ_LDS_IDX_ACTIVE SQ_LDS_BANK_CONFLICT SQ_INSTS_LDS SQ_WAIT_INST_LDS SQ_ACTIVE_INST_LDS "
                "SQ_ACTIVE_INST_VALU SQ_INSTS_VALU SQ_BUSY_CYCLES GRBM_GUI_ACTIVE")
UTIL_FILE = ROOT / "profiles" / "lds_valu_plain-1080p-300k-sh3.json"


def load_utilisation():
    """LDS utilisation and VALU busy per kernel from the committed rocprofv3 summary; {} when absent."""
    try:
        return json.loads(UTIL_FILE.read_text()).get("kernels", {})
    except (OSError, ValueError):
        return {}

TRAFFIC_FILE = ROOT / "profiles" / "pmc_traffic_plain-1080p-300k-sh3.json"


def load_traffic():
    """HBM bytes per launch per kernel from the committed rocprofv3 --pmc summary (FETCH_SIZE and
    WRITE_SIZE runs, scaled by the copy calibration); {} when absent."""
    try:
        data = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return {}
    return {k: int(v["bytes_per_launch"]) for k, v in data.get("kernels", {}).items()}


def copy_calibration(mib):
    """Plain device-to-device copies of known size (the runtime's copy kernel), for scaling the
    FETCH_SIZE / WRITE_SIZE counters: each copy reads and writes `mib` MiB."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    for _ in range(5):
        if hip.hipMemcpy(b, a, ctypes.c_size_t(n), 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": n, "bytes_written": n, "copies": 5}}), flush=True)
    hip.hipFree(a)
    hip.hipFree(b)


def _kernel_short(name):
    n = name.replace("(anonymous namespace)::", "").replace("void ", "")
    return n.split("(")[0].replace("gs::", "")


# Stages of a frame, in launch order. A dispatch belongs to the stage of the
# latest stage-opening kernel before it: k_project opens "project", the first
# radix kernel after it the depth sort, k_tile_counts the binning stage (tile
# counts, block scan, pair emission, pair radix sort, tile ranges, raster
# order), k_raster<...> or k_raster_rows2<...> the raster. Sort kernels are shared by the depth and pair sorts, so launch
# order (not the name) tells them apart.
STAGES = ["project", "sort", "bin", "raster"]
STAGE_KERNEL = {"project": "k_project", "sort": "depth sort", "bin": "bin(pair sort+emit)",
                "raster": "k_raster"}


def _stage_sequence(rows):
    """(stage, row) per dispatch of CSV rows, in dispatch order; the counting raster is its own stage."""
    out, stage = [], None
    for r in sorted(rows, key=lambda r: int(r["Dispatch_Id"])):
        k = _kernel_short(r["Kernel_Name"])
        raster = k.startswith("k_raster<") or k.startswith("k_raster_rows2<")
        if k.startswith("k_project"):
            stage = "project"
        elif stage == "project" and k.startswith(("k_os_", "k_rts_", "k_scan_", "k_bucket_")):
            stage = "sort"
        elif k == "k_tile_counts":
            stage = "bin"
        elif raster and "<true" in k:
            out.append(("counting", r))
            stage = None
            continue
        elif raster:
            stage = "raster"
        out.append((stage, r))
        if raster:
            stage = None
    return out


def summarize_profiles(trace_dir, fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, frames, out_dir):
    """rocprofv3 CSVs of the plain workload -> profiles/ summaries (kernel stats, stage times, HBM traffic)."""
    import csv
    import collections

    out_dir = pathlib.Path(out_dir)
    out_dir.mkdir(parents=True, exist_ok=True)

    def rows(d, suffix):
        path = next(pathlib.Path(d).glob(f"*{suffix}"))
        return list(csv.DictReader(open(path)))

    def copies_of(d):
        return [float(r["Counter_Value"]) * 1024.0 for r in rows(d, "counter_collection.csv")
                if "copyBuffer" in r["Kernel_Name"]]

    cf, cw = copies_of(copy_fetch_dir), copies_of(copy_write_dir)
    known = len(cf) * (1024 << 20)
    ff, wf = known / sum(cf), known / sum(cw)

    def per_stage_counter(d):
        acc, kern = collections.defaultdict(float), collections.defaultdict(set)
        nframes = 0
        for stage, r in _stage_sequence(rows(d, "counter_collection.csv")):
            if stage in (None, "counting"):
                continue
            k = _kernel_short(r["Kernel_Name"])
            nframes += k.startswith("k_project")
            acc[stage] += float(r["Counter_Value"]) * 1024.0
            kern[stage].add(k)
        return acc, kern, max(nframes, 1)

    fetch, kern, nf = per_stage_counter(fetch_dir)
    write, _, nw = per_stage_counter(write_dir)
    kernels = {}
    for stage in STAGES:
        fb, wb = fetch[stage] / nf * ff, write[stage] / nw * wf
        kernels[STAGE_KERNEL[stage]] = {"fetch_bytes": int(fb), "write_bytes": int(wb),
                                        "bytes_per_launch": int(fb + wb), "kernels": sorted(kern[stage])}
    # Device time per stage and frame from the kernel trace (sum of its dispatches' durations).
    times, counts = collections.defaultdict(float), collections.Counter()
    tframes = 0
    for stage, r in _stage_sequence(rows(trace_dir, "kernel_trace.csv")):
        if stage is None:
            continue
        k = _kernel_short(r["Kernel_Name"])
        tframes += k.startswith("k_project")
        times[stage] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3
        counts[stage] += 1
    tframes = max(tframes - counts["counting"], 1)  # the counting frame's raster is reported apart
    stage_us = {STAGE_KERNEL[s]: round(times[s] / tframes, 2) for s in STAGES}
    out = {"workload": "plain-1080p-300k-sh3", "frames": nf, "fetch_scale": round(ff, 4),
           "write_scale": round(wf, 4), "kernels": kernels, "stage_us_per_frame": stage_us,
           "commands": ["rocprofv3 --pmc FETCH_SIZE --output-format csv -d DIR -o plain -- python3 bench.py "
                        "--plain-only --steps 8 --warmup 2 --no-cpu-baseline",
                        "rocprofv3 --pmc WRITE_SIZE (same)",
                        "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE -- python3 bench.py --copy-calibration 1024"]}
    out_dir = pathlib.Path(out_dir)
    (out_dir / "pmc_traffic_plain-1080p-300k-sh3.json").write_text(json.dumps(out, indent=1) + "\n")
    stats = rows(trace_dir, "kernel_stats.csv")
    lines = ["# Kernel trace: plain-1080p-300k-sh3 (300,000 splats SH3, 1920x1080)", "",
             "Command (from /tmp, TMPDIR=/tmp):", "",
             "    rocprofv3 --kernel-trace --stats --output-format csv -d DIR -o plain -- \\",
             "        python3 bench.py --plain-only --steps 30 --warmup 5 --no-cpu-baseline", "",
             "Summarised with `python3 bench.py --summarize-profiles ...`. Calls cover warmup, timed, latency",
             "and one counting frame (k_raster<true, 1>).", "",
             "| kernel | calls | avg us | total % |", "|---|---:|---:|---:|"]
    for r in sorted(stats, key=lambda r: -float(r["Percentage"])):
        lines.append(f"| {_kernel_short(r['Name'])} | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
                     f"{float(r['Percentage']):.2f} |")
    lines += ["", f"# Device time per stage and frame ({tframes} frames, kernel trace, launch order)", "",
              "| stage | kernels | us per frame |", "|---|---|---:|"]
    for stage in STAGES:
        lines.append(f"| {STAGE_KERNEL[stage]} | {', '.join(sorted(kern[stage]))} | "
                     f"{stage_us[STAGE_KERNEL[stage]]:.2f} |")
    lines += ["", "# HBM traffic per frame (rocprofv3 --pmc, two runs)", "",
              f"FETCH_SIZE scaled by {ff:.3f} and WRITE_SIZE by {wf:.3f}: the factors that make the counters read",
              f"{len(cf)} plain device-to-device copies of 1 GiB (`bench.py --copy-calibration 1024`) as 1 GiB read",
              "and 1 GiB written.", "", "| stage | kernels | fetch MB | write MB | total MB |", "|---|---|---:|---:|---:|"]
    for name, v in kernels.items():
        lines.append(f"| {name} | {', '.join(v['kernels'])} | {v['fetch_bytes'] / 1e6:.1f} | "
                     f"{v['write_bytes'] / 1e6:.1f} | {v['bytes_per_launch'] / 1e6:.1f} |")
    (out_dir / "kernels_plain-1080p-300k-sh3.md").write_text("\n".join(lines) + "\n")
    print(json.dumps(out))


def bench_session(scene, view, steps):
    """A context following a slow orbit (0.003 rad per frame) of the 1080p scene: frames/s when every
    frame sorts (the plain path) and with session order reuse at sort interval 2, and the stale-order
    gate (mean |delta RGB|, share of pixels over 3/255) of the reused frames against exact ones."""
    import math

    from image_metrics import rgba_parity_metrics

    def cam(k):
        a = 0.7 + 0.003 * k
        return A.look_at((4.0 * math.cos(a), 0.8, 4.0 * math.sin(a)))

    cams = [cam(k) for k in range(steps)]
    out = {"orbit_rad_per_frame": 0.003, "frames": steps}
    for label, interval in (("sort_every_frame", None), ("sort_interval_2", 2)):
        ctx = A.Context(view.width, view.height)
        ctx.load(scene)
        ctx.set_stage_timing(False)
        if interval:
            ctx.set_sort_interval(interval)
        for c in cams[:10]:
            ctx.set_camera(c)
            ctx.render()
        ctx.wait()
        t0 = time.perf_counter()
        for c in cams:
            ctx.set_camera(c)
            ctx.render()
        ctx.wait()
        out[label] = {"frames_per_s": round(steps / (time.perf_counter() - t0), 1)}
        ctx.close()
    exact, sess = A.Context(view.width, view.height), A.Context(view.width, view.height)
    for c in (exact, sess):
        c.load(scene)
    sess.set_sort_interval(2)
    worst = [0.0, 0.0]
    for c in cams[:16]:
        for x in (exact, sess):
            x.set_camera(c)
            x.render()
        if not sess.sort_stats().flags & 1:
            mean, frac, _ = rgba_parity_metrics(sess.rgba8(), exact.rgba8())
            worst = [max(worst[0], mean), max(worst[1], frac)]
    exact.close()
    sess.close()
    out["sort_interval_2"]["worst_reused_mean_abs_rgb"] = round(worst[0], 6)
    out["sort_interval_2"]["worst_reused_frac_over_3_255"] = round(worst[1], 6)
    out["speedup"] = round(out["sort_interval_2"]["frames_per_s"] / out["sort_every_frame"]["frames_per_s"], 3)
    return out


def bench_packed(scene, view, cam, steps, warmup, direct_stage_ms):
    """The 1080p workload with the scene in the packed geometry path: frames/s, the projection's device
    time beside the fp32 path's, device bytes of the scene in both paths, and the packed-vs-direct
    image gate (mean |delta RGB|, share of pixels over 3/255) of the frame."""
    from image_metrics import rgba_parity_metrics

    out = {}
    frames = {}
    for label, path in (("direct", A.GEOMETRY_PATH_DIRECT), ("packed", A.GEOMETRY_PATH_PACKED_ATLAS)):
        ctx = A.Context(view.width, view.height)
        ctx.set_geometry_path(path)
        ctx.load(scene)
        ctx.set_camera(cam)
        ctx.render()
        frames[label] = ctx.rgba8()
        if path == A.GEOMETRY_PATH_PACKED_ATLAS:
            ctx.set_stage_timing(False)
            for _ in range(warmup):
                ctx.render()
            ctx.wait()
            t0 = time.perf_counter()
            for _ in range(steps):
                ctx.render()
            ctx.wait()
            out["frames_per_s"] = round(steps / (time.perf_counter() - t0), 3)
            ctx.set_stage_timing(True)
            st = []
            for _ in range(min(steps, 60)):
                ctx.render()
                ctx.wait()
                s = ctx.stage_times()
                st.append((s.project_ms, s.sort_ms, s.bin_ms, s.raster_ms))
            out["stage_ms"] = {nm: round(float(x), 4)
                               for nm, x in zip(("project", "sort", "bin", "raster"), np.median(np.array(st), 0))}
        ctx.close()
    out["direct_project_ms"] = round(float(direct_stage_ms), 4)
    for label, path in (("direct", A.GEOMETRY_PATH_DIRECT), ("packed", A.GEOMETRY_PATH_PACKED_ATLAS)):
        d = A.Scene(scene, geometry_path=path)
        out[f"{label}_scene_bytes"] = int(d.device_bytes)
        d.close()
    out["bytes_ratio"] = round(out["direct_scene_bytes"] / out["packed_scene_bytes"], 3)
    mean, frac, worst = rgba_parity_metrics(frames["packed"], frames["direct"])
    out["gate"] = {"mean_abs_rgb": round(mean, 6), "frac_over_3_255": round(frac, 6), "max_abs_rgb": round(worst, 6),
                   "passes": bool(mean <= 1 / 255 and frac <= 0.001)}
    return out


def pct(vals):
    vals = sorted(vals)
    return {"p50": round(nearest_rank(vals, 0.5), 3), "p99": round(nearest_rank(vals, 0.99), 3),
            "max": round(vals[-1], 3)}


def bench_served(requests, rate, closed_seconds, pipelines=3, max_batch=64, geometry_path=0):
    """The served workload through the server: Poisson arrivals, then a closed loop of 64 clients.
    geometry_path 1 holds the 16 scenes as packed records (68 bytes per splat instead of 236)."""
    import gc
    import queue
    import threading

    gc.collect()  # objects of earlier sections are released now, not by a collection during the timing
    scenes, reqs = synth.served(requests=requests, rate=rate)
    dev = [A.Scene(s, geometry_path=geometry_path) for s in scenes]
    del scenes
    srv = A.Server(max_batch_views=max_batch, pipelines=pipelines)
    ids = [srv.add_scene(d) for d in dev]
    cams = [(ids[v.scene], A.look_at(v.position, v.target), v.width, v.height) for (_, v) in reqs]
    # Pinned targets per resolution, enough for the requests in flight at the
    # offered rate (the pool must not throttle the arrivals).
    pools = {}
    for (_, _, w, h) in cams:
        if (w, h) not in pools:
            pools[(w, h)] = queue.Queue()
            for _ in range(128):
                pools[(w, h)].put(A.PinnedImage(w, h))
    # warm the pipelines' buffers on a few batches
    warm = [srv.submit(*cams[i % len(cams)]) for i in range(256)]
    for t in warm:
        srv.wait(t)

    def regrowth(before):
        # batches that had to reallocate (a stall of their own pipeline) or render again (pair overflow)
        st = srv.stats()
        return {"buffer_grows": int(st.buffer_grows - before.buffer_grows),
                "overflow_rerenders": int(st.overflow_rerenders - before.overflow_rerenders)}, st

    st0 = srv.stats()

    # Open loop: one thread submits at the scheduled arrival times, one waits in order.
    tickets = queue.Queue()
    lat, server_ms, batch_views = [], [], []
    t0 = time.perf_counter() + 0.05

    submitted_at = [0.0]

    def submitter():
        for (arrival, _), (sid, cam, w, h) in zip(reqs, cams):
            delay = t0 + arrival - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            buf = pools[(w, h)].get()
            tickets.put((t0 + arrival, srv.submit(sid, cam, w, h, buf), buf, (w, h)))
        submitted_at[0] = time.perf_counter() - t0
        tickets.put(None)

    def waiter():
        while True:
            item = tickets.get()
            if item is None:
                return
            due, t, buf, size = item
            r = srv.wait(t)
            lat.append((time.perf_counter() - due) * 1e3)
            server_ms.append(r.total_ms)
            batch_views.append(r.batch_views)
            pools[size].put(buf)

    th = [threading.Thread(target=submitter), threading.Thread(target=waiter)]
    for x in th:
        x.start()
    for x in th:
        x.join()
    span = max(time.perf_counter() - t0, 1e-9)
    open_loop = {"requests": len(reqs), "offered_per_s": rate,
                 "submitted_per_s": round(len(reqs) / max(submitted_at[0], 1e-9), 1),
                 "completed_per_s": round(len(lat) / span, 1),
                 "latency_ms": pct(lat), "server_ms": pct(server_ms),
                 "mean_batch_views": round(float(np.mean(batch_views)), 2)}
    open_loop["regrowth"], st1 = regrowth(st0)

    # Closed loop: 64 clients.
    done, clat, cserver, cviews = [0], [], [], []
    lock = threading.Lock()
    stop_at = time.perf_counter() + closed_seconds

    def client(k):
        buf = A.PinnedImage(1920, 1080)
        i = k
        while time.perf_counter() < stop_at:
            sid, cam, w, h = cams[i % len(cams)]
            ts = time.perf_counter()
            r = srv.wait(srv.submit(sid, cam, w, h, buf))
            with lock:
                clat.append((time.perf_counter() - ts) * 1e3)
                cserver.append(r.total_ms)
                cviews.append(r.batch_views)
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
              "latency_ms": pct(clat), "server_ms": pct(cserver),
              "mean_batch_views": round(float(np.mean(cviews)), 2) if cviews else 0.0}
    closed["regrowth"], st = regrowth(st1)
    srv.close()
    return {"scenes": len(dev), "splats": int(sum(d.count for d in dev)), "geometry_path": geometry_path,
            "resident_scene_bytes": int(st.resident_scene_bytes), "memory_budget_bytes": int(st.memory_budget_bytes),
            "pipelines": pipelines, "max_batch_views": max_batch,
            "poisson": open_loop, "closed64": closed, "batches": int(st.batches),
            "rejected": int(st.rejected)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--plain-only", action="store_true", help="skip the batched and served sections")
    ap.add_argument("--geometry-path", type=int, default=0, choices=(0, 1),
                    help="device form of the plain section's scene: 0 fp32 SoA (the headline), 1 packed records")
    ap.add_argument("--served-requests", type=int, default=2000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (PMC counter calibration)")
    ap.add_argument("--summarize-profiles", nargs=5, metavar=("TRACE", "FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="write profiles/ summaries from rocprofv3 output directories and exit")
    ap.add_argument("--profiles-out", default=str(ROOT / "profiles"), help="folder for --summarize-profiles")
    ap.add_argument("--summarize-counters", default=None, metavar="DIR",
                    help=f"write profiles/sq_counters_plain-1080p-300k-sh3.md from a rocprofv3 --pmc run of the "
                         f"plain workload with the counters {SQ_COUNTERS!r} and exit")
    ap.add_argument("--summarize-lds-counters", default=None, metavar="DIR",
                    help="summarise a rocprofv3 --pmc directory of the plain workload with the counters "
                         f"{LDS_COUNTERS!r} (LDS utilisation and VALU busy per kernel) and exit")
    ap.add_argument("--profile-batch", default=None, metavar="WORKLOAD",
                    help=f"only render --steps batches of one workload ({', '.join(sorted(BATCH_WORKLOADS))}), "
                         "for rocprofv3")
    ap.add_argument("--summarize-batch-profiles", nargs=4, metavar=("WORKLOAD", "TRACE", "FETCH", "WRITE"),
                    help="write profiles/kernels_WORKLOAD.md from rocprofv3 runs of --profile-batch and exit")
    ap.add_argument("--artifact-dir", default=None,
                    help="also write the plain workload as a gsplat-benchmark/v1 run directory (120 frames)")
    args = ap.parse_args()
    if args.summarize_profiles:
        summarize_profiles(*args.summarize_profiles, frames=0, out_dir=args.profiles_out)
        return
    if args.summarize_counters:
        summarize_counters(args.summarize_counters, args.profiles_out, SQ_COUNTERS)
        return
    if args.summarize_lds_counters:
        summarize_counters(args.summarize_lds_counters, args.profiles_out, LDS_COUNTERS, lds=True)
        return
    if args.summarize_batch_profiles:
        summarize_batch_profiles(*args.summarize_batch_profiles, out_dir=args.profiles_out, steps=args.steps,
                                 warmup=args.warmup)
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported by this build", file=sys.stderr)
        sys.exit(2)
    if not A.device_available():
        print("no gfx950 device visible", file=sys.stderr)
        sys.exit(1)
    if args.profile_batch:
        profile_batch(args.profile_batch, args.steps, args.warmup)
        return

    scene, view = synth.plain_1080p()
    cam = A.look_at(view.position, view.target)
    ctx = A.Context(view.width, view.height)
    ctx.set_geometry_path(args.geometry_path)
    ctx.load(scene)
    ctx.set_camera(cam)
    # The timed frames carry no per-stage timer events (a context records
    # them by default for its stats; the rate with them is reported beside).
    ctx.set_stage_timing(False)
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

    # Per-frame latency (render + wait, one frame at a time).
    lat = []
    for _ in range(min(args.steps, 60)):
        t = time.perf_counter()
        ctx.render()
        ctx.wait()
        lat.append((time.perf_counter() - t) * 1e3)
    lat.sort()
    # Stage device times, from frames with an event at each stage boundary
    # (each costs the device a drain, so the timed frames above have none).
    ctx.set_stage_timing(True)
    stages = []
    for _ in range(min(args.steps, 60)):
        ctx.render()
        ctx.wait()
        s = ctx.stage_times()
        stages.append((s.project_ms, s.sort_ms, s.bin_ms, s.raster_ms, s.total_ms))
    # the context's default: the same back-to-back frames with stage events
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.render()
    ctx.wait()
    fps_stage_events = args.steps / (time.perf_counter() - t0)
    ctx.set_stage_timing(False)
    st = ctx.stage_times()
    med = np.median(np.array(stages), axis=0)
    names = ["project", "sort", "bin", "raster"]
    dom = int(np.argmax(med[:4]))

    n, v = st.splat_count, st.visible_count
    pairs = int(st.tile_pairs)
    bin_pairs = int(st.bin_pairs)
    bin_px = 16 << st.bin_shift
    bin_tiles = -(-view.width // bin_px) * -(-view.height // bin_px)
    pair_bits = max((bin_tiles - 1).bit_length(), 1)
    pair_passes = 1 if pair_bits == 9 else -(-pair_bits // 8)  # 9-bit keys: one 9-bit pass
    if args.artifact_dir:
        write_plain_artifact(ctx, scene, view, args.artifact_dir)
    ctx.close()

    # One counting frame (outside the timed region) for the blend's algorithmic work E.
    os.environ["GSPLAT_COUNT_EVALS"] = "1"
    cctx = A.Context(view.width, view.height)
    cctx.load(scene)
    cctx.set_camera(cam)
    cctx.render()
    cst = cctx.stage_times()
    evals, tests = int(cst.blend_evals), int(cst.blend_tests)
    cctx.close()
    del os.environ["GSPLAT_COUNT_EVALS"]

    k = 3 * ((scene.sh_degree + 1) ** 2 - 1)
    # Algorithmic work per stage (units as stated; per frame):
    #  project  positions for all splats, covariance + alpha + cut + SH for visible, key/value writes for
    #           every slot, 48-byte record + 8-byte rect for visible
    #  sort     bucket sort (one view of <= 2^19 slots): key read for the bucket counts, one stable pass and
    #           the per-bucket LDS sort (key+value read and write each); else 8-bit LSD: 4 passes x (key+value
    #           read and write) over N slots
    #  raster   16 FLOP per (pixel, splat) evaluation passing coverage before the pixel's early stop
    proj_bytes = 12 * n + 8 * n + (24 + 4 + 4 + 12 + 4 * k + 48 + 8) * v
    sort_bytes = (4 + 16 + 16) * n if st.depth_sort == 2 else 4 * 16 * n
    #  bin      value + rect read per visible splat (counting, then emission), (binning tile, splat) pair
    #           writes, pair radix passes x (key+value read and write), one key read for the tile ranges
    pair_bytes = 2 * 12 * v + 8 * bin_pairs + pair_passes * 16 * bin_pairs + 4 * bin_pairs
    raster_flop = 16 * evals
    traffic = load_traffic()
    rooflines = {
        "k_project": {"bound": "hbm", "achieved": proj_bytes / (med[0] * 1e-3) / 1e9, "peak": MI355X_HBM_GBPS,
                      "unit": "GB/s", "algorithmic": proj_bytes},
        "depth sort": {"bound": "hbm", "achieved": sort_bytes / (med[1] * 1e-3) / 1e9,
                             "peak": MI355X_HBM_GBPS, "unit": "GB/s", "algorithmic": sort_bytes},
        "bin(pair sort+emit)": {"bound": "hbm", "achieved": pair_bytes / (med[2] * 1e-3) / 1e9,
                                "peak": MI355X_HBM_GBPS, "unit": "GB/s", "algorithmic": pair_bytes},
        "k_raster": {"bound": "valu", "achieved": raster_flop / (med[3] * 1e-3) / 1e12,
                     "peak": MI355X_FP32_TFLOPS, "unit": "TFLOP/s", "algorithmic": raster_flop},
    }
    util = load_utilisation()
    for name, r in rooflines.items():
        r["achieved"] = round(r["achieved"], 3)
        r["frac"] = round(r["achieved"] / r["peak"], 4)
        r["traffic"] = traffic.get(name)
    # the raster's issue limits, from the committed LDS/VALU counter pass of this workload
    if util.get("k_raster<false, 1>"):
        rooflines["k_raster"].update(util["k_raster<false, 1>"])
    dom_kernel = STAGE_KERNEL[names[dom]]
    roof = dict(rooflines[dom_kernel])
    roof = {"kernel": dom_kernel, "bound": roof["bound"], "achieved": roof["achieved"],
            "peak": roof["peak"], "unit": roof["unit"], "frac": roof["frac"], "traffic": roof["traffic"],
            **{k: roof[k] for k in ("lds_util", "valu_busy") if k in roof}}
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
        "config": {"workload": "plain-1080p-300k-sh3", "geometry_path": args.geometry_path, "splats": n,
                   "depth_sort": ["lookback radix", "reduce-then-scan radix", "bucket", "reused order"][st.depth_sort],
                   "width": view.width, "height": view.height,
                   "visible": v, "tile_pairs": pairs, "bin_pairs": bin_pairs, "bin_tile_px": bin_px,
                   "pair_sort_passes": pair_passes, "blend_evals": evals, "blend_tests": tests},
        "latency_ms": {"p50": round(nearest_rank(lat, 0.5), 4), "p99": round(nearest_rank(lat, 0.99), 4)},
        "frames_per_s_with_stage_events": round(fps_stage_events, 3),
        "stage_ms": {nm: round(float(x), 4) for nm, x in zip(names, med[:4])},
        "dominant_stage": names[dom],
        "mpix_per_s": round(fps * view.width * view.height / 1e6, 1),
        "msplats_sorted_per_s": round(v / (med[1] * 1e-3) / 1e6, 1),
        "roofline": roof,
        "rooflines": rooflines,
    }
    if not args.plain_only:
        s0, v0 = synth.plain_small()
        line["plain_640x480_16k"] = bench_plain_config("plain-640x480-16k-sh0", s0, v0, args.steps, args.warmup)
        print(f"plain_640x480_16k {line['plain_640x480_16k']}", file=sys.stderr, flush=True)
        del s0
        s2, v2 = synth.plain_4k()
        line["plain_4k_3m"] = bench_plain_config("plain-4k-3m-sh3-heavy-tail", s2, v2, max(args.steps // 10, 10),
                                                 max(args.warmup // 2, 3))
        print(f"plain_4k_3m {line['plain_4k_3m']}", file=sys.stderr, flush=True)
        del s2
        line["batched_1080p"] = bench_batched_1080p(scene, max(args.steps // 8, 5), max(args.warmup // 4, 2))
        print(f"batched_1080p {line['batched_1080p']}", file=sys.stderr, flush=True)
        line["batched_64v"] = bench_batched_64v(max(args.steps // 20, 5), max(args.warmup // 4, 2))
        print(f"batched_64v {line['batched_64v']}", file=sys.stderr, flush=True)
        line["served"] = bench_served(args.served_requests, 5000.0, 2.0)
        print(f"served {line['served']}", file=sys.stderr, flush=True)
        line["session"] = bench_session(scene, view, max(args.steps, 60))
        print(f"session {line['session']}", file=sys.stderr, flush=True)
        line["packed"] = bench_packed(scene, view, cam, args.steps, args.warmup, med[0])
        print(f"packed {line['packed']}", file=sys.stderr, flush=True)
        line["served_packed"] = bench_served(args.served_requests, 5000.0, 2.0,
                                             geometry_path=A.GEOMETRY_PATH_PACKED_ATLAS)
        print(f"served_packed {line['served_packed']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(scene, view, cam)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
