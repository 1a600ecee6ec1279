"""gsplat-benchmark/v1 run directories: measurement loop, writer and checker.

The format is the reference's benchmark artifact contract
(tests/perf/benchmark-artifact-v1.md): a run directory holds manifest.json,
frames.jsonl and summary.json, every record tagged with the schema, its
record type and one run id. A timing that cannot be observed is null and its
path is listed in manifest.unavailable_fields; it is never written as zero.
Summary distributions use nearest-rank percentiles and a left-to-right mean;
a missed frame is one whose frame_wall_ms exceeds the frame budget.

Per-frame values come from the boundaries the reference's runner uses
(tools/bench-runner/src/main.rs:107-139): call_ms is the render call,
frame_wall_ms the call plus the wait for the device, preprocess_ms and sort_ms
the frame stats, geometry_submit_ms the stats' raster time, gpu_wait_ms the
wait alone and gpu_complete_ms the whole frame.
"""

from __future__ import annotations

import datetime
import hashlib
import json
import math
import os
import pathlib
import platform
import re
import subprocess
import time
from typing import Any, Iterable, Optional

SCHEMA = "gsplat-benchmark/v1"
TIMING_FIELDS = ("call_ms", "frame_wall_ms", "preprocess_ms", "sort_ms", "geometry_submit_ms", "gpu_wait_ms",
                 "gpu_complete_ms")
ALWAYS_MEASURED = ("call_ms", "frame_wall_ms")
PERCENTILES = (("p50", 0.50), ("p90", 0.90), ("p95", 0.95), ("p99", 0.99))
TOLERANCE = 1e-9
_SHA256 = re.compile(r"^[0-9a-f]{64}$")


class ArtifactError(ValueError):
    """A run directory that breaks the v1 contract (the message names the rule)."""


# ----------------------------------------------------------------- statistics

def nearest_rank(ascending: list, fraction: float) -> float:
    """Value of 1-based rank max(ceil(fraction * n), 1) of ascending values
    (benchmark-artifact-v1.md, Summary; tools/bench-runner/src/artifact.rs:433-436)."""
    if not ascending:
        raise ArtifactError("nearest rank of no values")
    rank = max(math.ceil(fraction * len(ascending)), 1)
    return ascending[min(rank, len(ascending)) - 1]


def distribution(values: Iterable[Optional[float]]) -> Optional[dict]:
    """Summary distribution of the measured (non-null) values; None when none is measured."""
    measured = [v for v in values if v is not None]
    if not measured:
        return None
    total = 0.0
    for v in measured:  # left to right, in frame order
        total += float(v)
    ordered = sorted(measured)
    out = {"count": len(measured), "mean": total / len(measured)}
    for name, fraction in PERCENTILES:
        out[name] = nearest_rank(ordered, fraction)
    out["max"] = ordered[-1]
    return out


def _check_timing(name: str, value: Any) -> None:
    if value is None:
        return
    if isinstance(value, bool) or not isinstance(value, (int, float)) or not math.isfinite(value) or value < 0:
        raise ArtifactError(f"{name} must be a finite non-negative number or null")


def frame_record(run_id: str, frame_index: int, elapsed_ns: int, *, call_ms: float, frame_wall_ms: float,
                 preprocess_ms: Optional[float], sort_ms: Optional[float], geometry_submit_ms: Optional[float],
                 gpu_wait_ms: Optional[float], gpu_complete_ms: Optional[float], visible: int, drawn: int,
                 sort_refreshed: Optional[bool] = None) -> dict:
    rec = {"schema": SCHEMA, "record_type": "frame", "run_id": run_id, "frame_index": int(frame_index),
           "elapsed_ns": int(elapsed_ns), "call_ms": call_ms, "frame_wall_ms": frame_wall_ms,
           "preprocess_ms": preprocess_ms, "sort_ms": sort_ms, "geometry_submit_ms": geometry_submit_ms,
           "gpu_wait_ms": gpu_wait_ms, "gpu_complete_ms": gpu_complete_ms, "visible": int(visible),
           "drawn": int(drawn), "sort_refreshed": sort_refreshed}
    for name in TIMING_FIELDS:
        _check_timing(f"frame {frame_index} {name}", rec[name])
    for name in ALWAYS_MEASURED:
        if rec[name] is None:
            raise ArtifactError(f"frame {frame_index} {name} is required")
    return rec


def summary(run_id: str, frames: list, warmup_count: int, frame_budget_ms: float) -> dict:
    """The summary record of `frames` (artifact.rs:214-265 computes the same)."""
    if not run_id or not frames:
        raise ArtifactError("a summary needs a run id and at least one frame")
    if not (isinstance(frame_budget_ms, (int, float)) and math.isfinite(frame_budget_ms) and frame_budget_ms > 0):
        raise ArtifactError("frame budget must be finite and positive")
    for i, f in enumerate(frames):
        if f["run_id"] != run_id or f["frame_index"] != i:
            raise ArtifactError("frames must share the run id and count from zero without gaps")
        if i and f["elapsed_ns"] < frames[i - 1]["elapsed_ns"]:
            raise ArtifactError("elapsed_ns must be monotonic")
    return {"schema": SCHEMA, "record_type": "summary", "run_id": run_id, "sample_count": len(frames),
            "warmup_count": int(warmup_count), "frame_budget_ms": frame_budget_ms,
            "missed_frame_count": sum(1 for f in frames if f["frame_wall_ms"] > frame_budget_ms),
            "distributions": {name: distribution(f[name] for f in frames) for name in TIMING_FIELDS}}


# --------------------------------------------------------------------- writer

def utc_now() -> str:
    now = datetime.datetime.now(datetime.timezone.utc)
    return now.strftime("%Y-%m-%dT%H:%M:%S.") + f"{now.microsecond // 1000:03d}Z"


def sha256_bytes(data: bytes) -> str:
    return hashlib.sha256(data).hexdigest()


def build_info(repo: pathlib.Path) -> tuple:
    """(build object, unavailable paths): git commit and dirty state when git answers."""
    unavailable = []
    commit = dirty = None
    try:
        commit = subprocess.run(["git", "-C", str(repo), "rev-parse", "HEAD"], capture_output=True, text=True,
                                timeout=10, check=True).stdout.strip() or None
        status = subprocess.run(["git", "-C", str(repo), "status", "--porcelain"], capture_output=True, text=True,
                                timeout=10, check=True).stdout
        dirty = bool(status.strip())
    except (OSError, subprocess.SubprocessError):
        commit = dirty = None
    if commit is None:
        unavailable.append("build.repository_commit")
    if dirty is None:
        unavailable.append("build.dirty")
    return {"repository_commit": commit, "dirty": dirty, "profile": "release", "package_version": "0.1.0"}, unavailable


def environment_info(adapter: Optional[str] = None, driver: Optional[str] = None) -> tuple:
    """(environment object, unavailable paths) of this host."""
    device = None
    try:
        for line in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                device = line.split(":", 1)[1].strip()
                break
    except OSError:
        device = None
    env = {"platform": platform.machine() or "unknown", "os": f"{platform.system()} {platform.release()}".strip(),
           "device": device, "browser": None, "adapter": adapter, "driver": driver}
    unavailable = [f"environment.{k}" for k in ("device", "browser", "adapter", "driver") if env[k] is None]
    return env, unavailable


def write_run(directory, manifest: dict, frames: list, warmup_count: int) -> dict:
    """Writes a complete run directory, or nothing: the files go to a staging
    directory beside it, which is renamed into place. An existing directory is
    never reused. Returns the summary record."""
    directory = pathlib.Path(directory)
    if directory.exists():
        raise ArtifactError(f"{directory} exists; a run directory is written once")
    run_id = manifest.get("run_id", "")
    summ = summary(run_id, frames, warmup_count, manifest["display"]["frame_budget_ms"])
    missing = [name for name in TIMING_FIELDS if any(f[name] is None for f in frames)
               and f"frames[*].{name}" not in manifest["unavailable_fields"]]
    if missing:
        raise ArtifactError(f"null timings not listed as unavailable: {missing}")
    staging = directory.parent / f".{directory.name}.staging-{os.getpid()}"
    staging.mkdir(parents=True)
    dump = lambda obj: json.dumps(obj, allow_nan=False)  # noqa: E731 -- non-finite numbers are refused
    (staging / "manifest.json").write_text(json.dumps(manifest, indent=2, allow_nan=False) + "\n")
    (staging / "frames.jsonl").write_text("".join(dump(f) + "\n" for f in frames))
    (staging / "summary.json").write_text(json.dumps(summ, indent=2, allow_nan=False) + "\n")
    staging.rename(directory)
    return summ


def measure_context(ctx, iterations: int, warmup: int, run_id: str) -> tuple:
    """The reference runner's loop over a context (main.rs:100-139): warmup
    frames, then per frame the render call and the wait, each timed on the
    host, with the frame's stats. Returns (frames, measurement start UTC,
    measurement end UTC)."""
    if iterations < 1:
        raise ArtifactError("at least one measured frame")
    for _ in range(warmup):
        ctx.render()
        ctx.wait()
    started = utc_now()
    t0 = time.perf_counter_ns()
    frames = []
    for i in range(iterations):
        f0 = time.perf_counter_ns()
        ctx.render()
        call_ms = (time.perf_counter_ns() - f0) / 1e6
        ctx.wait()
        wall_ms = (time.perf_counter_ns() - f0) / 1e6
        st = ctx.stats()
        timed = st.preprocess_ms > 0 or st.sort_ms > 0 or st.raster_ms > 0  # stage timing on
        frames.append(frame_record(
            run_id, i, time.perf_counter_ns() - t0, call_ms=call_ms, frame_wall_ms=wall_ms,
            preprocess_ms=float(st.preprocess_ms) if timed else None, sort_ms=float(st.sort_ms) if timed else None,
            geometry_submit_ms=float(st.raster_ms) if timed else None, gpu_wait_ms=max(wall_ms - call_ms, 0.0),
            gpu_complete_ms=wall_ms, visible=st.visible_count, drawn=st.drawn_count, sort_refreshed=None))
    return frames, started, utc_now()


# -------------------------------------------------------------------- checker

def _obj(parent: dict, key: str, where: str) -> dict:
    v = parent.get(key)
    if not isinstance(v, dict):
        raise ArtifactError(f"{where}{key} must be an object")
    return v


def _str(parent: dict, key: str, where: str) -> str:
    v = parent.get(key)
    if not isinstance(v, str) or not v:
        raise ArtifactError(f"{where}{key} must be a non-empty string")
    return v


def _int(parent: dict, key: str, where: str) -> int:
    v = parent.get(key)
    if isinstance(v, bool) or not isinstance(v, int) or v < 0:
        raise ArtifactError(f"{where}{key} must be a non-negative integer")
    return v


def _num(parent: dict, key: str, where: str) -> float:
    v = parent.get(key)
    if isinstance(v, bool) or not isinstance(v, (int, float)) or not math.isfinite(v) or v <= 0:
        raise ArtifactError(f"{where}{key} must be a finite positive number")
    return v


def _load(path: pathlib.Path) -> Any:
    def refuse(token):
        raise ArtifactError(f"{path.name}: non-finite number {token}")
    try:
        return json.loads(path.read_text(), parse_constant=refuse)
    except (OSError, json.JSONDecodeError) as e:
        raise ArtifactError(f"{path.name}: {e}") from None


def _header(rec: Any, record_type: str, run_id: Optional[str], where: str) -> str:
    if not isinstance(rec, dict):
        raise ArtifactError(f"{where} must be a JSON object")
    if rec.get("schema") != SCHEMA:
        raise ArtifactError(f"{where} schema must be {SCHEMA}")
    if rec.get("record_type") != record_type:
        raise ArtifactError(f"{where} record_type must be {record_type}")
    rid = _str(rec, "run_id", f"{where} ")
    if run_id is not None and rid != run_id:
        raise ArtifactError(f"{where} run_id differs from the manifest's")
    return rid


def check_manifest(m: Any) -> tuple:
    run_id = _header(m, "manifest", None, "manifest")
    unavailable = m.get("unavailable_fields")
    if not isinstance(unavailable, list) or not all(isinstance(u, str) and u for u in unavailable):
        raise ArtifactError("unavailable_fields must be a list of non-empty strings")
    unavailable = set(unavailable)
    ident = _obj(m, "identity", "")
    for k in ("series_id", "started_at_utc", "ended_at_utc", "measurement_started_at_utc",
              "measurement_ended_at_utc"):
        _str(ident, k, "identity.")
    build = _obj(m, "build", "")
    for k, kind in (("repository_commit", str), ("dirty", bool)):
        v = build.get(k)
        if v is None:
            if f"build.{k}" not in unavailable:
                raise ArtifactError(f"null build.{k} must be listed in unavailable_fields")
        elif not isinstance(v, kind) or (kind is str and not v):
            raise ArtifactError(f"build.{k} has the wrong type")
    for k in ("profile", "package_version"):
        _str(build, k, "build.")
    data = _obj(m, "dataset", "")
    _str(data, "id", "dataset.")
    if not _SHA256.match(_str(data, "sha256", "dataset.")):
        raise ArtifactError("dataset.sha256 must be lowercase hex SHA-256")
    for k in ("bytes", "splat_count", "sh_degree"):
        _int(data, k, "dataset.")
    trace = _obj(m, "trace", "")
    _str(trace, "id", "trace.")
    if not _SHA256.match(_str(trace, "sha256", "trace.")):
        raise ArtifactError("trace.sha256 must be lowercase hex SHA-256")
    rend = _obj(m, "renderer", "")
    for k in ("implementation", "path", "backend", "sort_policy"):
        _str(rend, k, "renderer.")
    disp = _obj(m, "display", "")
    for k in ("width", "height"):
        if _int(disp, k, "display.") == 0:
            raise ArtifactError(f"display.{k} must be positive")
    for k in ("dpr", "refresh_hz", "frame_budget_ms"):
        _num(disp, k, "display.")
    for k in ("refresh_hz_source", "frame_budget_source"):
        _str(disp, k, "display.")
    env = _obj(m, "environment", "")
    for k in ("platform", "os"):
        _str(env, k, "environment.")
    for k in ("device", "browser", "adapter", "driver"):
        if k not in env or not (env[k] is None or isinstance(env[k], str)):
            raise ArtifactError(f"environment.{k} must be a string or null")
    return run_id, unavailable


def check_run(directory) -> None:
    """Raises ArtifactError unless `directory` is a valid v1 run."""
    d = pathlib.Path(directory)
    if not d.is_dir():
        raise ArtifactError(f"{d} is not a directory")
    run_id, unavailable = check_manifest(_load(d / "manifest.json"))
    frames = []
    try:
        lines = (d / "frames.jsonl").read_text().splitlines()
    except OSError as e:
        raise ArtifactError(f"frames.jsonl: {e}") from None
    for n, line in enumerate(lines):
        if not line.strip():
            continue

        def refuse(token):
            raise ArtifactError(f"frames.jsonl line {n + 1}: non-finite number {token}")
        try:
            f = json.loads(line, parse_constant=refuse)
        except json.JSONDecodeError as e:
            raise ArtifactError(f"frames.jsonl line {n + 1}: {e}") from None
        where = f"frame {len(frames)}"
        _header(f, "frame", run_id, where)
        if _int(f, "frame_index", f"{where} ") != len(frames):
            raise ArtifactError(f"{where}: frame_index must count from zero without gaps")
        _int(f, "elapsed_ns", f"{where} ")
        for k in TIMING_FIELDS:
            if k not in f:
                raise ArtifactError(f"{where}: {k} is missing")
            _check_timing(f"{where} {k}", f[k])
            if f[k] is None:
                if k in ALWAYS_MEASURED:
                    raise ArtifactError(f"{where}: {k} cannot be null")
                if f"frames[*].{k}" not in unavailable:
                    raise ArtifactError(f"{where}: null {k} must be listed in unavailable_fields")
        for k in ("visible", "drawn"):
            _int(f, k, f"{where} ")
        if "sort_refreshed" not in f or not (f["sort_refreshed"] is None or isinstance(f["sort_refreshed"], bool)):
            raise ArtifactError(f"{where}: sort_refreshed must be a boolean or null")
        if frames and f["elapsed_ns"] < frames[-1]["elapsed_ns"]:
            raise ArtifactError(f"{where}: elapsed_ns must be monotonic")
        frames.append(f)
    s = _load(d / "summary.json")
    _header(s, "summary", run_id, "summary")
    if _int(s, "sample_count", "summary.") != len(frames):
        raise ArtifactError("summary.sample_count differs from the frames")
    _int(s, "warmup_count", "summary.")
    budget = _num(s, "frame_budget_ms", "summary.")
    if _int(s, "missed_frame_count", "summary.") != sum(1 for f in frames if f["frame_wall_ms"] > budget):
        raise ArtifactError("summary.missed_frame_count differs from the frames")
    dists = _obj(s, "distributions", "summary.")
    for k in TIMING_FIELDS:
        if k not in dists:
            raise ArtifactError(f"summary.distributions.{k} is missing")
        want = distribution(f[k] for f in frames)
        got = dists[k]
        if want is None:
            if got is not None:
                raise ArtifactError(f"summary.distributions.{k} must be null: no frame measures it")
            continue
        if not isinstance(got, dict):
            raise ArtifactError(f"summary.distributions.{k} must be an object")
        if got.get("count") != want["count"]:
            raise ArtifactError(f"summary.distributions.{k}.count differs from the frames")
        for field in ("mean", "p50", "p90", "p95", "p99", "max"):
            v = got.get(field)
            if isinstance(v, bool) or not isinstance(v, (int, float)) or abs(v - want[field]) > TOLERANCE:
                raise ArtifactError(f"summary.distributions.{k}.{field} differs from the frames")
