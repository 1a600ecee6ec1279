"""The gsplat-benchmark/v1 writer and checker (gsplat-rs_amd/bench_artifact.py).

Pinned by the reference's own cases: the five-frame nearest-rank vector and
the missed-frame/null-metric summary of tools/bench-runner/src/artifact.rs:
569-590, and the valid run tests/perf/fixtures/v1/valid (copied as data under
tests/golden/artifact_v1_valid), which the checker must accept and must
reject once any contract rule of tests/perf/benchmark-artifact-v1.md is
broken in it.
"""
import json
import math
import pathlib
import shutil

import pytest

import bench_artifact as BA

VALID = pathlib.Path(__file__).parent / "golden" / "artifact_v1_valid"


def test_nearest_rank_five_frame_vector():
    d = BA.distribution([1.0, 2.0, 3.0, 4.0, 5.0])
    assert d == {"count": 5, "mean": 3.0, "p50": 3.0, "p90": 5.0, "p95": 5.0, "p99": 5.0, "max": 5.0}


def test_nearest_rank_by_brute_force():
    # rank r (1-based) of n values is the smallest r with r / n >= fraction
    for n in range(1, 40):
        vals = [float(i) for i in range(n)]
        for _, frac in BA.PERCENTILES:
            r = next(r for r in range(1, n + 1) if r >= frac * n - 1e-12)
            assert BA.nearest_rank(vals, frac) == vals[r - 1]
    assert BA.distribution([None, None]) is None
    assert BA.distribution([None, 2.0, None])["count"] == 1


def frames_of(values, run_id="run", gpu=False):
    return [BA.frame_record(run_id, i, 1000 * i, call_ms=v, frame_wall_ms=v, preprocess_ms=v / 10, sort_ms=v / 5,
                            geometry_submit_ms=v / 4, gpu_wait_ms=(v / 2 if gpu else None),
                            gpu_complete_ms=(v if gpu else None), visible=2, drawn=2) for i, v in enumerate(values)]


def test_summary_missed_frames_and_null_metrics():
    s = BA.summary("run", frames_of([1.0, 2.0, 3.0, 4.0, 5.0]), 2, 3.5)
    assert s["sample_count"] == 5 and s["missed_frame_count"] == 2
    assert s["distributions"]["gpu_wait_ms"] is None and s["distributions"]["gpu_complete_ms"] is None
    with pytest.raises(BA.ArtifactError):
        BA.summary("run", frames_of([1.0]), 0, 0.0)
    bad = frames_of([1.0, 2.0])
    bad[1]["frame_index"] = 5
    with pytest.raises(BA.ArtifactError):
        BA.summary("run", bad, 0, 3.5)
    with pytest.raises(BA.ArtifactError):
        BA.frame_record("run", 0, 0, call_ms=math.nan, frame_wall_ms=1.0, preprocess_ms=None, sort_ms=None,
                        geometry_submit_ms=None, gpu_wait_ms=None, gpu_complete_ms=None, visible=0, drawn=0)


def test_reference_valid_run_passes():
    BA.check_run(VALID)


def mutated(tmp_path, edit_manifest=None, edit_frames=None, edit_summary=None):
    d = tmp_path / "run"
    shutil.copytree(VALID, d)
    m = json.loads((d / "manifest.json").read_text())
    frames = [json.loads(x) for x in (d / "frames.jsonl").read_text().splitlines() if x.strip()]
    s = json.loads((d / "summary.json").read_text())
    if edit_manifest:
        edit_manifest(m)
    if edit_frames:
        edit_frames(frames)
    if edit_summary:
        edit_summary(s)
    (d / "manifest.json").write_text(json.dumps(m))
    (d / "frames.jsonl").write_text("".join(json.dumps(f) + "\n" for f in frames))
    (d / "summary.json").write_text(json.dumps(s))
    return d


def _set(obj, path, value):
    keys = path.split(".")
    for k in keys[:-1]:
        obj = obj[k]
    obj[keys[-1]] = value


@pytest.mark.parametrize("case", [
    ("manifest", lambda m: _set(m, "schema", "gsplat-benchmark/v2")),
    ("manifest", lambda m: m["unavailable_fields"].remove("frames[*].gpu_wait_ms")),
    ("manifest", lambda m: _set(m, "dataset.sha256", "A" * 64)),
    ("manifest", lambda m: _set(m, "display.width", 0)),
    ("manifest", lambda m: _set(m, "build.repository_commit", None)),
    ("manifest", lambda m: m["environment"].pop("driver")),
    ("frames", lambda fs: _set(fs[2], "frame_index", 3)),
    ("frames", lambda fs: _set(fs[3], "elapsed_ns", 1)),
    ("frames", lambda fs: _set(fs[1], "call_ms", None)),
    ("frames", lambda fs: _set(fs[1], "sort_ms", -1.0)),
    ("frames", lambda fs: _set(fs[4], "run_id", "other")),
    ("frames", lambda fs: fs.pop()),
    ("summary", lambda s: _set(s, "missed_frame_count", 1)),
    ("summary", lambda s: _set(s, "distributions.call_ms.mean", 3.0001)),
    ("summary", lambda s: _set(s, "distributions.sort_ms.p90", 0.8)),
    ("summary", lambda s: _set(s, "distributions.gpu_wait_ms", {"count": 0})),
    ("summary", lambda s: s["distributions"].pop("preprocess_ms")),
])
def test_broken_runs_are_rejected(tmp_path, case):
    where, edit = case
    d = mutated(tmp_path, **{f"edit_{where}": edit})
    with pytest.raises(BA.ArtifactError):
        BA.check_run(d)


def test_nan_literal_is_rejected(tmp_path):
    d = mutated(tmp_path)
    text = (d / "frames.jsonl").read_text().replace('"call_ms": 1.0', '"call_ms": NaN', 1)
    (d / "frames.jsonl").write_text(text)
    with pytest.raises(BA.ArtifactError):
        BA.check_run(d)


def manifest_for(run_id, unavailable):
    env, env_un = BA.environment_info()
    return {"schema": BA.SCHEMA, "record_type": "manifest", "run_id": run_id,
            "identity": {"series_id": "s", "started_at_utc": BA.utc_now(), "ended_at_utc": BA.utc_now(),
                         "measurement_started_at_utc": BA.utc_now(), "measurement_ended_at_utc": BA.utc_now()},
            "build": {"repository_commit": "abc", "dirty": False, "profile": "release", "package_version": "0.1.0"},
            "dataset": {"id": "d", "sha256": BA.sha256_bytes(b"d"), "bytes": 1, "splat_count": 1, "sh_degree": 0},
            "trace": {"id": "t", "sha256": BA.sha256_bytes(b"t")},
            "renderer": {"implementation": "x", "path": "p", "backend": "b", "sort_policy": "gpu_every_frame"},
            "display": {"width": 64, "height": 64, "dpr": 1.0, "refresh_hz": 60.0, "frame_budget_ms": 3.5,
                        "refresh_hz_source": "configured", "frame_budget_source": "configured"},
            "environment": env, "unavailable_fields": env_un + unavailable}


def test_written_run_checks_and_is_written_once(tmp_path):
    frames = frames_of([1.0, 2.0, 3.0, 4.0, 5.0, 0.5], run_id="w")
    m = manifest_for("w", ["frames[*].gpu_wait_ms", "frames[*].gpu_complete_ms", "frames[*].sort_refreshed"])
    s = BA.write_run(tmp_path / "out", m, frames, 3)
    assert s["missed_frame_count"] == 2 and s["distributions"]["call_ms"]["p50"] == 2.0  # rank ceil(0.5 * 6) of 0.5 1 2 3 4 5
    BA.check_run(tmp_path / "out")
    with pytest.raises(BA.ArtifactError):
        BA.write_run(tmp_path / "out", m, frames, 3)
    # a null timing that the manifest does not list is refused before writing
    m2 = manifest_for("w", [])
    with pytest.raises(BA.ArtifactError):
        BA.write_run(tmp_path / "out2", m2, frames, 3)
    assert not (tmp_path / "out2").exists()


def test_image_parity_metric_counts_pixels_not_channels():
    # render-wgpu/src/lib.rs:3539-3545: one of two pixels has one channel 4/255 off
    import numpy as np
    from image_metrics import passes_stale_order_gate, rgba_parity_metrics

    direct = np.array([0, 0, 0, 255, 0, 0, 0, 255], np.uint8)
    packed = np.array([4, 0, 0, 255, 0, 0, 0, 255], np.uint8)
    mean, frac, mx = rgba_parity_metrics(direct, packed)
    assert frac == 0.5 and abs(mean - 4 / 255 / 6) < 1e-12 and abs(mx - 4 / 255) < 1e-12
    assert not passes_stale_order_gate(direct, packed)
    assert passes_stale_order_gate(direct, direct)
