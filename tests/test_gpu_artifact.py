"""A gsplat-benchmark/v1 run measured on the device through the C ABI context
(the reference runner's loop, tools/bench-runner/src/main.rs:100-139) passes
the v1 checker; with stage timing off the stage timings are null and listed
as unavailable, never zero."""
import pytest

import bench_artifact as BA
import synth

pytestmark = pytest.mark.gpu


def manifest(run_id, unavailable):
    env, env_un = BA.environment_info()
    return {"schema": BA.SCHEMA, "record_type": "manifest", "run_id": run_id,
            "identity": {"series_id": "gpu-test", "started_at_utc": BA.utc_now(), "ended_at_utc": BA.utc_now(),
                         "measurement_started_at_utc": BA.utc_now(), "measurement_ended_at_utc": BA.utc_now()},
            "build": {"repository_commit": None, "dirty": None, "profile": "release", "package_version": "0.1.0"},
            "dataset": {"id": "tiny", "sha256": BA.sha256_bytes(b"tiny"), "bytes": 4, "splat_count": 3000,
                        "sh_degree": 3},
            "trace": {"id": "static", "sha256": BA.sha256_bytes(b"static")},
            "renderer": {"implementation": "gsplat-rs_amd", "path": "sorted_alpha_tiles", "backend": "hip-gfx950",
                         "sort_policy": "gpu_every_frame"},
            "display": {"width": 160, "height": 120, "dpr": 1.0, "refresh_hz": 60.0, "frame_budget_ms": 16.7,
                        "refresh_hz_source": "configured", "frame_budget_source": "configured"},
            "environment": env,
            "unavailable_fields": env_un + ["build.repository_commit", "build.dirty", "frames[*].sort_refreshed"]
            + unavailable}


@pytest.mark.parametrize("stage_timing", [True, False])
def test_measured_run_checks(abi, tmp_path, stage_timing):
    ctx = abi.Context(160, 120)
    ctx.load(synth.tiny_scene(3000, 515, sh_degree=3))
    ctx.set_camera(abi.look_at((0.2, 0.1, -2.5)))
    ctx.set_stage_timing(stage_timing)
    frames, m0, m1 = BA.measure_context(ctx, 6, 2, "gpu-run")
    ctx.close()
    assert all(f["visible"] == f["drawn"] > 0 for f in frames)
    if stage_timing:
        assert all(f["sort_ms"] is not None and f["sort_ms"] > 0 for f in frames)
        extra = []
    else:
        assert all(f["sort_ms"] is None for f in frames)
        extra = ["frames[*].preprocess_ms", "frames[*].sort_ms", "frames[*].geometry_submit_ms"]
    BA.write_run(tmp_path / "run", manifest("gpu-run", extra), frames, 2)
    BA.check_run(tmp_path / "run")
