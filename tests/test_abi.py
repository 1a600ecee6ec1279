"""C ABI contract (gsplat.h / gsplat_ext.h) that needs no GPU: symbols,
defaults, error strings, NULL handling (crates/gsplat-ffi-c/src/lib.rs:1806-2071)."""

import ctypes
import pathlib
import re

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names |= set(re.findall(r"\b(gsplat_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol(abi):
    lib = abi.lib()
    names = declared_functions()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(abi.SIGNATURES), set(names) ^ set(abi.SIGNATURES)


def test_struct_layouts(abi):
    assert ctypes.sizeof(abi.GsplatConfig) == 12
    assert ctypes.sizeof(abi.GsplatCamera) == 40
    assert ctypes.sizeof(abi.GsplatStats) == 24
    assert ctypes.sizeof(abi.GsplatSurfaceSortStats) == 48


def test_version(abi):
    assert abi.lib().gsplat_version_major() == 0
    assert abi.lib().gsplat_version_minor() == 1


def test_default_ffi_values_match_release_contract(abi):
    # ffi-c/src/lib.rs:1806-1817
    c = abi.lib().gsplat_config_default()
    assert (c.width, c.height, c.mode) == (1280, 720, 0)
    cam = abi.lib().gsplat_camera_default()
    assert list(cam.rotation_xyzw) == [0, 0, 0, 1]
    assert list(cam.position) == [0, 0, 0]
    assert cam.near_plane == ctypes.c_float(0.01).value
    assert cam.far_plane == 1000.0
    import numpy as np
    assert np.float32(cam.vertical_fov_radians) == np.float32(60.0) * (np.float32(np.pi) / np.float32(180.0))


def test_error_message_describes_known_and_unknown_codes(abi):
    m = abi.lib().gsplat_error_message
    assert [m(c).decode() for c in (0, 1, 2, 3, 4, 5, 100)] == [
        "ok", "invalid argument", "not found", "parse failed", "unsupported", "scene not loaded", "internal error"]
    assert m(-1).decode() == "unknown error"
    assert m(6).decode() == "unknown error"


def test_last_error_message_tracks_recent_failure(abi):
    rc = abi.lib().gsplat_context_load_scene_path(None, None)
    assert rc == abi.INVALID_ARGUMENT
    assert "gsplat_context_load_scene_path" in abi.last_error()


def test_context_create_rejects_null_out_pointer(abi):
    assert abi.lib().gsplat_context_create(abi.lib().gsplat_config_default(), None) == abi.INVALID_ARGUMENT


def test_context_create_rejects_non_release_render_mode(abi):
    cfg = abi.lib().gsplat_config_default()
    cfg.mode = 1
    h = abi.P(1234)
    assert abi.lib().gsplat_context_create(cfg, ctypes.byref(h)) == abi.INVALID_ARGUMENT
    assert not h.value
    assert abi.last_error() == "gsplat_context_create: unsupported render mode"


def test_context_create_rejects_zero_size(abi):
    cfg = abi.lib().gsplat_config_default()
    cfg.width = 0
    h = abi.P()
    assert abi.lib().gsplat_context_create(cfg, ctypes.byref(h)) == abi.INVALID_ARGUMENT


def test_context_functions_reject_null_handles_and_outputs(abi):
    lib = abi.lib()
    st = abi.GsplatStats()
    assert lib.gsplat_context_set_auto_camera(None) == abi.INVALID_ARGUMENT
    assert lib.gsplat_context_load_scene_path(None, None) == abi.INVALID_ARGUMENT
    assert lib.gsplat_context_render_frame(None) == abi.INVALID_ARGUMENT
    assert lib.gsplat_context_get_stats(None, ctypes.byref(st)) == abi.INVALID_ARGUMENT
    assert lib.gsplat_context_get_stats(None, None) == abi.INVALID_ARGUMENT
    assert lib.gsplat_context_set_camera(None, lib.gsplat_camera_default()) == abi.INVALID_ARGUMENT
    lib.gsplat_context_destroy(None)  # accepts NULL


def test_surface_functions_reject_null_renderer(abi):
    lib = abi.lib()
    st = abi.GsplatStats()
    inv = abi.INVALID_ARGUMENT
    assert lib.gsplat_surface_renderer_resize(None, 640, 480) == inv
    for name in ["set_sort_interval", "set_gpu_preproject", "set_gpu_preproject_double_buffer", "set_async_sort",
                 "set_async_geometry", "set_geometry_path", "set_instance_buffer_count", "set_frame_latency"]:
        assert getattr(lib, "gsplat_surface_renderer_" + name)(None, 1) == inv, name
    assert lib.gsplat_surface_renderer_reset_camera(None) == inv
    assert lib.gsplat_surface_renderer_orbit(None, 0.1, 0.1) == inv
    assert lib.gsplat_surface_renderer_zoom(None, 1.1) == inv
    assert lib.gsplat_surface_renderer_pan(None, 0.1, 0.1) == inv
    assert lib.gsplat_surface_renderer_render_frame(None) == inv
    assert lib.gsplat_surface_renderer_get_stats(None, ctypes.byref(st)) == inv
    assert lib.gsplat_surface_renderer_get_stats(None, None) == inv
    lib.gsplat_surface_renderer_destroy(None)


def test_surface_renderer_is_unsupported_on_headless_device(abi):
    out = abi.P()
    rc = abi.lib().gsplat_surface_renderer_create_android(abi.P(1), b"x.ply", 64, 64, ctypes.byref(out))
    assert rc == abi.UNSUPPORTED and not out.value


def test_look_at_identity_and_normalised(abi):
    # ffi-c/src/lib.rs:1869-1896: camera at target - d*z looks along +z with identity rotation
    c = abi.look_at((1.0, 2.0, -7.0), (1.0, 2.0, 3.0))
    assert list(c.rotation_xyzw) == [0, 0, 0, 1]
    q = abi.look_at((2.0, 1.0, -4.0), (0.0, 0.0, 0.0)).rotation_xyzw
    assert abs(sum(v * v for v in q) - 1.0) < 1e-4


@pytest.mark.skipif("gpu" in __import__("os").environ.get("PYTEST_ADDOPTS", ""), reason="cpu-only check")
def test_create_without_device_reports_unsupported(abi):
    if abi.device_available():
        pytest.skip("a device is visible")
    h = abi.P()
    rc = abi.lib().gsplat_context_create(abi.lib().gsplat_config_default(), ctypes.byref(h))
    assert rc == abi.UNSUPPORTED and not h.value
    assert abi.last_error() == "gsplat_context_create: gpu rasterizer unavailable"
