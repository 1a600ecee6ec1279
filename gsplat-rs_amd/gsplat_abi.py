"""ctypes binding of libgsplat_ffi_c.so (include/gsplat.h + include/gsplat_ext.h).

This is the binding a Python caller of the C ABI writes; tests and bench.py
use it. Loading fails loudly when the shared library has not been built:
there is no fallback implementation.
"""

from __future__ import annotations

import ctypes
import pathlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB_PATH = HERE / "libgsplat_ffi_c.so"

OK = 0
INVALID_ARGUMENT = 1
NOT_FOUND = 2
PARSE_FAILED = 3
UNSUPPORTED = 4
SCENE_NOT_LOADED = 5
INTERNAL = 100


class GsplatConfig(ctypes.Structure):
    _fields_ = [("width", ctypes.c_uint32), ("height", ctypes.c_uint32), ("mode", ctypes.c_uint32)]


class GsplatStats(ctypes.Structure):
    _fields_ = [
        ("frame_ms", ctypes.c_float),
        ("preprocess_ms", ctypes.c_float),
        ("sort_ms", ctypes.c_float),
        ("raster_ms", ctypes.c_float),
        ("visible_count", ctypes.c_uint32),
        ("drawn_count", ctypes.c_uint32),
    ]


class GsplatSurfaceSortStats(ctypes.Structure):
    _fields_ = [
        ("camera_revision", ctypes.c_uint64),
        ("applied_order_revision", ctypes.c_uint64),
        ("scheduled_revision", ctypes.c_uint64),
        ("completed_revision", ctypes.c_uint64),
        ("presented_order_revision_lag", ctypes.c_uint32),
        ("observed_result_revision_lag", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
    ]


class GsplatCamera(ctypes.Structure):
    _fields_ = [
        ("position", ctypes.c_float * 3),
        ("rotation_xyzw", ctypes.c_float * 4),
        ("vertical_fov_radians", ctypes.c_float),
        ("near_plane", ctypes.c_float),
        ("far_plane", ctypes.c_float),
    ]

    def as_dict(self) -> dict:
        return dict(
            position=list(self.position),
            rotation_xyzw=list(self.rotation_xyzw),
            fov=self.vertical_fov_radians,
            near=self.near_plane,
            far=self.far_plane,
        )


class GsplatSceneView(ctypes.Structure):
    _fields_ = [
        ("count", ctypes.c_uint32),
        ("sh_degree", ctypes.c_uint32),
        ("positions", ctypes.c_void_p),
        ("opacity", ctypes.c_void_p),
        ("scale_xyz", ctypes.c_void_p),
        ("rotation_xyzw", ctypes.c_void_p),
        ("color_dc", ctypes.c_void_p),
        ("sh_rest", ctypes.c_void_p),
    ]


class GsplatStageTimes(ctypes.Structure):
    _fields_ = [
        ("project_ms", ctypes.c_float),
        ("sort_ms", ctypes.c_float),
        ("bin_ms", ctypes.c_float),
        ("raster_ms", ctypes.c_float),
        ("total_ms", ctypes.c_float),
        ("splat_count", ctypes.c_uint32),
        ("visible_count", ctypes.c_uint32),
        ("tiles_x", ctypes.c_uint32),
        ("tiles_y", ctypes.c_uint32),
        ("tile_pairs", ctypes.c_uint64),
        ("blend_evals", ctypes.c_uint64),
        ("blend_tests", ctypes.c_uint64),
        ("bin_pairs", ctypes.c_uint64),
        ("bin_shift", ctypes.c_uint32),
        ("depth_sort", ctypes.c_uint32),
    ]


class GsplatPlyLimits(ctypes.Structure):
    _fields_ = [
        ("max_input_bytes", ctypes.c_uint64),
        ("max_header_bytes", ctypes.c_uint64),
        ("max_vertices", ctypes.c_uint64),
        ("max_vertex_properties", ctypes.c_uint64),
        ("max_scene_bytes", ctypes.c_uint64),
    ]


class GsplatSceneInfo(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint32), ("sh_degree", ctypes.c_uint32), ("device_bytes", ctypes.c_uint64),
                ("geometry_path", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class GsplatViewDesc(ctypes.Structure):
    _fields_ = [
        ("scene", ctypes.c_void_p),
        ("camera", GsplatCamera),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
    ]


class GsplatTuning(ctypes.Structure):
    _fields_ = [("bin_shift", ctypes.c_uint32), ("raster_variant", ctypes.c_uint32),
                ("sort_rts_min_tiles", ctypes.c_uint32), ("frame_ms", ctypes.c_float),
                ("blend_min", ctypes.c_uint32), ("bin_auto", ctypes.c_uint32),
                ("rows2_min_tiles", ctypes.c_uint32)]


class GsplatServerConfig(ctypes.Structure):
    _fields_ = [("memory_budget_bytes", ctypes.c_uint64), ("max_batch_views", ctypes.c_uint32),
                ("max_queue", ctypes.c_uint32), ("pipelines", ctypes.c_uint32)]


class GsplatRenderRequest(ctypes.Structure):
    _fields_ = [
        ("scene_id", ctypes.c_uint32),
        ("camera", GsplatCamera),
        ("width", ctypes.c_uint32),
        ("height", ctypes.c_uint32),
        ("out_rgba8", ctypes.c_void_p),
        ("out_len", ctypes.c_uint64),
    ]


class GsplatRequestResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("visible_count", ctypes.c_uint32),
        ("batch_views", ctypes.c_uint32),
        ("queue_ms", ctypes.c_float),
        ("device_ms", ctypes.c_float),
        ("total_ms", ctypes.c_float),
    ]


class GsplatServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "completed", "rejected", "batches", "views_in_batches",
                                              "resident_scene_bytes", "memory_budget_bytes")] + \
               [("scenes", ctypes.c_uint32), ("queued", ctypes.c_uint32), ("overflow_rerenders", ctypes.c_uint64),
                ("buffer_grows", ctypes.c_uint64)]


P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I32 = ctypes.c_int32
F32 = ctypes.c_float
PU32 = ctypes.POINTER(ctypes.c_uint32)

# name -> (restype, argtypes); every symbol the headers declare.
SIGNATURES = {
    # gsplat.h
    "gsplat_version_major": (U32, []),
    "gsplat_version_minor": (U32, []),
    "gsplat_error_message": (ctypes.c_char_p, [I32]),
    "gsplat_last_error_message": (ctypes.c_char_p, []),
    "gsplat_config_default": (GsplatConfig, []),
    "gsplat_camera_default": (GsplatCamera, []),
    "gsplat_context_create": (I32, [GsplatConfig, ctypes.POINTER(P)]),
    "gsplat_context_destroy": (None, [P]),
    "gsplat_context_set_camera": (I32, [P, GsplatCamera]),
    "gsplat_context_set_auto_camera": (I32, [P]),
    "gsplat_context_load_scene_path": (I32, [P, ctypes.c_char_p]),
    "gsplat_context_render_frame": (I32, [P]),
    "gsplat_context_get_stats": (I32, [P, ctypes.POINTER(GsplatStats)]),
    "gsplat_context_set_sort_interval": (I32, [P, U32]),
    "gsplat_context_force_sort_refresh": (I32, [P]),
    "gsplat_context_get_sort_stats": (I32, [P, ctypes.POINTER(GsplatSurfaceSortStats)]),
    "gsplat_surface_renderer_create_android": (I32, [P, ctypes.c_char_p, U32, U32, ctypes.POINTER(P)]),
    "gsplat_surface_renderer_create_android_with_geometry_path": (
        I32, [P, ctypes.c_char_p, U32, U32, U32, ctypes.POINTER(P)]),
    "gsplat_surface_renderer_create_uikit": (I32, [P, P, ctypes.c_char_p, U32, U32, ctypes.POINTER(P)]),
    "gsplat_surface_renderer_create_uikit_with_geometry_path": (
        I32, [P, P, ctypes.c_char_p, U32, U32, U32, ctypes.POINTER(P)]),
    "gsplat_surface_renderer_destroy": (None, [P]),
    "gsplat_surface_renderer_resize": (I32, [P, U32, U32]),
    "gsplat_surface_renderer_set_sort_interval": (I32, [P, U32]),
    "gsplat_surface_renderer_set_geometry_path": (I32, [P, U32]),
    "gsplat_surface_renderer_set_gpu_preproject": (I32, [P, U32]),
    "gsplat_surface_renderer_set_gpu_preproject_double_buffer": (I32, [P, U32]),
    "gsplat_surface_renderer_set_async_sort": (I32, [P, U32]),
    "gsplat_surface_renderer_set_async_geometry": (I32, [P, U32]),
    "gsplat_surface_renderer_set_instance_buffer_count": (I32, [P, U32]),
    "gsplat_surface_renderer_set_frame_latency": (I32, [P, U32]),
    "gsplat_surface_renderer_reset_camera": (I32, [P]),
    "gsplat_surface_renderer_orbit": (I32, [P, F32, F32]),
    "gsplat_surface_renderer_zoom": (I32, [P, F32]),
    "gsplat_surface_renderer_pan": (I32, [P, F32, F32]),
    "gsplat_surface_renderer_render_frame": (I32, [P]),
    "gsplat_surface_renderer_get_stats": (I32, [P, ctypes.POINTER(GsplatStats)]),
    "gsplat_surface_renderer_get_sort_stats": (I32, [P, ctypes.POINTER(GsplatSurfaceSortStats)]),
    # gsplat_ext.h
    "gsplat_context_load_scene_memory": (I32, [P, ctypes.POINTER(GsplatSceneView)]),
    "gsplat_context_set_size": (I32, [P, U32, U32]),
    "gsplat_context_get_camera": (I32, [P, ctypes.POINTER(GsplatCamera)]),
    "gsplat_context_wait": (I32, [P]),
    "gsplat_context_readback_rgba8": (I32, [P, P, U64]),
    "gsplat_context_readback_rgba32f": (I32, [P, P, U64]),
    "gsplat_context_get_stage_times": (I32, [P, ctypes.POINTER(GsplatStageTimes)]),
    "gsplat_context_read_depth_keys": (I32, [P, P, U64]),
    "gsplat_context_read_sorted_indices": (I32, [P, P, U64, PU32]),
    "gsplat_context_read_projection": (I32, [P, P, U64]),
    "gsplat_context_read_tile_counts": (I32, [P, P, U64, PU32, PU32]),
    "gsplat_sort_values_by_keys": (I32, [P, P, U64]),
    "gsplat_sort_values_by_keys_noalloc": (I32, [P, P, U64]),
    "gsplat_ply_limits_default": (GsplatPlyLimits, []),
    "gsplat_ply_load": (I32, [ctypes.c_char_p, ctypes.POINTER(P)]),
    "gsplat_ply_parse_bytes": (I32, [P, U64, ctypes.POINTER(GsplatPlyLimits), ctypes.POINTER(P)]),
    "gsplat_ply_parse_text": (I32, [P, U64, ctypes.POINTER(GsplatPlyLimits), ctypes.POINTER(P)]),
    "gsplat_ply_scene_view": (I32, [P, ctypes.POINTER(GsplatSceneView)]),
    "gsplat_ply_free": (None, [P]),
    "gsplat_device_available": (U32, []),
    "gsplat_camera_look_at": (GsplatCamera, [ctypes.POINTER(F32), ctypes.POINTER(F32), F32, F32, F32]),
    "gsplat_scene_create": (I32, [ctypes.POINTER(GsplatSceneView), ctypes.POINTER(P)]),
    "gsplat_scene_create_from_path": (I32, [ctypes.c_char_p, ctypes.POINTER(P)]),
    "gsplat_scene_create_with_geometry_path": (I32, [ctypes.POINTER(GsplatSceneView), U32, ctypes.POINTER(P)]),
    "gsplat_scene_create_from_path_with_geometry_path": (I32, [ctypes.c_char_p, U32, ctypes.POINTER(P)]),
    "gsplat_pack_scene": (I32, [ctypes.POINTER(GsplatSceneView), P, U64, P, U64, PU32]),
    "gsplat_context_set_geometry_path": (I32, [P, U32]),
    "gsplat_context_get_geometry_path": (I32, [P, PU32]),
    "gsplat_scene_get_info": (I32, [P, ctypes.POINTER(GsplatSceneInfo)]),
    "gsplat_scene_destroy": (None, [P]),
    "gsplat_batch_create": (I32, [ctypes.POINTER(P)]),
    "gsplat_batch_destroy": (None, [P]),
    "gsplat_batch_render": (I32, [P, ctypes.POINTER(GsplatViewDesc), U32]),
    "gsplat_batch_wait": (I32, [P]),
    "gsplat_batch_readback_rgba8": (I32, [P, U32, P, U64]),
    "gsplat_batch_readback_rgba32f": (I32, [P, U32, P, U64]),
    "gsplat_batch_get_view_stats": (I32, [P, U32, ctypes.POINTER(GsplatStats)]),
    "gsplat_batch_read_sorted_indices": (I32, [P, U32, P, U64, PU32]),
    "gsplat_batch_read_tile_counts": (I32, [P, U32, P, U64, PU32, PU32]),
    "gsplat_batch_get_stage_times": (I32, [P, ctypes.POINTER(GsplatStageTimes)]),
    "gsplat_batch_get_tuning": (I32, [P, ctypes.POINTER(GsplatTuning)]),
    "gsplat_batch_set_tuning": (I32, [P, ctypes.POINTER(GsplatTuning)]),
    "gsplat_batch_autotune": (I32, [P, ctypes.POINTER(GsplatViewDesc), U32, U32, ctypes.POINTER(GsplatTuning)]),
    "gsplat_context_get_tuning": (I32, [P, ctypes.POINTER(GsplatTuning)]),
    "gsplat_context_set_stage_timing": (I32, [P, U32]),
    "gsplat_batch_set_stage_timing": (I32, [P, U32]),
    "gsplat_context_set_tuning": (I32, [P, ctypes.POINTER(GsplatTuning)]),
    "gsplat_context_autotune": (I32, [P, U32, ctypes.POINTER(GsplatTuning)]),
    "gsplat_server_config_default": (GsplatServerConfig, []),
    "gsplat_server_create": (I32, [ctypes.POINTER(GsplatServerConfig), ctypes.POINTER(P)]),
    "gsplat_server_destroy": (None, [P]),
    "gsplat_server_add_scene": (I32, [P, P, PU32]),
    "gsplat_server_remove_scene": (I32, [P, U32]),
    "gsplat_server_submit": (I32, [P, ctypes.POINTER(GsplatRenderRequest), ctypes.POINTER(U64)]),
    "gsplat_server_wait": (I32, [P, U64, ctypes.POINTER(GsplatRequestResult)]),
    "gsplat_server_get_stats": (I32, [P, ctypes.POINTER(GsplatServerStats)]),
    "gsplat_host_alloc": (P, [U64]),
    "gsplat_host_free": (I32, [P]),
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is not built; run `make -C gsplat-rs_amd` (or __graft_entry__.build())")
        l = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


class GsplatError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"[{code}] {message}")
        self.code = code
        self.message = message


def check(rc: int) -> None:
    if rc != OK:
        raise GsplatError(rc, last_error())


def last_error() -> str:
    return lib().gsplat_last_error_message().decode()


def device_available() -> bool:
    return bool(lib().gsplat_device_available())


def camera(position, rotation_xyzw=(0.0, 0.0, 0.0, 1.0), fov=None, near=0.01, far=1000.0) -> GsplatCamera:
    c = lib().gsplat_camera_default()
    c.position[:] = [float(v) for v in position]
    c.rotation_xyzw[:] = [float(v) for v in rotation_xyzw]
    if fov is not None:
        c.vertical_fov_radians = float(fov)
    c.near_plane = float(near)
    c.far_plane = float(far)
    return c


def look_at(position, target=(0.0, 0.0, 0.0), fov=None, near=0.01, far=1000.0) -> GsplatCamera:
    p = (F32 * 3)(*[float(v) for v in position])
    t = (F32 * 3)(*[float(v) for v in target])
    if fov is None:
        fov = lib().gsplat_camera_default().vertical_fov_radians
    return lib().gsplat_camera_look_at(p, t, F32(fov), F32(near), F32(far))


def scene_view(scene) -> GsplatSceneView:
    """`scene` has positions/opacity/scale_xyz/rotation_xyzw/color_dc/sh_rest numpy float32 arrays."""
    v = GsplatSceneView()
    v.count = int(scene.count)
    v.sh_degree = int(scene.sh_degree)
    v.positions = scene.positions.ctypes.data
    v.opacity = scene.opacity.ctypes.data
    v.scale_xyz = scene.scale_xyz.ctypes.data
    v.rotation_xyzw = scene.rotation_xyzw.ctypes.data
    v.color_dc = scene.color_dc.ctypes.data
    v.sh_rest = None if scene.sh_rest is None else scene.sh_rest.ctypes.data
    return v


class Context:
    """Owns one GsplatContext (one renderer on the GPU)."""

    def __init__(self, width: int = 1280, height: int = 720):
        cfg = lib().gsplat_config_default()
        cfg.width, cfg.height = width, height
        h = P()
        check(lib().gsplat_context_create(cfg, ctypes.byref(h)))
        self.h = h
        self.width, self.height = width, height
        self.count = 0

    def close(self):
        if self.h:
            lib().gsplat_context_destroy(self.h)
            self.h = P()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def load(self, scene):
        v = scene_view(scene)
        check(lib().gsplat_context_load_scene_memory(self.h, ctypes.byref(v)))
        self.count = int(scene.count)

    def load_path(self, path: str):
        check(lib().gsplat_context_load_scene_path(self.h, str(path).encode()))

    def set_camera(self, cam: GsplatCamera):
        check(lib().gsplat_context_set_camera(self.h, cam))

    def set_size(self, width: int, height: int):
        check(lib().gsplat_context_set_size(self.h, width, height))
        self.width, self.height = width, height

    def set_sort_interval(self, interval: int):
        """Session order reuse: refresh the depth order every `interval` camera changes."""
        check(lib().gsplat_context_set_sort_interval(self.h, interval))

    def force_sort_refresh(self):
        check(lib().gsplat_context_force_sort_refresh(self.h))

    def set_geometry_path(self, path: int):
        """Device form of the context's scene (GEOMETRY_PATH_DIRECT or GEOMETRY_PATH_PACKED_ATLAS)."""
        check(lib().gsplat_context_set_geometry_path(self.h, path))

    def geometry_path(self) -> int:
        out = ctypes.c_uint32()
        check(lib().gsplat_context_get_geometry_path(self.h, ctypes.byref(out)))
        return out.value

    def sort_stats(self) -> GsplatSurfaceSortStats:
        s = GsplatSurfaceSortStats()
        check(lib().gsplat_context_get_sort_stats(self.h, ctypes.byref(s)))
        return s

    def get_camera(self) -> GsplatCamera:
        c = GsplatCamera()
        check(lib().gsplat_context_get_camera(self.h, ctypes.byref(c)))
        return c

    def render(self):
        check(lib().gsplat_context_render_frame(self.h))

    def wait(self):
        check(lib().gsplat_context_wait(self.h))

    def stats(self) -> GsplatStats:
        s = GsplatStats()
        check(lib().gsplat_context_get_stats(self.h, ctypes.byref(s)))
        return s

    def stage_times(self) -> GsplatStageTimes:
        s = GsplatStageTimes()
        check(lib().gsplat_context_get_stage_times(self.h, ctypes.byref(s)))
        return s

    def set_stage_timing(self, on: bool = True):
        """Per-stage device times for the following frames (an event per stage boundary)."""
        check(lib().gsplat_context_set_stage_timing(self.h, 1 if on else 0))

    def rgba8(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.uint8)
        check(lib().gsplat_context_readback_rgba8(self.h, out.ctypes.data, out.size))
        return out

    def rgba32f(self) -> np.ndarray:
        out = np.zeros((self.height, self.width, 4), np.float32)
        check(lib().gsplat_context_readback_rgba32f(self.h, out.ctypes.data, out.size))
        return out

    def depth_keys(self, count: int) -> np.ndarray:
        out = np.zeros(max(count, 1), np.uint32)
        check(lib().gsplat_context_read_depth_keys(self.h, out.ctypes.data, out.size))
        return out[:count]

    def sorted_indices(self, count: int) -> np.ndarray:
        out = np.zeros(max(count, 1), np.uint32)
        n = ctypes.c_uint32()
        check(lib().gsplat_context_read_sorted_indices(self.h, out.ctypes.data, out.size, ctypes.byref(n)))
        return out[: n.value].copy()

    def projection(self, count: int) -> np.ndarray:
        out = np.zeros((max(count, 1), 12), np.float32)
        check(lib().gsplat_context_read_projection(self.h, out.ctypes.data, out.size))
        return out[:count]

    def autotune(self, reps: int = 3) -> GsplatTuning:
        t = GsplatTuning()
        check(lib().gsplat_context_autotune(self.h, reps, ctypes.byref(t)))
        return t

    def tuning(self) -> GsplatTuning:
        t = GsplatTuning()
        check(lib().gsplat_context_get_tuning(self.h, ctypes.byref(t)))
        return t

    def set_tuning(self, t: GsplatTuning):
        check(lib().gsplat_context_set_tuning(self.h, ctypes.byref(t)))

    def tile_counts(self) -> np.ndarray:
        tx, ty = (self.width + 15) // 16, (self.height + 15) // 16
        out = np.zeros(tx * ty, np.uint32)
        ox, oy = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().gsplat_context_read_tile_counts(self.h, out.ctypes.data, out.size, ctypes.byref(ox),
                                                    ctypes.byref(oy)))
        return out.reshape(oy.value, ox.value)


GEOMETRY_PATH_DIRECT, GEOMETRY_PATH_PACKED_ATLAS, GEOMETRY_PATH_PAGED_ACTIVE_ATLAS = 0, 1, 2


class Scene:
    """A scene resident in device memory (GsplatScene), shared by batches.

    geometry_path: GEOMETRY_PATH_DIRECT (fp32) or GEOMETRY_PATH_PACKED_ATLAS
    (the packed 20-byte records plus the i8 SH sidecar).
    """

    def __init__(self, scene=None, path=None, geometry_path=GEOMETRY_PATH_DIRECT):
        h = P()
        if path is not None:
            check(lib().gsplat_scene_create_from_path_with_geometry_path(str(path).encode(), geometry_path,
                                                                         ctypes.byref(h)))
        else:
            v = scene_view(scene)
            check(lib().gsplat_scene_create_with_geometry_path(ctypes.byref(v), geometry_path, ctypes.byref(h)))
        self.h = h
        info = GsplatSceneInfo()
        check(lib().gsplat_scene_get_info(self.h, ctypes.byref(info)))
        self.count, self.sh_degree, self.device_bytes = info.count, info.sh_degree, info.device_bytes
        self.geometry_path = info.geometry_path

    def close(self):
        if self.h:
            lib().gsplat_scene_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Owns one GsplatBatch: renders lists of (scene, camera, width, height)."""

    def __init__(self):
        h = P()
        check(lib().gsplat_batch_create(ctypes.byref(h)))
        self.h = h
        self.views = []

    def close(self):
        if self.h:
            lib().gsplat_batch_destroy(self.h)
            self.h = P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _descs(views):
        arr = (GsplatViewDesc * max(len(views), 1))()
        for i, (scene, cam, w, h) in enumerate(views):
            arr[i].scene = scene.h if scene is not None else None
            arr[i].camera = cam
            arr[i].width, arr[i].height = w, h
        return arr

    def render(self, views):
        """views: sequence of (Scene, GsplatCamera, width, height)."""
        check(lib().gsplat_batch_render(self.h, self._descs(views), len(views)))
        self.views = [(s, c, w, h) for (s, c, w, h) in views]

    def autotune(self, views, reps: int = 3) -> GsplatTuning:
        t = GsplatTuning()
        check(lib().gsplat_batch_autotune(self.h, self._descs(views), len(views), reps, ctypes.byref(t)))
        self.views = [(s, c, w, h) for (s, c, w, h) in views]
        return t

    def tuning(self) -> GsplatTuning:
        t = GsplatTuning()
        check(lib().gsplat_batch_get_tuning(self.h, ctypes.byref(t)))
        return t

    def set_tuning(self, t: GsplatTuning):
        check(lib().gsplat_batch_set_tuning(self.h, ctypes.byref(t)))

    def wait(self):
        check(lib().gsplat_batch_wait(self.h))

    def _size(self, v):
        _, _, w, h = self.views[v]
        return w, h

    def rgba8(self, v: int) -> np.ndarray:
        w, h = self._size(v)
        out = np.zeros((h, w, 4), np.uint8)
        check(lib().gsplat_batch_readback_rgba8(self.h, v, out.ctypes.data, out.size))
        return out

    def rgba32f(self, v: int) -> np.ndarray:
        w, h = self._size(v)
        out = np.zeros((h, w, 4), np.float32)
        check(lib().gsplat_batch_readback_rgba32f(self.h, v, out.ctypes.data, out.size))
        return out

    def view_stats(self, v: int) -> GsplatStats:
        s = GsplatStats()
        check(lib().gsplat_batch_get_view_stats(self.h, v, ctypes.byref(s)))
        return s

    def stage_times(self) -> GsplatStageTimes:
        s = GsplatStageTimes()
        check(lib().gsplat_batch_get_stage_times(self.h, ctypes.byref(s)))
        return s

    def set_stage_timing(self, on: bool = True):
        """Per-stage device times for the following frames (an event per stage boundary)."""
        check(lib().gsplat_batch_set_stage_timing(self.h, 1 if on else 0))

    def sorted_indices(self, v: int, count: int) -> np.ndarray:
        out = np.zeros(max(count, 1), np.uint32)
        n = ctypes.c_uint32()
        check(lib().gsplat_batch_read_sorted_indices(self.h, v, out.ctypes.data, out.size, ctypes.byref(n)))
        return out[: n.value].copy()

    def tile_counts(self, v: int) -> np.ndarray:
        w, h = self._size(v)
        out = np.zeros(((w + 15) // 16) * ((h + 15) // 16), np.uint32)
        ox, oy = ctypes.c_uint32(), ctypes.c_uint32()
        check(lib().gsplat_batch_read_tile_counts(self.h, v, out.ctypes.data, out.size, ctypes.byref(ox),
                                                  ctypes.byref(oy)))
        return out.reshape(oy.value, ox.value)


class PinnedImage:
    """RGBA8 target in pinned host memory (gsplat_host_alloc)."""

    def __init__(self, width: int, height: int):
        self.nbytes = width * height * 4
        self.ptr = lib().gsplat_host_alloc(self.nbytes)
        if not self.ptr:
            raise GsplatError(INTERNAL, last_error())
        buf = (ctypes.c_uint8 * self.nbytes).from_address(self.ptr)
        self.array = np.ctypeslib.as_array(buf).reshape(height, width, 4)

    def close(self):
        """Frees the block; raises GsplatError (and keeps it) while a request still targets it."""
        if self.ptr:
            check(lib().gsplat_host_free(self.ptr))
            self.array = None
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Owns one GsplatServer (continuous batching over resident scenes)."""

    def __init__(self, memory_budget_bytes: int = 0, max_batch_views: int = 64, max_queue: int = 4096,
                 pipelines: int = 3):
        cfg = lib().gsplat_server_config_default()
        cfg.memory_budget_bytes, cfg.max_batch_views, cfg.max_queue = memory_budget_bytes, max_batch_views, max_queue
        cfg.pipelines = pipelines
        h = P()
        check(lib().gsplat_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        # ticket -> target buffer: a target stays referenced (not freed by the
        # garbage collector) until wait() returns for its ticket
        self._targets = {}

    def close(self):
        if self.h:
            lib().gsplat_server_destroy(self.h)
            self.h = P()
            self._targets.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, scene: "Scene") -> int:
        sid = ctypes.c_uint32()
        check(lib().gsplat_server_add_scene(self.h, scene.h, ctypes.byref(sid)))
        return sid.value

    def remove_scene(self, scene_id: int):
        check(lib().gsplat_server_remove_scene(self.h, scene_id))

    def submit(self, scene_id: int, cam: GsplatCamera, width: int, height: int, out=None) -> int:
        """`out`: None, a PinnedImage, or a C-contiguous uint8 numpy array of width*height*4 bytes."""
        req = GsplatRenderRequest()
        req.scene_id, req.camera, req.width, req.height = scene_id, cam, width, height
        if isinstance(out, PinnedImage):
            req.out_rgba8, req.out_len = out.ptr, out.nbytes
        elif out is not None:
            req.out_rgba8, req.out_len = out.ctypes.data, out.nbytes
        t = ctypes.c_uint64()
        check(lib().gsplat_server_submit(self.h, ctypes.byref(req), ctypes.byref(t)))
        if out is not None:
            self._targets[t.value] = out
        return t.value

    def wait(self, ticket: int, raise_on_error: bool = True) -> GsplatRequestResult:
        r = GsplatRequestResult()
        rc = lib().gsplat_server_wait(self.h, ticket, ctypes.byref(r))
        self._targets.pop(ticket, None)  # the ticket is released (or was never known)
        if raise_on_error:
            check(rc)
        return r

    def stats(self) -> GsplatServerStats:
        s = GsplatServerStats()
        check(lib().gsplat_server_get_stats(self.h, ctypes.byref(s)))
        return s


def sort_values_by_keys(keys, values) -> np.ndarray:
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.uint32))
    v = np.ascontiguousarray(np.asarray(values, dtype=np.uint32)).copy()
    if k.shape != v.shape:
        raise ValueError("key/value length mismatch")
    check(lib().gsplat_sort_values_by_keys(k.ctypes.data if k.size else None, v.ctypes.data if v.size else None,
                                           k.size))
    return v


PACKED_CONSTS = 272


def pack_scene(scene):
    """The packed records of a scene (gsplat_pack_scene): (words, consts).

    words has shape (words_per_splat, count): word w of splat i at words[w, i].
    """
    v = scene_view(scene)
    per = ctypes.c_uint32()
    check(lib().gsplat_pack_scene(ctypes.byref(v), None, 0, None, 0, ctypes.byref(per)))
    words = np.zeros((per.value, v.count), np.uint32)
    consts = np.zeros(PACKED_CONSTS, np.float32)
    if v.count:
        check(lib().gsplat_pack_scene(ctypes.byref(v), words.ctypes.data, words.size, consts.ctypes.data, consts.size,
                                      ctypes.byref(per)))
    return words, consts


class PlyScene:
    """A parsed PLY scene owned by the library (gsplat_ply_*)."""

    def __init__(self, handle):
        self.h = handle
        v = GsplatSceneView()
        check(lib().gsplat_ply_scene_view(self.h, ctypes.byref(v)))
        n = v.count
        self.count = n
        self.sh_degree = v.sh_degree

        def arr(ptr, k):
            if not ptr or n == 0:
                return np.zeros((n, k), np.float32)
            return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(ctypes.c_float)), shape=(n * k,)).reshape(
                n, k).copy()

        self.positions = arr(v.positions, 3)
        self.opacity = arr(v.opacity, 1).reshape(n)
        self.scale_xyz = arr(v.scale_xyz, 3)
        self.rotation_xyzw = arr(v.rotation_xyzw, 4)
        self.color_dc = arr(v.color_dc, 3)
        k = 3 * ((v.sh_degree + 1) ** 2 - 1)
        self.sh_rest = arr(v.sh_rest, k) if v.sh_rest else None
        lib().gsplat_ply_free(self.h)
        self.h = None


def ply_load(path) -> PlyScene:
    h = P()
    check(lib().gsplat_ply_load(str(path).encode(), ctypes.byref(h)))
    return PlyScene(h)


def ply_parse(data: bytes, limits: GsplatPlyLimits | None = None, text: bool = False) -> PlyScene:
    h = P()
    buf = ctypes.create_string_buffer(data, len(data)) if data else None
    fn = lib().gsplat_ply_parse_text if text else lib().gsplat_ply_parse_bytes
    check(fn(ctypes.cast(buf, P) if buf is not None else None, len(data),
             ctypes.byref(limits) if limits is not None else None, ctypes.byref(h)))
    return PlyScene(h)


def ply_limits_default() -> GsplatPlyLimits:
    return lib().gsplat_ply_limits_default()
