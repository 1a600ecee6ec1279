"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker. The product never imports it.
"""

from __future__ import annotations

import ctypes
import os
import pathlib

import numpy as np

_HERE = pathlib.Path(__file__).resolve().parent
# GSPLAT_ORACLE_LIB names another build of oracle.cpp (the mutation self-check
# of the pinning tests loads builds with a planted mistake this way).
_LIB_PATH = pathlib.Path(os.environ.get("GSPLAT_ORACLE_LIB") or (_HERE / "liboracle.so"))


class OracleCamera(ctypes.Structure):
    _fields_ = [
        ("position", ctypes.c_float * 3),
        ("rotation_xyzw", ctypes.c_float * 4),
        ("vertical_fov_radians", ctypes.c_float),
        ("near_plane", ctypes.c_float),
        ("far_plane", ctypes.c_float),
    ]


class OracleScene(ctypes.Structure):
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


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not _LIB_PATH.exists():
            raise RuntimeError(f"oracle not built: {_LIB_PATH} (run `make -C oracle`)")
        _lib = ctypes.CDLL(str(_LIB_PATH))
        _lib.oracle_render.restype = ctypes.c_uint32
        _lib.oracle_render_window.restype = ctypes.c_uint32
        _lib.oracle_render_prepared.restype = ctypes.c_uint32
        _lib.oracle_preprocess.restype = ctypes.c_uint32
        _lib.oracle_coverage_cut.restype = ctypes.c_float
        _lib.oracle_coverage_cut.argtypes = [ctypes.c_float]
        _lib.oracle_camera_valid.restype = ctypes.c_int
        _lib.oracle_sort_values_by_keys.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]
        _lib.oracle_blend_evals.restype = ctypes.c_uint64
    return _lib


def _ptr(a: np.ndarray | None):
    if a is None:
        return None
    return ctypes.c_void_p(a.ctypes.data)


def make_camera(cam) -> OracleCamera:
    """`cam` is a dict with position, rotation_xyzw, fov, near, far."""
    c = OracleCamera()
    c.position[:] = [float(v) for v in cam["position"]]
    c.rotation_xyzw[:] = [float(v) for v in cam["rotation_xyzw"]]
    c.vertical_fov_radians = float(cam["fov"])
    c.near_plane = float(cam["near"])
    c.far_plane = float(cam["far"])
    return c


class Scene:
    """Holds contiguous float32 arrays in the reference SceneBuffers layout."""

    def __init__(self, positions, opacity, scale_xyz, rotation_xyzw, color_dc, sh_degree=0, sh_rest=None):
        f = lambda a: np.ascontiguousarray(np.asarray(a, dtype=np.float32))  # noqa: E731
        self.positions = f(positions).reshape(-1, 3)
        self.count = self.positions.shape[0]
        self.opacity = f(opacity).reshape(self.count)
        self.scale_xyz = f(scale_xyz).reshape(self.count, 3)
        self.rotation_xyzw = f(rotation_xyzw).reshape(self.count, 4)
        self.color_dc = f(color_dc).reshape(self.count, 3)
        self.sh_degree = int(sh_degree)
        self.sh_rest = None if sh_rest is None else f(sh_rest).reshape(self.count, -1)

    def c(self) -> OracleScene:
        s = OracleScene()
        s.count = self.count
        s.sh_degree = self.sh_degree
        s.positions = self.positions.ctypes.data
        s.opacity = self.opacity.ctypes.data
        s.scale_xyz = self.scale_xyz.ctypes.data
        s.rotation_xyzw = self.rotation_xyzw.ctypes.data
        s.color_dc = self.color_dc.ctypes.data
        s.sh_rest = None if self.sh_rest is None else self.sh_rest.ctypes.data
        return s


def precompute(scene: Scene):
    alpha = np.zeros(scene.count, np.float32)
    cov = np.zeros((scene.count, 6), np.float32)
    sc = scene.c()
    lib().oracle_precompute(ctypes.byref(sc), _ptr(alpha), _ptr(cov))
    return alpha, cov


def view_rows(rotation_xyzw) -> np.ndarray:
    q = (ctypes.c_float * 4)(*[float(v) for v in rotation_xyzw])
    out = (ctypes.c_float * 9)()
    lib().oracle_view_rows(q, out)
    return np.array(out, dtype=np.float32).reshape(3, 3)


def camera_valid(cam) -> bool:
    c = make_camera(cam)
    return bool(lib().oracle_camera_valid(ctypes.byref(c)))


def preprocess(scene: Scene, cam):
    keys = np.zeros(max(scene.count, 1), np.uint32)
    idx = np.zeros(max(scene.count, 1), np.uint32)
    sc, c = scene.c(), make_camera(cam)
    n = lib().oracle_preprocess(ctypes.byref(sc), ctypes.byref(c), _ptr(keys), _ptr(idx))
    return keys[:n].copy(), idx[:n].copy()


def sort_values_by_keys(keys, values) -> np.ndarray:
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.uint32))
    v = np.ascontiguousarray(np.asarray(values, dtype=np.uint32)).copy()
    assert k.shape == v.shape
    if k.size:
        lib().oracle_sort_values_by_keys(_ptr(k), _ptr(v), ctypes.c_uint64(k.size))
    return v


def project(scene: Scene, cam, width: int, height: int) -> np.ndarray:
    out = np.zeros((max(scene.count, 1), 12), np.float32)
    sc, c = scene.c(), make_camera(cam)
    lib().oracle_project(ctypes.byref(sc), ctypes.byref(c), ctypes.c_uint32(width), ctypes.c_uint32(height), _ptr(out))
    return out[: scene.count]


def coverage_cut(alpha: float) -> float:
    return float(lib().oracle_coverage_cut(ctypes.c_float(alpha)))


def tile_counts(proj12: np.ndarray, width: int, height: int, tile: int = 16) -> np.ndarray:
    tx, ty = (width + tile - 1) // tile, (height + tile - 1) // tile
    out = np.zeros(tx * ty, np.uint32)
    p = np.ascontiguousarray(proj12, dtype=np.float32)
    lib().oracle_tile_counts(ctypes.c_uint32(p.shape[0]), _ptr(p), ctypes.c_uint32(width), ctypes.c_uint32(height),
                             ctypes.c_uint32(tile), _ptr(out))
    return out.reshape(ty, tx)


def render(scene: Scene, cam, width: int, height: int, threads: int = 0):
    """Returns (rgba float32 [H, W, 4], visible_count)."""
    out = np.zeros((height, width, 4), np.float32)
    sc, c = scene.c(), make_camera(cam)
    vis = lib().oracle_render(ctypes.byref(sc), ctypes.byref(c), ctypes.c_uint32(width), ctypes.c_uint32(height),
                              ctypes.c_int(threads), _ptr(out))
    if vis == 0xFFFFFFFF:
        raise ValueError("invalid camera or size")
    return out, int(vis)


def render_prepared(scene: Scene, prepared, cam, width: int, height: int, threads: int = 0):
    """render() from precompute(scene)'s load-time values (computed once per scene, as the reference does)."""
    alpha, cov = prepared
    out = np.zeros((height, width, 4), np.float32)
    sc, c = scene.c(), make_camera(cam)
    vis = lib().oracle_render_prepared(ctypes.byref(sc), _ptr(alpha), _ptr(cov), ctypes.byref(c),
                                       ctypes.c_uint32(width), ctypes.c_uint32(height), ctypes.c_int(threads),
                                       _ptr(out))
    if vis == 0xFFFFFFFF:
        raise ValueError("invalid camera or size")
    return out, int(vis)


def render_window(scene: Scene, cam, width: int, height: int, x0: int, y0: int, x1: int, y1: int,
                  threads: int = 0):
    """Pixels [x0, x1) x [y0, y1) of render(): (rgba float32 [y1-y0, x1-x0, 4], visible_count)."""
    out = np.zeros((y1 - y0, x1 - x0, 4), np.float32)
    sc, c = scene.c(), make_camera(cam)
    vis = lib().oracle_render_window(ctypes.byref(sc), ctypes.byref(c), ctypes.c_uint32(width),
                                     ctypes.c_uint32(height), ctypes.c_uint32(x0), ctypes.c_uint32(y0),
                                     ctypes.c_uint32(x1), ctypes.c_uint32(y1), ctypes.c_int(threads), _ptr(out))
    if vis == 0xFFFFFFFF:
        raise ValueError("invalid camera, size or window")
    return out, int(vis)


def blend_evals(scene: Scene, cam, width: int, height: int) -> int:
    sc, c = scene.c(), make_camera(cam)
    return int(lib().oracle_blend_evals(ctypes.byref(sc), ctypes.byref(c), ctypes.c_uint32(width),
                                        ctypes.c_uint32(height)))


def last_timings() -> list[float]:
    out = (ctypes.c_double * 4)()
    lib().oracle_last_timings(out)
    return list(out)


def to_rgba8(rgba: np.ndarray) -> np.ndarray:
    """One final UNORM8 rounding of the fp32 target: round(clamp(v,0,1)*255)."""
    return np.floor(np.clip(rgba, 0.0, 1.0) * np.float32(255.0) + np.float32(0.5)).astype(np.uint8)


def default_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
