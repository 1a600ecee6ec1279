"""The session order-reuse policy of a context (gs::OrderReusePolicy), on the
CPU: the refresh plan of SurfaceFrameState (render-wgpu/src/surface_session.rs:
556-596: refresh when forced, without an order, or after `interval` camera
changes), camera revisions counted on changed cameras only (:734-744), and the
bounds of a presented order (:17-32, 1092-1105, 1245-1262): at most 2
revisions of lag and a camera within 2% of the scene diagonal and 0.01 rad."""

import ctypes
import math
import pathlib
import subprocess

import pytest

import gsplat_abi

ROOT = pathlib.Path(__file__).resolve().parents[1]
REUSE, REFRESH, FALLBACK = 0, 1, 3


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("orp") / "liborp.so"
    subprocess.run(["g++", "-O1", "-std=c++17", "-shared", "-fPIC", "-I", str(ROOT / "gsplat-rs_amd/csrc"),
                    "-I", str(ROOT / "include"), str(ROOT / "tests/native/order_reuse_check.cpp"), "-o", str(out)],
                   check=True)
    L = ctypes.CDLL(str(out))
    L.orp_new.restype = ctypes.c_void_p
    for name, args in (("orp_free", [ctypes.c_void_p]), ("orp_set_interval", [ctypes.c_void_p, ctypes.c_uint]),
                       ("orp_set_diagonal", [ctypes.c_void_p, ctypes.c_float]), ("orp_force", [ctypes.c_void_p]),
                       ("orp_camera", [ctypes.c_void_p, ctypes.POINTER(gsplat_abi.GsplatCamera)]),
                       ("orp_frame", [ctypes.c_void_p, ctypes.c_int]),
                       ("orp_stats", [ctypes.c_void_p, ctypes.POINTER(gsplat_abi.GsplatSurfaceSortStats)])):
        getattr(L, name).argtypes = args
    return L


class Policy:
    def __init__(self, L, interval=2, diagonal=5.0):
        self.L, self.h = L, L.orp_new()
        L.orp_set_diagonal(self.h, diagonal)
        L.orp_set_interval(self.h, interval)

    def camera(self, yaw=0.0, x=0.0):
        c = gsplat_abi.GsplatCamera()
        c.position[:] = [x, 0.0, -4.0]
        c.rotation_xyzw[:] = [0.0, math.sin(yaw / 2), 0.0, math.cos(yaw / 2)]
        c.vertical_fov_radians, c.near_plane, c.far_plane = 1.0, 0.01, 1000.0
        self.L.orp_camera(self.h, ctypes.byref(c))

    def frame(self, has_order=True):
        return self.L.orp_frame(self.h, int(has_order))

    def stats(self):
        s = gsplat_abi.GsplatSurfaceSortStats()
        self.L.orp_stats(self.h, ctypes.byref(s))
        return s


def test_interval_two_refreshes_every_second_change(lib):
    p = Policy(lib, interval=2)
    p.camera(0.0)
    assert p.frame(has_order=False) == REFRESH  # no order yet
    plan = []
    for k in range(1, 9):
        p.camera(0.002 * k)  # small steps inside the pose envelope
        plan.append(p.frame())
    assert plan == [REUSE, REFRESH] * 4
    s = p.stats()
    assert s.camera_revision == 9 and s.applied_order_revision == 9 and s.presented_order_revision_lag == 0
    assert s.flags & 1 and not s.flags & (1 << 7)
    p.camera(0.0185)
    assert p.frame() == REUSE
    s = p.stats()
    assert s.presented_order_revision_lag == 1 and s.applied_order_revision == 9 and not s.flags & 1


def test_unchanged_camera_is_no_new_revision(lib):
    p = Policy(lib, interval=1)
    p.camera(0.0)
    assert p.frame(False) == REFRESH
    for _ in range(3):
        p.camera(0.0)  # the same camera again
        assert p.frame() == REUSE
    assert p.stats().camera_revision == 1
    p.camera(0.001)
    assert p.frame() == REFRESH  # interval 1: every change refreshes


def test_revision_lag_bound(lib):
    p = Policy(lib, interval=10)
    p.camera(0.0)
    assert p.frame(False) == REFRESH
    got = []
    for k in range(1, 8):
        p.camera(0.001 * k)
        got.append(p.frame())
    # reused at lag 1 and 2, refreshed (bounds) at lag 3, and so on
    assert got == [REUSE, REUSE, FALLBACK, REUSE, REUSE, FALLBACK, REUSE]
    assert p.stats().presented_order_revision_lag == 1


def test_pose_envelope(lib):
    p = Policy(lib, interval=10, diagonal=5.0)  # translation bound 0.1
    p.camera(0.0, x=0.0)
    assert p.frame(False) == REFRESH
    p.camera(0.0, x=0.099)
    assert p.frame() == REUSE
    p.camera(0.0, x=0.101)
    assert p.frame() == FALLBACK  # 0.101 from the order's camera
    p.camera(0.009, x=0.101)
    assert p.frame() == REUSE  # 0.009 rad
    p.camera(0.0115, x=0.101)
    assert p.frame() == FALLBACK  # 0.0115 rad
    assert p.stats().flags & (1 << 7)


def test_forced_refresh_and_interval_change(lib):
    p = Policy(lib, interval=3)
    p.camera(0.0)
    assert p.frame(False) == REFRESH
    p.camera(0.001)
    assert p.frame() == REUSE
    lib.orp_force(p.h)
    assert p.frame() == REFRESH
    lib.orp_set_interval(p.h, 4)  # a new interval refreshes
    assert p.frame() == REFRESH
    lib.orp_set_interval(p.h, 4)  # the same interval does not
    assert p.frame() == REUSE
    assert p.frame(has_order=False) == REFRESH
