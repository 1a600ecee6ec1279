"""Shared helpers of the test suite (scene conversion, image metrics)."""

import math

import numpy as np

DEFAULT_FOV = np.float32(60.0) * (np.float32(math.pi) / np.float32(180.0))


def oracle_scene(oracle, s):
    return oracle.Scene(s.positions, s.opacity, s.scale_xyz, s.rotation_xyzw, s.color_dc, s.sh_degree, s.sh_rest)


def cam_dict(c):
    return dict(position=list(c.position), rotation_xyzw=list(c.rotation_xyzw), fov=c.vertical_fov_radians,
                near=c.near_plane, far=c.far_plane)


from image_metrics import rgba_parity_metrics  # noqa: E402,F401  (the reference's image-parity metric)


def minimal_ascii_rows():
    """Rows of tests/golden/minimal_ascii.ply (fixture copied from the reference's tests/datasets)."""
    return np.array([[0.0, 0.0, 0.5, 0.9, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.9, 0.2, 0.1],
                     [0.1, 0.1, 1.2, 0.8, 1.1, 1.0, 0.9, 1.0, 0.0, 0.0, 0.0, 0.7, 0.3, 0.2],
                     [0.0, 0.0, -0.1, 0.5, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.0, 0.1, 0.1, 0.1]], np.float32)


def minimal_ascii_scene(oracle):
    """minimal_ascii.ply after the loader's RDF->RUF conversion (y, qx, qz negated; wxyz->xyzw)."""
    r = minimal_ascii_rows()
    pos = r[:, 0:3].copy()
    pos[:, 1] = -pos[:, 1]
    q = r[:, [8, 9, 10, 7]].copy()
    q[:, 0] = -q[:, 0]
    q[:, 2] = -q[:, 2]
    return oracle.Scene(pos, r[:, 3], r[:, 4:7], q, r[:, 11:14])


def default_cam():
    return dict(position=[0, 0, 0], rotation_xyzw=[0, 0, 0, 1], fov=DEFAULT_FOV, near=0.01, far=1000.0)


def lcg_keys(n, seed):
    """lcg_next of sort/src/lib.rs:597-600."""
    out = np.zeros(n, np.uint32)
    s = seed & 0xFFFFFFFF
    for i in range(n):
        s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
        out[i] = s
    return out


def lcg_keys_fast(n, seed):
    """Same sequence as lcg_keys, vectorised by blocks of the affine recurrence."""
    if n < 4096:
        return lcg_keys(n, seed)
    a, c, m = 1664525, 1013904223, 1 << 32
    out = np.zeros(n, np.uint64)
    s = seed
    # powers a^k and sums for k = 1..B
    B = 1024
    ak = np.zeros(B, np.uint64)
    ck = np.zeros(B, np.uint64)
    x, y = 1, 0
    for k in range(B):
        x = (x * a) % m
        y = (y * a + c) % m
        ak[k], ck[k] = x, y
    for start in range(0, n, B):
        cnt = min(B, n - start)
        vals = (ak[:cnt] * np.uint64(s) + ck[:cnt]) % np.uint64(m)
        out[start:start + cnt] = vals
        s = int(vals[-1])
    return out.astype(np.uint32)
