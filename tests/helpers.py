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


def decode_packed(words, consts, sh_degree):
    """The packed scene codec's CPU decoders (render-wgpu/src/packed_atlas.rs: SceneBounds::decode_u16
    :86-97, LogScaleRange::decode_u10 :184-189, unpack_quat_smallest_three :311-337, unpack_color_rgb10
    :350-357, dequantize_sh_rest :409-419) over coefficient-major words (gsplat_pack_scene)."""
    f = np.float32
    lo, span = consts[0:3], consts[3:6]
    lmin, lspan = consts[6], consts[7]
    pos = np.stack([words[0] & 0xFFFF, words[0] >> 16, words[1] & 0xFFFF], 1).astype(f)
    pos = lo + (pos / f(65535.0)) * span
    op = (words[1] >> 16) & 0xFF
    log = np.stack([(words[2] >> (10 * a)) & 0x3FF for a in range(3)], 1).astype(f)
    log = lmin + (log / f(1023.0)) * lspan
    q = words[3]
    comps = np.stack([((q >> (2 + 10 * l)) & 0x3FF).astype(f) / f(1023.0) * f(2.0) - f(1.0) for l in range(3)], 1)
    comps = comps * np.float32(0.70710678118654752440)
    dropped = np.sqrt(np.maximum(f(1.0) - (comps ** 2).sum(1), f(0.0)))
    big = q & 3
    quat = np.zeros((len(q), 4), f)
    for k in range(len(q)):
        others = [c for c in range(4) if c != big[k]]
        quat[k, others] = comps[k]
        quat[k, big[k]] = dropped[k]
    quat /= np.sqrt((quat ** 2).sum(1, keepdims=True))
    rgb = np.stack([(words[4] >> (10 * a)) & 0x3FF for a in range(3)], 1).astype(f) / f(1023.0)
    rest = np.zeros((len(q), 45), f)
    if sh_degree:
        side = words[5:17].T.copy().view(np.int8).reshape(len(q), 48)[:, :45].astype(f)
        rest = side / f(127.0) * np.repeat(consts[8:11], 15)[None, :]
    return dict(pos=pos, op=op, log=log, quat=quat, rgb=rgb, rest=rest, cut=consts[16:16 + 256])


def packed_decoded_scene(oracle, words, consts, sh_degree):
    """An oracle scene holding what a packed scene decodes to: decoded positions, log-scales and
    unit quaternions, the logit of the decoded alpha, the DC coefficient whose degree-0 term is the
    decoded RGB10 colour less 0.5, and the dequantized SH coefficients."""
    d = decode_packed(words, consts, sh_degree)
    alpha = np.clip(d["op"].astype(np.float64) / 255.0, 1e-4, 1.0 - 1e-4)
    opacity = np.log(alpha / (1.0 - alpha)).astype(np.float32)
    dc = ((d["rgb"].astype(np.float64) - 0.5) / 0.28209479177387814).astype(np.float32)
    rest = None
    if sh_degree:
        pc = (sh_degree + 1) ** 2 - 1
        rest = np.ascontiguousarray(d["rest"].reshape(-1, 3, 15)[:, :, :pc].reshape(-1, 3 * pc))
    return oracle.Scene(d["pos"], opacity, d["log"], d["quat"], dc, sh_degree, rest)
