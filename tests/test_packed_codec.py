"""The packed scene codec's host encoder (gsplat_pack_scene), on the CPU.

Mirrors the reference's own codec tests (render-wgpu/src/packed_atlas.rs:
629-915: record sizes, metadata scans, position / quaternion / SH / colour /
opacity / log-scale round trips, word placement) through a decoder written
here from the reference's CPU decoders (SceneBounds::decode_u16 :86-97,
LogScaleRange::decode_u10 :184-189, unpack_quat_smallest_three :311-337,
unpack_color_rgb10 :350-357, dequantize_sh_rest :409-419), and pins the exact
words of hand-derived cases (identity rotation, the opacity 0 code, range end
points, the half-away-from-zero rounding of f32::round, the smallest-three
choice when the largest component is negative).
"""

import ctypes

import numpy as np
import pytest

import synth
from helpers import decode_packed as decode

C0 = np.float32(0.2820948)
INV_SQRT2 = np.float32(0.70710678118654752440)  # smallest-three component range


def sample_scene(count, sh_degree):
    """sample_scene of packed_atlas.rs:633-657."""
    i = np.arange(count, dtype=np.float32)
    pos = np.stack([i * np.float32(0.01), np.sin(i * np.float32(0.02)), np.float32(1.0) + i * np.float32(0.001)], 1)
    rest = None
    if sh_degree:
        stride = ((sh_degree + 1) ** 2 - 1) * 3
        rest = ((np.arange(count * stride) % 17).astype(np.float32) - np.float32(8.0)) * np.float32(0.02)
        rest = rest.reshape(count, stride)
    return synth.Scene(pos.astype(np.float32), np.zeros(count, np.float32),
                       np.tile(np.float32([-2.0, -1.5, -1.0]), (count, 1)),
                       np.tile(np.float32([0, 0, 0, 1]), (count, 1)),
                       np.tile(np.float32([0.2, -0.1, 0.4]), (count, 1)), sh_degree, rest)


def test_record_sizes_and_degree3_reduction(abi):
    words, consts = abi.pack_scene(sample_scene(1024, 3))
    assert words.shape == (17, 1024) and consts.shape == (abi.PACKED_CONSTS,)
    assert 4 * 5 == 20 and 4 * 12 == 48  # hot record, SH sidecar
    direct_bytes = 4 * (14 + 45)  # the fp32 SoA of an SH3 splat: 59 floats
    assert direct_bytes / (4 * words.shape[0]) >= 3.0
    for degree, per in ((0, 5), (1, 17), (2, 17), (3, 17)):
        assert abi.pack_scene(sample_scene(4, degree))[0].shape[0] == per


def test_exact_words_of_hand_derived_cases(abi):
    s = sample_scene(1, 0)
    words, consts = abi.pack_scene(s)
    # A single splat: degenerate bounds become [0, 1] per axis (from_positions, :56-77).
    assert list(consts[0:6]) == [0, 0, 0, 1, 1, 1]
    x, y, z = s.positions[0]
    assert (x, y, z) == (0.0, 0.0, 1.0)
    assert words[0, 0] == 0 and words[1, 0] & 0xFFFF == 65535
    # opacity logit 0 -> alpha 0.5 -> 127.5 rounds away from zero to 128.
    assert words[1, 0] >> 16 == 128
    # log-scales -2, -1.5, -1 over [min, max] = [-2, -1]: 0, 511.5 -> 512, 1023.
    assert (consts[6], consts[7]) == (-2.0, 1.0)
    assert words[2, 0] == 0 | (512 << 10) | (1023 << 20)
    # identity (0, 0, 0, 1): dropped index 3, the others map 0 -> 511.5 -> 512.
    assert words[3, 0] == 3 | (512 << 2) | (512 << 12) | (512 << 22)
    # colour: round(clamp(C0 * dc + 0.5) * 1023) per channel.
    want = 0
    for a, dc in enumerate((0.2, -0.1, 0.4)):
        v = np.float32(C0 * np.float32(dc) + np.float32(0.5))
        want |= int(np.floor(v * np.float32(1023.0) + np.float32(0.5))) << (10 * a)
    assert words[4, 0] == want
    # SH scales of a degree-0 scene stay at their floor.
    assert np.all(consts[8:11] == np.float32(1e-6)) and np.all(consts[11:16] == 0)


def test_quaternion_smallest_three_choice(abi):
    s = sample_scene(3, 0)
    # largest |component| negative (the case packed_atlas.rs:643 notes): negated whole.
    s.rotation_xyzw[:] = [[0.9, -0.2, 0.1, -0.3], [-0.7, 0.1, 0.2, 0.1], [0.1, 0.8, -0.2, -0.4]]
    words, consts = abi.pack_scene(s)
    assert list(words[3] & 3) == [0, 0, 1]
    q = np.asarray(s.rotation_xyzw, np.float32)
    q = q / np.sqrt((q ** 2).sum(1, keepdims=True))
    q1 = -q[1]  # largest is -0.7: flipped positive
    lanes = [(int(words[3, 1]) >> (2 + 10 * l)) & 0x3FF for l in range(3)]
    for lane, comp in zip(lanes, q1[1:]):
        assert lane == int(np.round(np.clip(comp / INV_SQRT2 * 0.5 + 0.5, 0, 1) * 1023))


def test_metadata_scans_match_packing(abi):
    s = sample_scene(8, 3)
    s.scale_xyz[0] = [-7.0, -2.0, 4.0]
    s.scale_xyz[7] = [5.0, -3.0, 2.0]
    words, consts = abi.pack_scene(s)
    assert consts[6] == -7.0 and consts[7] == np.float32(5.0) - np.float32(-7.0)
    rest = s.sh_rest.reshape(8, 3, 15)
    assert np.array_equal(consts[8:11], np.abs(rest).max(axis=(0, 2)))
    pos = s.positions
    assert np.array_equal(consts[0:3], pos.min(0))
    assert np.array_equal(consts[3:6], np.maximum(pos.max(0) - pos.min(0), np.float32(1e-6)))


def test_position_roundtrip_within_bounds_error(abi):
    s = sample_scene(8, 0)
    d = decode(*abi.pack_scene(s), 0)
    assert np.abs(d["pos"] - s.positions).max() < 1e-3


@pytest.mark.parametrize("q", [[0.0, 0.0, 0.0, 1.0], [0.1, -0.2, 0.3, 0.9], [0.5, 0.5, 0.5, 0.5], [-0.3, 0.1, 0.2, 0.9],
                               [0.9, -0.2, 0.1, -0.3], [-0.7, 0.1, 0.2, 0.1], [0.1, 0.8, -0.2, -0.4]])
def test_quaternion_roundtrip_preserves_orientation(abi, q):
    s = sample_scene(1, 0)
    s.rotation_xyzw[0] = q
    d = decode(*abi.pack_scene(s), 0)
    a = np.float32(q) / np.sqrt(np.sum(np.float32(q) ** 2))
    b = d["quat"][0]
    assert min(np.abs(a - b).sum(), np.abs(a + b).sum()) < 0.05


def test_sh_sidecar_roundtrip_keeps_sign_and_scale(abi):
    s = sample_scene(4, 3)
    d = decode(*abi.pack_scene(s), 3)
    assert np.abs(d["rest"] - s.sh_rest).max() < 0.03
    assert np.all(np.sign(d["rest"][np.abs(s.sh_rest) > 0.02]) == np.sign(s.sh_rest[np.abs(s.sh_rest) > 0.02]))


@pytest.mark.parametrize("degree", [0, 1, 2, 3])
def test_sh_sidecar_degree_transitions(abi, degree):
    s = sample_scene(4, degree)
    words, consts = abi.pack_scene(s)
    d = decode(words, consts, degree)
    if degree == 0:
        assert words.shape[0] == 5
        return
    pc = (degree + 1) ** 2 - 1
    src = s.sh_rest.reshape(4, 3, pc)
    got = d["rest"].reshape(4, 3, 15)
    assert np.abs(got[:, :, :pc] - src).max() < 0.03
    assert np.all(got[:, :, pc:] == 0)
    # the three pad bytes stay zero
    assert np.all(words[16] >> 8 == 0)


def test_colour_opacity_and_log_scale_codecs(abi, oracle):
    s = sample_scene(5, 0)
    s.opacity[:] = [-4.0, -1.0, 0.0, 1.0, 3.0]
    s.scale_xyz[:, 0] = [-4.0, -1.0, 0.0, 1.0, 3.0]
    s.scale_xyz[:, 1] = -4.0
    s.scale_xyz[:, 2] = 4.0
    s.color_dc[:] = [[-5.0, 0.0, 5.0]] * 5  # clamps to 0, 0.5, 1
    words, consts = abi.pack_scene(s)
    d = decode(words, consts, 0)
    alpha = 1.0 / (1.0 + np.exp(-s.opacity.astype(np.float64)))
    assert np.array_equal(d["op"], np.round(alpha * 255).astype(np.uint32))
    assert np.abs(d["log"] - s.scale_xyz).max() < 0.01
    assert np.abs(d["rgb"] - [0.0, 0.5, 1.0]).max() < 1.0 / 1023.0 + 1e-6
    # coverage cut of each opacity code: the oracle's cut of the decoded alpha, bit for bit.
    codes = np.arange(256, dtype=np.float32)
    alphas = np.clip(codes / np.float32(255.0), np.float32(1e-4), np.float32(1.0) - np.float32(1e-4))
    want = np.float32([oracle.coverage_cut(float(a)) for a in alphas])
    assert np.array_equal(d["cut"].view(np.uint32), want.view(np.uint32))


def test_words_two_to_four_hold_scales_rotation_colour(abi):
    s = synth.tiny_scene(300, 5, sh_degree=2)
    words, consts = abi.pack_scene(s)
    d = decode(words, consts, 2)
    assert np.abs(d["log"] - s.scale_xyz).max() <= consts[7] / 1023 * 0.5 + 1e-5
    span = np.maximum(s.positions.max(0) - s.positions.min(0), 1e-6)
    assert np.all(np.abs(d["pos"] - s.positions) <= span / 65535 * 0.5 + 1e-6)
    q = s.rotation_xyzw / np.linalg.norm(s.rotation_xyzw, axis=1, keepdims=True)
    err = np.minimum(np.abs(d["quat"] - q).sum(1), np.abs(d["quat"] + q).sum(1))
    assert err.max() < 0.01


def test_degree4_and_bad_inputs_are_refused(abi):
    s = synth.tiny_scene(8, 1, sh_degree=4)
    with pytest.raises(abi.GsplatError) as e:
        abi.pack_scene(s)
    assert e.value.code == abi.UNSUPPORTED
    v = abi.scene_view(sample_scene(4, 0))
    per = ctypes.c_uint32()
    words = np.zeros(5 * 4 - 1, np.uint32)
    consts = np.zeros(272, np.float32)
    rc = abi.lib().gsplat_pack_scene(ctypes.byref(v), words.ctypes.data, words.size, consts.ctypes.data, 272,
                                     ctypes.byref(per))
    assert rc == abi.INVALID_ARGUMENT
    assert abi.lib().gsplat_pack_scene(None, None, 0, None, 0, ctypes.byref(per)) == abi.INVALID_ARGUMENT
