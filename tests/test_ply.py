"""PLY loader of the product (CPU) against the reference's exact-value tests
(crates/gsplat-io-ply/src/lib.rs:1005-1366) and fixtures."""

import pathlib
import struct

import numpy as np
import pytest

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"

HEADER14 = ("ply\nformat {fmt} 1.0\nelement vertex {n}\nproperty float x\nproperty float y\nproperty float z\n"
            "property float opacity\nproperty float scale_0\nproperty float scale_1\nproperty float scale_2\n"
            "property float rot_0\nproperty float rot_1\nproperty float rot_2\nproperty float rot_3\n"
            "property float f_dc_0\nproperty float f_dc_1\nproperty float f_dc_2\n")
VALID = HEADER14.format(fmt="ascii", n=1) + "end_header\n0.0 0.1 1.0 0.9 1.0 1.1 1.2 1.0 0.0 0.0 0.0 0.2 0.3 0.4\n"
NON_ID_QUAT = HEADER14.format(fmt="ascii", n=1) + \
    "end_header\n0.0 0.1 1.0 0.9 1.0 1.1 1.2 0.9 0.2 0.3 0.4 0.2 0.3 0.4\n"
REST1 = "".join(f"property float f_rest_{i}\n" for i in range(9))
DEG1 = HEADER14.format(fmt="ascii", n=1) + REST1 + \
    "end_header\n0.0 0.1 1.0 0.9 1.0 1.1 1.2 1.0 0.0 0.0 0.0 0.2 0.3 0.4 0.01 0.02 0.03 0.04 0.05 0.06 0.07 0.08 0.09\n"
INCOMPLETE = HEADER14.format(fmt="ascii", n=1) + "property float f_rest_0\n" + \
    "end_header\n0.0 0.1 1.0 0.9 1.0 1.1 1.2 1.0 0.0 0.0 0.0 0.2 0.3 0.4 0.125\n"
BASE = [0.0, 0.1, 1.0, 0.9, 1.0, 1.1, 1.2, 1.0, 0.0, 0.0, 0.0, 0.2, 0.3, 0.4]


def f32(v):
    return np.float32(v)


def err(abi, fn):
    with pytest.raises(abi.GsplatError) as e:
        fn()
    return e.value


def test_parses_valid_ascii_ply(abi):
    s = abi.ply_parse(VALID.encode(), text=True)
    assert s.count == 1 and s.sh_degree == 0 and s.sh_rest is None
    assert s.positions[0, 1] == f32(-0.1) and s.positions[0, 2] == 1.0
    assert list(s.rotation_xyzw[0]) == [0, 0, 0, 1]


def test_flips_quaternion_components_for_rdf_to_ruf(abi):
    s = abi.ply_parse(NON_ID_QUAT.encode(), text=True)
    assert list(s.rotation_xyzw[0]) == [f32(-0.2), f32(0.3), f32(-0.4), f32(0.9)]


def test_incomplete_sh_rest_is_ignored(abi):
    s = abi.ply_parse(INCOMPLETE.encode(), text=True)
    assert s.sh_degree == 0 and s.sh_rest is None


def test_parses_sh_rest_when_present_and_complete(abi):
    s = abi.ply_parse(DEG1.encode(), text=True)
    assert s.sh_degree == 1 and s.sh_rest.shape == (1, 9)
    sh = s.sh_rest[0]
    assert sh[0] == f32(-0.01) and sh[1] == f32(0.02) and sh[2] == f32(0.03) and sh[3] == f32(-0.04)
    assert sh[8] == f32(0.09)


def test_rejects_missing_required_field(abi):
    e = err(abi, lambda: abi.ply_parse(VALID.replace("property float opacity\n", "").encode(), text=True))
    assert e.code == abi.PARSE_FAILED and e.message.endswith("missing required field `opacity`")


def test_rejects_non_ascii_format_in_text_parser(abi):
    bad = VALID.replace("format ascii 1.0", "format binary_little_endian 1.0")
    assert err(abi, lambda: abi.ply_parse(bad.encode(), text=True)).code == abi.UNSUPPORTED


def test_rejects_ascii_non_finite_value(abi):
    bad = VALID.replace("0.0 0.1", "NaN 0.1", 1)
    e = err(abi, lambda: abi.ply_parse(bad.encode(), text=True))
    assert e.code == abi.PARSE_FAILED and e.message.endswith("failed to parse numeric value")


def test_normalizes_ascii_infinite_opacity_logits(abi):
    pos = VALID.replace("0.0 0.1 1.0 0.9", "0.0 0.1 1.0 inf", 1)
    neg = VALID.replace("0.0 0.1 1.0 0.9", "0.0 0.1 1.0 -inf", 1)
    assert abi.ply_parse(pos.encode(), text=True).opacity[0] == 16.0
    assert abi.ply_parse(neg.encode(), text=True).opacity[0] == -16.0
    nan = VALID.replace("0.0 0.1 1.0 0.9", "0.0 0.1 1.0 NaN", 1)
    assert err(abi, lambda: abi.ply_parse(nan.encode(), text=True)).code == abi.PARSE_FAILED


def test_rejects_ascii_vertex_count_mismatch(abi):
    bad = VALID.replace("element vertex 1", "element vertex 2")
    e = err(abi, lambda: abi.ply_parse(bad.encode(), text=True))
    assert e.message.endswith("vertex rows do not match declared vertex count")


def test_rejects_vertex_count_above_default_limit(abi):
    bad = VALID.replace("element vertex 1", "element vertex 5000001")
    e = err(abi, lambda: abi.ply_parse(bad.encode(), text=True))
    assert e.code == abi.UNSUPPORTED
    assert e.message.endswith("PLY resource limit exceeded for vertices: requested 5000001, limit 5000000")


def test_enforces_custom_input_and_header_limits(abi):
    lim = abi.ply_limits_default()
    lim.max_input_bytes = len(VALID) - 1
    e = err(abi, lambda: abi.ply_parse(VALID.encode(), lim))
    assert e.message.endswith(f"input bytes: requested {len(VALID)}, limit {len(VALID) - 1}")
    lim = abi.ply_limits_default()
    lim.max_header_bytes = 3
    e = err(abi, lambda: abi.ply_parse(VALID.encode(), lim, text=True))
    assert "PLY resource limit exceeded for header bytes" in e.message


def test_enforces_custom_property_and_scene_limits(abi):
    lim = abi.ply_limits_default()
    lim.max_vertex_properties = 13
    e = err(abi, lambda: abi.ply_parse(VALID.encode(), lim, text=True))
    assert e.message.endswith("vertex properties: requested 14, limit 13")
    lim = abi.ply_limits_default()
    lim.max_scene_bytes = 55
    e = err(abi, lambda: abi.ply_parse(VALID.encode(), lim, text=True))
    assert e.message.endswith("decoded scene bytes: requested 56, limit 55")


def test_parses_header_without_trailing_newline(abi):
    s = abi.ply_parse((HEADER14.format(fmt="ascii", n=0) + "end_header").encode())
    assert s.count == 0


@pytest.mark.parametrize("fmt,pack", [("binary_little_endian", "<14f"), ("binary_big_endian", ">14f")])
def test_parses_binary_ply(abi, fmt, pack):
    data = (HEADER14.format(fmt=fmt, n=1) + "end_header\n").encode() + struct.pack(pack, *BASE)
    s = abi.ply_parse(data)
    assert s.count == 1 and s.sh_degree == 0
    assert s.positions[0, 1] == f32(-0.1) and s.opacity[0] == f32(0.9)
    assert list(s.rotation_xyzw[0]) == [0, 0, 0, 1]


def test_rejects_truncated_binary_body(abi):
    data = (HEADER14.format(fmt="binary_little_endian", n=2) + "end_header\n").encode() + struct.pack("<14f", *BASE)
    e = err(abi, lambda: abi.ply_parse(data))
    assert e.message.endswith("vertex rows do not match declared vertex count")


def test_rejects_binary_body_size_overflow_before_allocation(abi):
    data = (HEADER14.format(fmt="binary_little_endian", n=2 ** 64 - 1) + "end_header\n").encode()
    lim = abi.ply_limits_default()
    lim.max_vertices = 2 ** 64 - 1
    e = err(abi, lambda: abi.ply_parse(data, lim))
    assert e.code == abi.UNSUPPORTED and e.message.endswith("overflow while computing binary body bytes")


def test_rejects_binary_non_finite_value(abi):
    data = (HEADER14.format(fmt="binary_little_endian", n=1) + "end_header\n").encode() + \
        struct.pack("<14f", float("nan"), *BASE[1:])
    assert err(abi, lambda: abi.ply_parse(data)).code == abi.PARSE_FAILED


def test_normalizes_binary_infinite_opacity_logit(abi):
    vals = list(BASE)
    vals[3] = float("inf")
    data = (HEADER14.format(fmt="binary_little_endian", n=1) + "end_header\n").encode() + struct.pack("<14f", *vals)
    assert abi.ply_parse(data).opacity[0] == 16.0


def test_parses_binary_sh_rest_degree_1(abi):
    data = (HEADER14.format(fmt="binary_little_endian", n=1) + REST1 + "end_header\n").encode() + \
        struct.pack("<14f", *BASE) + struct.pack("<9f", *[0.01 * (i + 1) for i in range(9)])
    s = abi.ply_parse(data)
    assert s.sh_degree == 1 and s.sh_rest[0, 0] == f32(-0.01) and s.sh_rest[0, 8] == f32(0.09)


def test_integer_and_double_property_types(abi):
    hdr = HEADER14.format(fmt="binary_little_endian", n=1).replace("property float x", "property double x") \
        .replace("property float opacity", "property int opacity").replace("property float f_dc_0",
                                                                          "property uchar f_dc_0")
    body = struct.pack("<d", 0.0) + struct.pack("<2f", 0.1, 1.0) + struct.pack("<i", -3) + \
        struct.pack("<7f", 1.0, 1.1, 1.2, 1.0, 0.0, 0.0, 0.0) + struct.pack("<B", 200) + struct.pack("<2f", 0.3, 0.4)
    s = abi.ply_parse((hdr + "end_header\n").encode() + body)
    assert s.opacity[0] == -3.0 and s.color_dc[0, 0] == 200.0


def test_minimal_fixtures_match(abi):
    a = abi.ply_load(GOLDEN / "minimal_ascii.ply")
    b = abi.ply_load(GOLDEN / "minimal_binary.ply")
    assert a.count == 3 and b.count == 3
    for name in ("positions", "opacity", "scale_xyz", "rotation_xyzw", "color_dc"):
        assert np.array_equal(getattr(a, name), getattr(b, name)), name
    assert a.positions[1, 1] == f32(-0.1)


def test_missing_file_is_not_found(abi):
    e = err(abi, lambda: abi.ply_load(GOLDEN / "does_not_exist.ply"))
    assert e.code == abi.NOT_FOUND and e.message.endswith("I/O error while reading PLY")


def test_list_property_is_unsupported(abi):
    bad = VALID.replace("property float x\n", "property list uchar int x\n")
    assert err(abi, lambda: abi.ply_parse(bad.encode(), text=True)).code == abi.UNSUPPORTED


def test_rot_layout_env_override(abi, monkeypatch):
    monkeypatch.setenv("GSPLAT_ROT_LAYOUT", "xyzw")
    s = abi.ply_parse(NON_ID_QUAT.encode(), text=True)
    # raw (0.9, 0.2, 0.3, 0.4) taken as xyzw, then x and z negated
    assert list(s.rotation_xyzw[0]) == [f32(-0.9), f32(0.2), f32(-0.3), f32(0.4)]
