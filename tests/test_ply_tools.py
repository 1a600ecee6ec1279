"""PLY writer and point-count ladder (gsplat-rs_amd/ply_tools.py), on the CPU.

The ladder follows the reference's integer midpoint rule (tests/datasets/
README.md:8-25); its known answers are those of the reference's own fixture
test (tests/datasets/test_dataset_tools.py:84-186: a 10-record PLY whose
4-record tier holds ids [1, 3, 6, 8], the header kept but for the vertex
count, a trailing element copied). The writer is pinned by round trips
through the product's PLY loader (gsplat_ply_load), which must return the
written floats bit for bit for every SH degree and both binary byte orders.
"""

import hashlib
import json
import pathlib
import struct

import numpy as np
import pytest

import ply_tools
import synth


def fixture_ply(count=10):
    """The reference fixture: int id + float x per record, then a 2-record metadata element."""
    header = ("ply\nformat binary_little_endian 1.0\ncomment fixture header must be retained\n"
              f"element vertex {count}\nproperty int id\nproperty float x\nelement metadata 2\n"
              "property ushort tag\nend_header\n").encode("ascii")
    body = b"".join(struct.pack("<if", i, i + 0.25) for i in range(count))
    suffix = struct.pack("<HH", 17, 29)
    return header + body + suffix, header, suffix


def read_ids(path, layout, n):
    with open(path, "rb") as f:
        f.seek(layout.vertex_offset)
        return [struct.unpack("<if", f.read(8))[0] for _ in range(n)]


def test_ladder_reference_fixture(tmp_path):
    data, header, suffix = fixture_ply()
    src = tmp_path / "scene.ply"
    src.write_bytes(data)
    prov = tmp_path / "source.json"
    prov.write_text(json.dumps({"id": "fixture", "license": "fixture-only", "local_path": str(src),
                                "sha256": hashlib.sha256(data).hexdigest(), "bytes": len(data), "splat_count": 10}))
    out = tmp_path / "ladder"
    m = ply_tools.generate_ladder(src, [7, 4, 4], out, source_manifest=prov)
    four = out / "scene-n4.ply"
    lay = ply_tools.inspect_binary_ply(four)
    assert lay.elements[lay.vertex].count == 4 and lay.elements[lay.vertex].properties == ("id", "x")
    assert lay.header == header.replace(b"element vertex 10", b"element vertex 4")
    assert four.read_bytes().endswith(suffix)
    assert read_ids(four, lay, 4) == [1, 3, 6, 8]
    seven = out / "scene-n7.ply"
    assert read_ids(seven, ply_tools.inspect_binary_ply(seven), 7) == [0, 2, 3, 5, 6, 7, 9]
    assert m["schema"] == "gsplat-ply-ladder/v1"
    assert m["selection"]["algorithm"] == "stratified-midpoint-integer/v1"
    assert [t["splat_count"] for t in m["tiers"]] == [4, 7]
    assert (m["tiers"][0]["selection_first_index"], m["tiers"][0]["selection_last_index"]) == (1, 8)
    assert m["source"]["provenance"]["license"] == "fixture-only"
    assert m["tiers"][0]["sha256"] == hashlib.sha256(four.read_bytes()).hexdigest()
    before = (out / "ladder.json").read_bytes()
    again = ply_tools.generate_ladder(src, [4, 7], out, source_manifest=prov, overwrite=True)
    assert [t["sha256"] for t in again["tiers"]] == [t["sha256"] for t in m["tiers"]]
    assert (out / "ladder.json").read_bytes() == before


def test_ladder_rejects_what_it_cannot_subset(tmp_path):
    asc = tmp_path / "ascii.ply"
    asc.write_text("ply\nformat ascii 1.0\nelement vertex 1\nproperty float x\nend_header\n0\n")
    with pytest.raises(ply_tools.PlyToolError, match="input must use binary"):
        ply_tools.inspect_binary_ply(asc)
    lst = tmp_path / "list.ply"
    lst.write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 1\nproperty list uchar int idx\n"
                    b"end_header\n\x00")
    with pytest.raises(ply_tools.PlyToolError, match="list properties"):
        ply_tools.inspect_binary_ply(lst)
    data, _, _ = fixture_ply()
    short = tmp_path / "short.ply"
    short.write_bytes(data[:-3])
    with pytest.raises(ply_tools.PlyToolError, match="payload size"):
        ply_tools.inspect_binary_ply(short)
    src = tmp_path / "scene.ply"
    src.write_bytes(data)
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps({"sha256": "0" * 64, "bytes": len(data), "splat_count": 10}))
    with pytest.raises(ply_tools.PlyToolError, match="source manifest mismatch"):
        ply_tools.generate_ladder(src, [4], tmp_path / "bad-ladder", source_manifest=bad)
    assert not (tmp_path / "bad-ladder" / "scene-n4.ply").exists()
    with pytest.raises(ply_tools.PlyToolError, match="tier counts"):
        ply_tools.generate_ladder(src, [11], tmp_path / "big")
    ply_tools.generate_ladder(src, [3], tmp_path / "once")
    with pytest.raises(ply_tools.PlyToolError, match="refusing to overwrite"):
        ply_tools.generate_ladder(src, [3], tmp_path / "once")


@pytest.mark.parametrize("m,n", [(1, 1), (1, 7), (7, 7), (3, 10), (50_000, 279_199), (1_000_000, 5_000_000)])
def test_midpoint_index_against_exact_integers(m, n):
    i = np.arange(m)
    got = ply_tools.midpoint_index(i, m, n)
    step = max(1, m // 997)
    for k in list(range(0, m, step)) + [m - 1]:  # Python's unbounded integers
        assert int(got[k]) == ((2 * k + 1) * n) // (2 * m)
    assert np.all(np.diff(got) > 0) and got[0] >= 0 and got[-1] < n  # distinct, ascending, in range
    if m == n:
        assert np.array_equal(got, i)


def test_parse_count():
    assert [ply_tools.parse_count(s) for s in ("50k", "1.5m", "300", "2M")] == [50_000, 1_500_000, 300, 2_000_000]
    for s in ("0", "1.2345k", "x"):
        with pytest.raises(Exception):
            ply_tools.parse_count(s)


@pytest.mark.parametrize("degree,fmt", [(0, "binary_little_endian"), (1, "binary_big_endian"),
                                        (3, "binary_little_endian"), (4, "binary_little_endian"), (2, "ascii")])
def test_written_scene_loads_back_bit_for_bit(abi, tmp_path, degree, fmt):
    s = synth.tiny_scene(257 if fmt == "ascii" else 3001, 40 + degree, sh_degree=degree)
    path = ply_tools.write_scene_ply(tmp_path / "scene.ply", s, fmt)
    got = abi.ply_load(path)
    assert got.count == s.count and got.sh_degree == degree
    for name in ("positions", "opacity", "scale_xyz", "rotation_xyzw", "color_dc"):
        a, b = np.asarray(getattr(got, name)).reshape(-1), np.asarray(getattr(s, name)).reshape(-1)
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), name
    if degree:
        assert np.array_equal(np.asarray(got.sh_rest).reshape(-1).view(np.uint32),
                              np.asarray(s.sh_rest).reshape(-1).view(np.uint32))
    else:
        assert got.sh_rest is None


def test_ladder_of_a_scene_is_its_midpoint_subset(abi, tmp_path):
    s = synth.tiny_scene(5000, 77, sh_degree=3)
    src = ply_tools.write_scene_ply(tmp_path / "full.ply", s)
    ply_tools.generate_ladder(src, [1234], tmp_path / "ladder")
    tier = abi.ply_load(tmp_path / "ladder" / "full-n1234.ply")
    idx = ply_tools.midpoint_index(np.arange(1234), 1234, 5000)
    assert tier.count == 1234
    assert np.array_equal(np.asarray(tier.positions).view(np.uint32), s.positions[idx].view(np.uint32))
    assert np.array_equal(np.asarray(tier.sh_rest).reshape(1234, -1).view(np.uint32),
                          s.sh_rest[idx].view(np.uint32))
