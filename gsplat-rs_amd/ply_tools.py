"""PLY scene files for the SortedAlpha path: a 3DGS writer and point-count tiers.

write_scene_ply    writes a scene (SceneBuffers layout, right-up-forward axes)
                   as a 3DGS-style PLY that gsplat_ply_load /
                   gsplat_context_load_scene_path read back to the same floats:
                   it applies the inverse of the loader's RDF->RUF conversion
                   (crates/gsplat-io-ply/src/lib.rs:76-114: y, quaternion x and z,
                   and the SH_FLIP_RDF_TO_RUF coefficient signs negated; sign
                   flips are exact) and stores quaternions w first (:59-74).
generate_ladder    point-count tiers of a fixed-record binary PLY by the
                   reference's integer midpoint stratification
                   (tests/datasets/README.md:8-25, tests/datasets/ply_ladder.py):
                   output record i of m is source record
                   floor(((2 i + 1) n) / (2 m)); every header line but
                   `element vertex` and every non-vertex element is kept byte for
                   byte, and the tiers' ladder.json uses the reference's
                   gsplat-ply-ladder/v1 layout, so its tooling reads it.

The records are selected with one vectorised gather over a memory map of the
vertex block, never decoding a float, so a tier is byte-identical on every
platform.
"""

from __future__ import annotations

import argparse
import dataclasses
import hashlib
import json
import os
import pathlib
import re
import sys
import tempfile

import numpy as np

LADDER_SCHEMA = "gsplat-ply-ladder/v1"
SELECTION_ALGORITHM = "stratified-midpoint-integer/v1"
HEADER_LIMIT = 1 << 20
SCALAR_BYTES = {"char": 1, "int8": 1, "uchar": 1, "uint8": 1, "short": 2, "int16": 2, "ushort": 2, "uint16": 2,
                "int": 4, "int32": 4, "uint": 4, "uint32": 4, "float": 4, "float32": 4, "double": 8, "float64": 8}
# Sign of SH rest coefficient k (k < 15) under the RDF -> RUF axis change
# (SH_FLIP_RDF_TO_RUF, io-ply/src/lib.rs:54-56); degree-4 coefficients keep theirs.
SH_FLIP = np.array([-1, 1, 1, -1, -1, 1, 1, 1, -1, -1, -1, 1, 1, 1, 1], np.float32)


class PlyToolError(ValueError):
    """An input the tools cannot handle without changing its record layout."""


# ------------------------------------------------------------------ writer

def write_scene_ply(path, scene, fmt: str = "binary_little_endian") -> pathlib.Path:
    """Writes `scene` (positions/opacity/scale_xyz/rotation_xyzw/color_dc/sh_rest float32 arrays,
    sh_degree) as a PLY of float properties x y z f_dc_0..2 f_rest_* opacity scale_0..2 rot_0..3."""
    if fmt not in ("binary_little_endian", "binary_big_endian", "ascii"):
        raise PlyToolError(f"unsupported format {fmt}")
    n = int(scene.positions.shape[0])
    pc = (scene.sh_degree + 1) ** 2 - 1
    pos = np.array(scene.positions, np.float32).reshape(n, 3)
    pos[:, 1] = -pos[:, 1]
    q = np.array(scene.rotation_xyzw, np.float32).reshape(n, 4)
    rot = np.stack([q[:, 3], -q[:, 0], q[:, 1], -q[:, 2]], axis=1)  # w, x, y, z in RDF
    cols = [pos, np.asarray(scene.color_dc, np.float32).reshape(n, 3)]
    names = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2"]
    if pc:
        rest = np.array(scene.sh_rest, np.float32).reshape(n, 3, pc)
        flip = np.ones(pc, np.float32)
        flip[:min(pc, 15)] = SH_FLIP[:min(pc, 15)]
        rest = rest * flip  # exact: the signs are +-1
        cols.append(rest.reshape(n, 3 * pc))
        names += [f"f_rest_{k}" for k in range(3 * pc)]
    cols += [np.asarray(scene.opacity, np.float32).reshape(n, 1), np.asarray(scene.scale_xyz, np.float32).reshape(n, 3),
             rot]
    names += ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"]
    body = np.ascontiguousarray(np.concatenate(cols, axis=1), dtype=np.float32)
    header = ["ply", f"format {fmt} 1.0", f"element vertex {n}", *[f"property float {p}" for p in names],
              "end_header", ""]
    path = pathlib.Path(path)
    with path.open("wb") as f:
        f.write("\n".join(header).encode("ascii"))
        if fmt == "ascii":
            for row in body:
                f.write((" ".join(repr(float(v)) for v in row) + "\n").encode("ascii"))
        else:
            f.write(body.astype(">f4" if fmt == "binary_big_endian" else "<f4").tobytes())
    return path


# ---------------------------------------------------------- layout + ladder

@dataclasses.dataclass(frozen=True)
class Element:
    name: str
    count: int
    record_bytes: int
    properties: tuple


@dataclasses.dataclass(frozen=True)
class BinaryLayout:
    header: bytes
    format: str
    elements: tuple
    vertex: int  # index of the vertex element

    @property
    def vertex_offset(self) -> int:
        return len(self.header) + sum(e.count * e.record_bytes for e in self.elements[:self.vertex])

    @property
    def payload_bytes(self) -> int:
        return sum(e.count * e.record_bytes for e in self.elements)


def _header_bytes(path: pathlib.Path) -> bytes:
    with path.open("rb") as f:
        head = f.read(HEADER_LIMIT + 1)
    m = re.search(rb"(?:^|\n)end_header\r?\n", head[:HEADER_LIMIT + 16])
    if m is None or m.end() > HEADER_LIMIT:
        raise PlyToolError(f"PLY header exceeds {HEADER_LIMIT} bytes or lacks end_header")
    return head[:m.end()]


def inspect_binary_ply(path) -> BinaryLayout:
    """Fixed-record layout of a binary PLY (scalar properties only), checked against the file size."""
    path = pathlib.Path(path)
    header = _header_bytes(path)
    try:
        lines = header.decode("ascii").splitlines()
    except UnicodeDecodeError as e:
        raise PlyToolError("PLY header must be ASCII") from e
    if not lines or lines[0] != "ply":
        raise PlyToolError("missing PLY magic")
    fmt, elems = None, []
    for line in lines[1:]:
        f = line.split()
        if not f or f[0] in ("comment", "obj_info"):
            continue
        if f[0] == "format":
            if len(f) != 3 or f[2] != "1.0":
                raise PlyToolError(f"unsupported PLY format declaration: {line}")
            fmt = f[1]
        elif f[0] == "element":
            if len(f) != 3 or not f[2].isdigit():
                raise PlyToolError(f"malformed element declaration: {line}")
            elems.append([f[1], int(f[2]), 0, []])
        elif f[0] == "property":
            if not elems:
                raise PlyToolError("property declared before an element")
            if len(f) >= 2 and f[1] == "list":
                raise PlyToolError("list properties have variable row sizes and cannot be subset safely")
            if len(f) != 3 or f[1] not in SCALAR_BYTES:
                raise PlyToolError(f"unsupported scalar property declaration: {line}")
            elems[-1][2] += SCALAR_BYTES[f[1]]
            elems[-1][3].append(f[2])
        elif f[0] == "end_header":
            break
    if fmt not in ("binary_little_endian", "binary_big_endian"):
        raise PlyToolError("input must use binary_little_endian 1.0 or binary_big_endian 1.0")
    elements = tuple(Element(e[0], e[1], e[2], tuple(e[3])) for e in elems)
    vertex = [i for i, e in enumerate(elements) if e.name == "vertex"]
    if len(vertex) != 1:
        raise PlyToolError("PLY must contain exactly one vertex element")
    if elements[vertex[0]].record_bytes == 0 or any(e.count and not e.record_bytes for e in elements):
        raise PlyToolError("an element with records has no properties")
    layout = BinaryLayout(header, fmt, elements, vertex[0])
    if path.stat().st_size != len(header) + layout.payload_bytes:
        raise PlyToolError(f"PLY payload size does not match its fixed-width header: expected="
                           f"{len(header) + layout.payload_bytes} actual={path.stat().st_size}")
    return layout


def midpoint_index(i, m: int, n: int):
    """Source record of output record i of m over n records: floor(((2i + 1) n) / (2m)), integers only
    (i may be a numpy integer array)."""
    if not 0 < m <= n:
        raise ValueError("expected 0 < output_count <= source_count")
    return ((2 * np.asarray(i, dtype=np.int64) + 1) * n) // (2 * m)


def _sha256_file(path: pathlib.Path) -> str:
    h = hashlib.sha256()
    with path.open("rb") as f:
        for chunk in iter(lambda: f.read(1 << 23), b""):
            h.update(chunk)
    return h.hexdigest()


def _portable(path: pathlib.Path) -> str:
    return path.resolve().as_posix()


def generate_ladder(source, counts, output_dir, *, source_manifest=None, overwrite=False) -> dict:
    """Writes OUTPUT_DIR/<stem>-n<count>.ply for each count and OUTPUT_DIR/ladder.json; returns the manifest."""
    source = pathlib.Path(source).resolve()
    out = pathlib.Path(output_dir).resolve()
    layout = inspect_binary_ply(source)
    v = layout.elements[layout.vertex]
    counts = sorted(set(int(c) for c in counts))
    if not counts:
        raise PlyToolError("at least one output count is required")
    if counts[0] <= 0 or counts[-1] > v.count:
        raise PlyToolError(f"tier counts must be in 1..{v.count}, got {counts}")
    prov = None
    if source_manifest is not None:
        try:
            prov = json.loads(pathlib.Path(source_manifest).read_text(encoding="utf-8"))
        except (OSError, json.JSONDecodeError) as e:
            raise PlyToolError(f"cannot read source manifest {source_manifest}: {e}") from e
        if not isinstance(prov, dict):
            raise PlyToolError("source manifest must contain a JSON object")
        local = prov.get("local_path")
        if isinstance(local, str):  # relative paths are taken from the manifest's folder
            where = pathlib.Path(local)
            where = where if where.is_absolute() else pathlib.Path(source_manifest).resolve().parent / where
            if where.resolve() != source:
                raise PlyToolError(f"source manifest local_path resolves to {where}, not {source}")
    source_sha, source_bytes = _sha256_file(source), source.stat().st_size
    if prov is not None:
        bad = [f"{k} expected={prov[k]} actual={a}" for k, a in
               (("sha256", source_sha), ("bytes", source_bytes), ("splat_count", v.count))
               if prov.get(k) is not None and prov[k] != a]
        if bad:
            raise PlyToolError("source manifest mismatch: " + "; ".join(bad))
    out.mkdir(parents=True, exist_ok=True)
    paths = [out / f"{source.stem}-n{c}.ply" for c in counts]
    manifest_path = out / "ladder.json"
    existing = [p for p in (*paths, manifest_path) if p.exists()]
    if existing and not overwrite:
        raise PlyToolError("refusing to overwrite existing output(s): " + ", ".join(map(str, existing)))
    records = np.memmap(source, dtype=np.uint8, mode="r", offset=layout.vertex_offset, shape=(v.count, v.record_bytes))
    with source.open("rb") as f:
        f.seek(len(layout.header))
        before = f.read(layout.vertex_offset - len(layout.header))
        f.seek(layout.vertex_offset + v.count * v.record_bytes)
        after = f.read()
    pattern = re.compile(rb"(?m)^(element[ \t]+vertex[ \t]+)([0-9]+)([ \t]*\r?)$")
    tiers = []
    for c, p in zip(counts, paths):
        header, k = pattern.subn(lambda m: m.group(1) + str(c).encode() + m.group(3), layout.header)
        if k != 1:
            raise PlyToolError("could not rewrite exactly one vertex count in PLY header")
        idx = midpoint_index(np.arange(c), c, v.count)
        fd, tmp = tempfile.mkstemp(prefix=f".{p.name}.", suffix=".tmp", dir=out)
        try:
            with os.fdopen(fd, "wb") as f:
                f.write(header)
                f.write(before)
                for lo in range(0, c, 1 << 18):  # gather in slices: bounded memory for large tiers
                    f.write(np.ascontiguousarray(records[idx[lo:lo + (1 << 18)]]).tobytes())
                f.write(after)
            os.replace(tmp, p)
        finally:
            if os.path.exists(tmp):
                os.unlink(tmp)
        tiers.append({"splat_count": c, "local_path": _portable(p), "sha256": _sha256_file(p),
                      "bytes": p.stat().st_size, "selection_first_index": int(idx[0]),
                      "selection_last_index": int(idx[-1])})
    del records
    manifest = {
        "schema": LADDER_SCHEMA,
        "source": {"local_path": _portable(source), "sha256": source_sha, "bytes": source_bytes,
                   "format": layout.format, "splat_count": v.count, "vertex_record_bytes": v.record_bytes,
                   "vertex_properties": list(v.properties),
                   "header_sha256": hashlib.sha256(layout.header).hexdigest()},
        "selection": {"algorithm": SELECTION_ALGORITHM,
                      "formula": "floor(((2 * output_index + 1) * source_count) / (2 * output_count))",
                      "ordering": "ascending source record index",
                      "nesting": "not guaranteed; pair backends on the exact same tier file",
                      "purpose": "performance scaling only; not a replacement for full-scene quality runs"},
        "tiers": tiers,
    }
    if prov is not None:
        manifest["source"]["provenance_manifest"] = _portable(pathlib.Path(source_manifest))
        manifest["source"]["provenance"] = prov
    manifest_path.write_text(json.dumps(manifest, indent=2) + "\n", encoding="utf-8")
    return manifest


def parse_count(text: str) -> int:
    """'50k' -> 50000, '1.5m' -> 1500000, '300' -> 300."""
    m = re.fullmatch(r"([0-9]+(?:\.[0-9]+)?)([kKmM]?)", text.strip())
    if m is None:
        raise argparse.ArgumentTypeError(f"invalid point count: {text!r}")
    scale = {"": 1, "k": 1000, "m": 1000000}[m.group(2).lower()]
    whole, _, frac = m.group(1).partition(".")
    digits = int(whole + frac) * scale
    div = 10 ** len(frac)
    if digits % div or digits // div <= 0:
        raise argparse.ArgumentTypeError(f"point count must be a positive integer: {text!r}")
    return digits // div


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Deterministic point-count tiers of a binary PLY scene.")
    ap.add_argument("source", type=pathlib.Path)
    ap.add_argument("--counts", nargs="+", type=parse_count, required=True)
    ap.add_argument("--output-dir", type=pathlib.Path, required=True)
    ap.add_argument("--source-manifest", type=pathlib.Path)
    ap.add_argument("--overwrite", action="store_true")
    a = ap.parse_args(argv)
    try:
        m = generate_ladder(a.source, a.counts, a.output_dir, source_manifest=a.source_manifest,
                            overwrite=a.overwrite)
    except PlyToolError as e:
        print(f"ply_tools: {e}", file=sys.stderr)
        return 2
    for t in m["tiers"]:
        print(t["local_path"])
    return 0


if __name__ == "__main__":
    sys.exit(main())
