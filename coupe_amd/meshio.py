"""Mesh, weight and partition files of the coupe tools, for mesh-part style runs.

* MEDIT meshes, ASCII and binary (mesh-io/src/medit/parser.rs:197-345 and
  :386-575): the sections the tools read (Vertices, Edges, Triangles,
  Quadrilaterals, Quadrangles, Tetrahedra, Hexahedra; Corners, Ridges and
  RequiredVertices are skipped).  Element node indices are stored 0-based,
  as mesh-io does.
* MeWe weight files (mesh-io/src/weight.rs, ABNF in tools/doc/weight-gen.1.scd
  "WEIGHT FILE").
* MePe partition files (mesh-io/src/partition.rs, ABNF in
  tools/doc/mesh-part.1.scd "PARTITION FILE").

Parsing is host work (numpy); the points the tools partition, the element
barycentres, are computed on the GPU by ``coupe_amd.tools.barycentres``.
"""
from __future__ import annotations

import io
import struct
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

# name -> (dimension, nodes per element); mesh-io/src/lib.rs:24-56
ELEMENT_TYPES = {
    "vertex": (0, 1),
    "edges": (1, 2),
    "triangles": (2, 3),
    "quadrilaterals": (2, 4),
    "quadrangles": (2, 4),
    "tetrahedra": (3, 4),
    "hexahedra": (3, 8),
}
# binary MEDIT keywords (mesh-io/src/medit/mod.rs:14-23)
CODE_DIMENSION, CODE_VERTEX, CODE_END = 3, 4, 54
CODE_ELEMENTS = {5: "edges", 6: "triangles", 7: "quadrilaterals", 8: "tetrahedra", 9: "hexahedra"}


class FormatError(ValueError):
    """A file that does not follow its format (coupe's tools fail on it too)."""


@dataclass
class Mesh:
    """mesh_io::Mesh: coordinates (n_nodes x dimension), node references,
    and blocks of elements of one type: (type name, 0-based node indices
    (n x nodes per element), element references)."""

    dimension: int
    coordinates: np.ndarray
    node_refs: np.ndarray
    topology: List[Tuple[str, np.ndarray, np.ndarray]] = field(default_factory=list)

    @property
    def node_count(self) -> int:
        return int(self.coordinates.shape[0])

    @property
    def element_count(self) -> int:
        return sum(int(nodes.shape[0]) for _, nodes, _ in self.topology)


# ---------------------------------------------------------------------------
# MEDIT
# ---------------------------------------------------------------------------
def read_mesh(source) -> Mesh:
    """Mesh::from_reader (mesh-io/src/lib.rs:127-140): binary if the file
    starts with the 4-byte magic 1 (either byte order), ASCII if it starts
    with "MeshVersionFormatted" (any case, after blanks)."""
    data = source if isinstance(source, (bytes, bytearray)) else open(source, "rb").read()
    if len(data) >= 4 and data[:4] in (b"\x01\x00\x00\x00", b"\x00\x00\x00\x01"):
        return parse_binary(bytes(data))
    head = data[:64].lstrip().lower()
    if head.startswith(b"meshversionformatted"):
        return parse_ascii(bytes(data).decode("utf-8"))
    raise FormatError("unknown format")


def _lines(text):
    """Non-blank lines with their 1-based line numbers."""
    for no, line in enumerate(text.split("\n"), 1):
        if line.strip():
            yield no, line


def parse_ascii(text: str) -> Mesh:
    """ASCII MEDIT (parser.rs:197-345).  Keywords are case-insensitive; a
    vertex or element is one line (extra tokens on a vertex line are an
    error, on an element line they are ignored after the reference); junk
    after a section keyword on the same line is skipped."""
    lines = list(_lines(text))
    pos = 0
    tokens: List[Tuple[int, str]] = []   # pending tokens of the current line

    def next_token():
        nonlocal pos
        while not tokens:
            if pos >= len(lines):
                raise FormatError("unexpected end of file")
            no, line = lines[pos]
            pos += 1
            tokens.extend((no, t) for t in line.split())
        return tokens.pop(0)

    def next_line():
        nonlocal pos
        if tokens:   # the rest of a line already split into tokens
            rest = [t for _, t in tokens]
            tokens.clear()
            return rest
        if pos >= len(lines):
            raise FormatError("unexpected end of file")
        _, line = lines[pos]
        pos += 1
        return line.split()

    def expect_int(tok):
        no, t = tok
        try:
            v = int(t)
        except ValueError:
            raise FormatError(f"at line {no}: expected an integer, found {t!r}") from None
        if v < 0:
            raise FormatError(f"at line {no}: expected an unsigned integer, found {t!r}")
        return v

    no, t = next_token()
    if t.lower() != "meshversionformatted":
        raise FormatError(f"at line {no}: expected token MeshVersionFormatted, found {t!r}")
    expect_int(next_token())
    no, t = next_token()
    if t.lower() != "dimension":
        raise FormatError(f"at line {no}: expected token Dimension, found {t!r}")
    dim = expect_int(next_token())
    mesh = Mesh(dim, np.zeros((0, dim)), np.zeros(0, dtype=np.int64))
    while True:
        no, t = next_token()
        key = t.lower()
        if key == "end":
            break
        if key == "vertices":
            n = expect_int(next_token())
            rows = [next_line() for _ in range(n)]
            coords = np.empty((n, dim), dtype=np.float64)
            refs = []
            for i, words in enumerate(rows):
                if len(words) < dim:
                    raise FormatError("unexpected end of vertex line")
                coords[i] = [float(x) for x in words[:dim]]
                if len(words) > dim:
                    refs.append(int(words[dim]))
                if len(words) > dim + 1:
                    raise FormatError(f"expected newline, found {words[dim + 1]!r}")
            mesh.coordinates = coords
            mesh.node_refs = np.array(refs, dtype=np.int64)
        elif key in ELEMENT_TYPES and key != "vertex":
            k = ELEMENT_TYPES[key][1]
            line_no = no
            while True:   # skip junk on the keyword's line (parser.rs:303-322)
                no2, t2 = next_token()
                try:
                    n = int(t2)
                    if n < 0:
                        raise ValueError
                    break
                except ValueError:
                    if no2 != line_no:
                        raise FormatError(f"at line {line_no}: expected the element count") from None
            nodes = np.empty((n, k), dtype=np.int64)
            refs = []
            for i in range(n):
                words = next_line()
                if len(words) < k:
                    raise FormatError("unexpected end of element line")
                idx = [int(x) for x in words[:k]]
                if min(idx) < 1:
                    raise FormatError("element node indices start at 1")
                nodes[i] = [x - 1 for x in idx]
                if len(words) > k:
                    refs.append(int(words[k]))
            mesh.topology.append((key, nodes, np.array(refs, dtype=np.int64)))
        elif key in ("corners", "ridges", "requiredvertices"):
            n = expect_int(next_token())
            for _ in range(n):
                next_line()
        else:
            raise FormatError(f"at line {no}: expected token element, found {t!r}")
    return mesh


def parse_binary(data: bytes) -> Mesh:
    """Binary MEDIT (parser.rs:386-575): magic 1 (its byte order sets the
    file's), version 1..4 setting the sizes of ints (i32/i64), floats
    (f32/f64) and field positions (i32/i64)."""
    r = io.BytesIO(data)

    def take(n):
        b = r.read(n)
        if len(b) != n:
            raise EOFError
        return b

    magic = take(4)
    end = "<" if magic == b"\x01\x00\x00\x00" else ">"
    key = lambda: struct.unpack(end + "i", take(4))[0]   # noqa: E731
    version = key()
    sizes = {1: ("i", "f", "i"), 2: ("i", "d", "i"), 3: ("i", "d", "q"), 4: ("q", "d", "q")}
    if version not in sizes:
        raise FormatError(f"expected token 4, found {version}")
    fi, ff, fp = sizes[version]
    rd = lambda f: struct.unpack(end + f, take(struct.calcsize(f)))[0]   # noqa: E731
    if key() != CODE_DIMENSION:
        raise FormatError("expected the dimension field")
    rd(fp)
    dim = key()
    mesh = Mesh(dim, np.zeros((0, dim)), np.zeros(0, dtype=np.int64))
    while True:
        try:
            code = key()
        except EOFError:
            break
        if code == CODE_END:
            break
        if code == CODE_VERTEX:
            rd(fp)
            n = rd(fi)
            rec = np.dtype([("c", end + ff, (dim,)), ("r", end + fi)])
            arr = np.frombuffer(take(rec.itemsize * n), dtype=rec, count=n)
            mesh.coordinates = arr["c"].astype(np.float64).reshape(n, dim)
            mesh.node_refs = arr["r"].astype(np.int64)
            continue
        if code not in CODE_ELEMENTS:
            raise FormatError(f"expected token 54, found {code}")
        name = CODE_ELEMENTS[code]
        k = ELEMENT_TYPES[name][1]
        rd(fp)
        n = rd(fi)
        rec = np.dtype([("v", end + fi, (k,)), ("r", end + fi)])
        arr = np.frombuffer(take(rec.itemsize * n), dtype=rec, count=n)
        nodes = arr["v"].astype(np.int64).reshape(n, k)
        if n and nodes.min() < 1:
            raise FormatError("element node indices start at 1")
        mesh.topology.append((name, nodes - 1, arr["r"].astype(np.int64)))
    return mesh


# ---------------------------------------------------------------------------
# MeWe weights
# ---------------------------------------------------------------------------
WEIGHT_VERSION = 1
FLAG_INTEGER = 1


def read_weights(source):
    """weight::read (mesh-io/src/weight.rs:78-111): returns an array of shape
    (weight count, criterion count), int64 for integer files, float64
    otherwise.  A file with no criterion is an empty integer array."""
    data = source if isinstance(source, (bytes, bytearray)) else open(source, "rb").read()
    if len(data) < 8 or data[:4] != b"MeWe":
        raise FormatError("bad file header")
    if data[4] != WEIGHT_VERSION:
        raise FormatError("unsupported file version")
    integer = bool(data[5] & FLAG_INTEGER)
    crit = struct.unpack("<H", data[6:8])[0]
    if crit == 0:
        return np.zeros((0, 0), dtype=np.int64)
    if len(data) < 16:
        raise FormatError("read/write error")
    count = struct.unpack("<Q", data[8:16])[0]
    need = 16 + 8 * crit * count
    if len(data) < need:
        raise FormatError("read/write error")
    return np.frombuffer(data, dtype="<i8" if integer else "<f8", count=crit * count, offset=16).reshape(
        count, crit).astype(np.int64 if integer else np.float64)


def write_weights(path_or_file, weights) -> bytes:
    """weight::write_integers / write_floats (weight.rs:114-191): int64
    arrays are written as integers, anything else as floats.  The bytes are
    returned and, when a path or file is given, written there."""
    w = np.asarray(weights)
    if w.ndim == 1:
        w = w[:, None]
    integer = np.issubdtype(w.dtype, np.integer)
    flags = FLAG_INTEGER if integer else 0
    if w.shape[0] == 0:
        out = b"MeWe" + bytes([WEIGHT_VERSION, flags]) + bytes(10)
    else:
        if w.shape[1] >= 1 << 16:
            raise ValueError("too many criteria")
        body = np.ascontiguousarray(w, dtype="<i8" if integer else "<f8").tobytes()
        out = b"MeWe" + bytes([WEIGHT_VERSION, flags]) + struct.pack("<HQ", w.shape[1], w.shape[0]) + body
    _emit(path_or_file, out)
    return out


# ---------------------------------------------------------------------------
# MePe partitions
# ---------------------------------------------------------------------------
def read_partition(source) -> np.ndarray:
    """partition::read (mesh-io/src/partition.rs:45-68)."""
    data = source if isinstance(source, (bytes, bytearray)) else open(source, "rb").read()
    if len(data) < 4 or data[:4] != b"MePe":
        raise FormatError("bad file header")
    if len(data) < 12:
        raise FormatError("read/write error")
    count = struct.unpack("<Q", data[4:12])[0]
    if len(data) < 12 + 8 * count:
        raise FormatError("read/write error")
    return np.frombuffer(data, dtype="<u8", count=count, offset=12).astype(np.uint64)


def write_partition(path_or_file, ids) -> bytes:
    """partition::write (mesh-io/src/partition.rs:70-86)."""
    a = np.ascontiguousarray(np.asarray(ids), dtype="<u8")
    out = b"MePe" + struct.pack("<Q", a.shape[0]) + a.tobytes()
    _emit(path_or_file, out)
    return out


def _emit(path_or_file, data: bytes):
    if path_or_file is None:
        return
    if hasattr(path_or_file, "write"):
        path_or_file.write(data)
    else:
        with open(path_or_file, "wb") as f:
            f.write(data)
