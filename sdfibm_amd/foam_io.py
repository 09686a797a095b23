"""OpenFOAM file I/O for the coupling step: binary/ASCII polyMesh readers and
``List<scalar>`` field readers/writers.

This is the data-format layer either side of the coupling path: the reference
solver reads its mesh through OpenFOAM's ``fvMesh`` (src/meshinfo.h:22-29) and
writes ``As`` through ``volScalarField::write`` (src/solidcloud.cpp:276-286);
the VOF initialiser writes ``alpha`` the same way (tool_vof/solidcloud.cpp:141-173).
OpenFOAM itself is not available here, so the two formats the path touches are
read and written directly:

* ``polyMesh/{points,faces,owner,neighbour}`` in ``format binary`` (32-bit
  labels, 64-bit scalars) or ``format ascii``;
* ``volScalarField`` files whose ``internalField`` is ``nonuniform List<scalar>``.

Everything is numpy; nothing here touches the GPU.
"""
from __future__ import annotations

import os
import re
from dataclasses import dataclass

import numpy as np

__all__ = [
    "read_foam_header",
    "read_points",
    "read_label_list",
    "read_faces",
    "read_scalar_field",
    "write_scalar_field",
    "PolyMesh",
    "read_polymesh",
]

_HEADER_END = re.compile(rb"\n// \*[^\n]*\n")


def _split_header(raw: bytes):
    """Return (header dict, byte offset just past the FoamFile block)."""
    i = raw.find(b"FoamFile")
    if i < 0:
        raise ValueError("not an OpenFOAM file: no FoamFile header")
    j = raw.find(b"}", i)
    block = raw[raw.find(b"{", i) + 1:j].decode("ascii", "replace")
    hdr = {}
    for line in block.split(";"):
        parts = line.split(None, 1)
        if len(parts) == 2:
            hdr[parts[0].strip()] = parts[1].strip().strip('"')
    return hdr, j + 1


def read_foam_header(path: str) -> dict:
    with open(path, "rb") as f:
        raw = f.read(4096)
    return _split_header(raw)[0]


def _read_list_at(raw: bytes, pos: int, binary: bool, dtype, ncomp: int):
    """Parse one ``N ( ... )`` list starting at or after ``pos``.

    Returns (array of shape (N,) or (N, ncomp), position after the closing ')').
    """
    m = re.compile(rb"\s*(\d+)\s*\(").search(raw, pos)
    if m is None:
        raise ValueError("list size not found")
    n = int(m.group(1))
    start = m.end()
    if binary:
        itemsize = np.dtype(dtype).itemsize * ncomp
        nbytes = n * itemsize
        arr = np.frombuffer(raw, dtype=dtype, count=n * ncomp, offset=start).copy()
        end = start + nbytes
        if raw[end:end + 1] != b")":
            raise ValueError("binary list not terminated by ')': label/scalar size mismatch?")
        end += 1
    else:
        # ASCII: vectors are written as "(x y z)" per entry; scalars/labels bare.
        depth, k = 1, start
        while depth > 0:
            c = raw[k:k + 1]
            if c == b"(":
                depth += 1
            elif c == b")":
                depth -= 1
            k += 1
        body = raw[start:k - 1].replace(b"(", b" ").replace(b")", b" ")
        arr = np.array(body.split(), dtype=dtype)
        end = k
    if ncomp > 1:
        arr = arr.reshape(n, ncomp)
    elif arr.size != n:
        raise ValueError(f"expected {n} entries, got {arr.size}")
    return arr, end


def _load(path: str):
    with open(path, "rb") as f:
        raw = f.read()
    hdr, off = _split_header(raw)
    binary = hdr.get("format", "ascii") == "binary"
    return raw, hdr, off, binary


def read_points(path: str) -> np.ndarray:
    """``vectorField`` → float64 array (nPoints, 3)."""
    raw, _, off, binary = _load(path)
    arr, _ = _read_list_at(raw, off, binary, np.float64, 3)
    return arr


def read_label_list(path: str) -> np.ndarray:
    """``labelList`` (owner / neighbour) → int32 array."""
    raw, _, off, binary = _load(path)
    arr, _ = _read_list_at(raw, off, binary, np.int32, 1)
    return arr


def read_faces(path: str) -> list:
    """``faceCompactList`` (binary) or ``faceList`` (ASCII) → (offsets, flat point labels)."""
    raw, hdr, off, binary = _load(path)
    cls = hdr.get("class", "")
    if cls == "faceCompactList":
        offsets, pos = _read_list_at(raw, off, binary, np.int32, 1)
        flat, _ = _read_list_at(raw, pos, binary, np.int32, 1)
        return offsets.astype(np.int64), flat
    # ASCII faceList: "N( 4(a b c d) 4(...) ... )"
    m = re.compile(rb"\s*(\d+)\s*\(").search(raw, off)
    n = int(m.group(1))
    faces = re.findall(rb"(\d+)\(([\d\s]+)\)", raw[m.end():])
    if len(faces) != n:
        raise ValueError("faceList parse mismatch")
    offsets = np.zeros(n + 1, dtype=np.int64)
    flat = []
    for i, (cnt, body) in enumerate(faces):
        lab = [int(t) for t in body.split()]
        if len(lab) != int(cnt):
            raise ValueError("face size mismatch")
        flat.extend(lab)
        offsets[i + 1] = offsets[i] + len(lab)
    return offsets, np.array(flat, dtype=np.int32)


def read_scalar_field(path: str) -> np.ndarray:
    """``volScalarField`` with ``internalField nonuniform List<scalar>`` → float64 array.

    A ``uniform v`` internal field raises: its length is the mesh's, not the file's.
    """
    raw, _, off, binary = _load(path)
    k = raw.find(b"internalField", off)
    if k < 0:
        raise ValueError("no internalField")
    if raw[k:k + 80].split()[1] != b"nonuniform":
        raise ValueError("internalField is uniform; length not stored in the file")
    arr, _ = _read_list_at(raw, k, binary, np.float64, 1)
    return arr


def write_scalar_field(path: str, values: np.ndarray, name: str, binary: bool = True,
                       boundary: str | None = None) -> None:
    """Write ``values`` as a ``volScalarField`` ``internalField`` (what ``field.write()`` produces
    for the VOF initialiser and for ``As`` after ``initialCorrect``).

    ``boundary`` is the text of the ``boundaryField`` block; by default every patch is written as
    ``".*" { type zeroGradient; }``.
    """
    v = np.ascontiguousarray(values, dtype=np.float64)
    fmt = "binary" if binary else "ascii"
    head = (
        "FoamFile\n{\n    version     2.0;\n    format      %s;\n    class       volScalarField;\n"
        "    location    \"0\";\n    object      %s;\n}\n"
        "// * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * * //\n\n"
        "dimensions      [0 0 0 0 0 0 0];\n\ninternalField   nonuniform List<scalar> \n%d\n("
        % (fmt, name, v.size)
    )
    if boundary is None:
        boundary = '    ".*"\n    {\n        type            zeroGradient;\n    }\n'
    with open(path, "wb") as f:
        f.write(head.encode())
        if binary:
            f.write(v.tobytes())
        else:
            f.write(b"\n" + b"\n".join(repr(float(x)).encode() for x in v) + b"\n")
        f.write((")\n;\n\nboundaryField\n{\n%s}\n" % boundary).encode())


@dataclass
class PolyMesh:
    """A polyMesh as read from disk: geometry is derived in :mod:`sdfibm_amd.mesh`."""

    points: np.ndarray       # (nPoints, 3) float64
    face_offsets: np.ndarray  # (nFaces+1,) int64
    face_points: np.ndarray  # flat int32
    owner: np.ndarray        # (nFaces,) int32
    neighbour: np.ndarray    # (nInternalFaces,) int32

    @property
    def n_cells(self) -> int:
        top = int(self.owner.max())
        if self.neighbour.size:
            top = max(top, int(self.neighbour.max()))
        return top + 1

    @property
    def n_faces(self) -> int:
        return self.owner.size

    def face(self, i: int) -> np.ndarray:
        return self.face_points[self.face_offsets[i]:self.face_offsets[i + 1]]


def read_polymesh(case_dir: str) -> PolyMesh:
    d = os.path.join(case_dir, "constant", "polyMesh")
    if not os.path.isdir(d):
        d = case_dir
    pts = read_points(os.path.join(d, "points"))
    off, flat = read_faces(os.path.join(d, "faces"))
    own = read_label_list(os.path.join(d, "owner"))
    nei = read_label_list(os.path.join(d, "neighbour"))
    return PolyMesh(pts, off, flat, own, nei)
