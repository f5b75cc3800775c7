"""File formats of the partition -> voxelize -> SVO-build path, in Python.

* ``.tri`` text header + ``.tridata`` records: tri_tools.h:75-127, tri_util.h:7-11.
* ``.octree`` text header + ``.octreenodes`` (24 B) + ``.octreedata`` (32 B):
  octree_io.h:49-110, Node.h:16-18, VoxelData.h:10-18.

The header writers reproduce the reference's iostream output byte for byte for the
values used here (integers, and floats that C++'s default 6-significant-digit
formatting prints like Python's ``%g``).
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

NODE_DTYPE = np.dtype([("data", "<u8"), ("children_base", "<u8"), ("offsets", "i1", (8,))])
DATA_DTYPE = np.dtype([("morton", "<u8"), ("color", "<f4", (3,)), ("normal", "<f4", (3,))])
assert NODE_DTYPE.itemsize == 24 and DATA_DTYPE.itemsize == 32


def _fmt_float(v: float) -> str:
    """C++ ostream default float formatting (%g, precision 6)."""
    return "%g" % float(np.float32(v))


@dataclass
class TriInfo:
    base: str
    n_triangles: int
    geo_only: int
    bbox: tuple = (0.0, 0.0, 0.0, 1.0, 1.0, 1.0)
    version: int = 1


def write_tri(base: str, records: np.ndarray, payload: bool,
              bbox=(0.0, 0.0, 0.0, 1.0, 1.0, 1.0)) -> str:
    """Write ``base.tri`` + ``base.tridata`` (tri_tools.h:118-127); returns the .tri path."""
    rec = np.ascontiguousarray(records, dtype="<f4")
    assert rec.ndim == 2 and rec.shape[1] == (21 if payload else 9)
    with open(base + ".tri", "w") as f:
        f.write("#tri 1\n")
        f.write("ntriangles %d\n" % rec.shape[0])
        f.write("geo_only %d\n" % (0 if payload else 1))
        f.write("bbox  " + " ".join(_fmt_float(b) for b in bbox) + "\n")
        f.write("END\n")
    rec.tofile(base + ".tridata")
    return base + ".tri"


def read_tri(path: str) -> tuple[TriInfo, np.ndarray]:
    """Parse a .tri header (tri_tools.h:75-115) and load its .tridata records."""
    base = path[: path.rfind(".")]
    toks = open(path).read().split()
    if not toks or toks[0] != "#tri":
        raise ValueError("first token is not #tri")
    version = int(toks[1])
    i = 2
    n = 0
    geo = 0
    bbox = (0.0,) * 6
    done = False
    while i < len(toks):
        k = toks[i]
        if k == "END":
            done = True
            break
        if k == "ntriangles":
            n = int(toks[i + 1]); i += 2
        elif k == "geo_only":
            geo = int(toks[i + 1]); i += 2
        elif k == "bbox":
            bbox = tuple(float(np.float32(t)) for t in toks[i + 1:i + 7]); i += 7
        else:
            i += 1
    if not done:
        raise ValueError("error reading header")
    stride = 9 if geo else 21
    rec = np.fromfile(base + ".tridata", dtype="<f4")
    if rec.size < n * stride:
        raise ValueError("%s.tridata holds fewer than %d triangles" % (base, n))
    return TriInfo(base, n, geo, bbox, version), rec[: n * stride].reshape(n, stride)


@dataclass
class Octree:
    gridlength: int
    nodes: np.ndarray  # NODE_DTYPE
    data: np.ndarray   # DATA_DTYPE
    header_text: str = field(default="")


def octree_header_text(gridlength: int, n_nodes: int, n_data: int) -> str:
    """octree_io.h:74-83."""
    return "#octreeheader 1\ngridlength %d\nn_nodes %d\nn_data %d\nEND\n" % (
        gridlength, n_nodes, n_data)


def read_octree(base: str) -> Octree:
    """Load ``base.octree`` + nodes + data (octree_io.h:56-110)."""
    text = open(base + ".octree").read()
    toks = text.split()
    if toks[0] != "#octreeheader":
        raise ValueError("not an octree header")
    vals = {toks[i]: toks[i + 1] for i in range(2, len(toks) - 1, 2) if toks[i] != "END"}
    G = int(vals["gridlength"])
    nn = int(vals["n_nodes"])
    nd = int(vals["n_data"])
    nodes = np.fromfile(base + ".octreenodes", dtype=NODE_DTYPE)
    data = np.fromfile(base + ".octreedata", dtype=DATA_DTYPE)
    if len(nodes) != nn or len(data) != nd:
        raise ValueError("octree files disagree with header")
    return Octree(G, nodes, data, text)


def octree_files(base: str) -> list[str]:
    return [base + ext for ext in (".octree", ".octreenodes", ".octreedata")]


def files_identical(base_a: str, base_b: str) -> bool:
    for fa, fb in zip(octree_files(base_a), octree_files(base_b)):
        if os.path.getsize(fa) != os.path.getsize(fb):
            return False
        with open(fa, "rb") as a, open(fb, "rb") as b:
            while True:
                x = a.read(1 << 24)
                y = b.read(1 << 24)
                if x != y:
                    return False
                if not x:
                    break
    return True
