"""STL surfaces (crates/core/src/stl.rs): binary and ASCII reading with
vertex welding, so the voxelizer gets one connected surface, and binary
writing.

Binary is recognised by its size (84 + 50 n bytes, stl.rs:17-24), since
binary files may also begin with "solid". Welding merges vertices that fall
on the same point of a grid of 1e-6 of the part's largest size (nearest
grid point, halves away from zero) and drops triangles that lose a corner
(stl.rs:60-92)."""
from __future__ import annotations

import struct

import numpy as np

from .shapes import TriMesh


def _is_binary(b: bytes) -> bool:
    if len(b) < 84:
        return False
    n = struct.unpack_from("<I", b, 80)[0]
    return 84 + n * 50 == len(b)


def _read_binary(b: bytes) -> np.ndarray:
    n = struct.unpack_from("<I", b, 80)[0]
    rec = np.frombuffer(b, dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]), count=n, offset=84)
    return rec["v"].astype(np.float64)


def _read_ascii(b: bytes) -> np.ndarray:
    try:
        s = b.decode("utf-8")
    except UnicodeDecodeError:
        raise ValueError("not a valid STL file") from None
    out, cur = [], []
    for line in s.splitlines():
        it = line.split()
        if it and it[0] == "vertex":
            v = []
            for x in it[1:4]:
                try:
                    v.append(float(x))
                except ValueError:
                    pass
            if len(v) != 3:
                raise ValueError(f"bad vertex line in STL: {line}")
            cur.append(v)
            if len(cur) == 3:
                out.append(cur)
                cur = []
    return np.asarray(out, np.float64).reshape(-1, 3, 3)


def weld(tris: np.ndarray) -> TriMesh:
    """stl.rs:60-92: one vertex per 1e-6-of-size grid point, first-seen order."""
    tris = np.asarray(tris, np.float64).reshape(-1, 3, 3)
    if len(tris) == 0:
        return TriMesh()
    pts = tris.reshape(-1, 3)
    size = max(float((pts.max(axis=0) - pts.min(axis=0)).max()), 1e-12)
    q = 1e-6 * size
    v = pts / q
    keys = (np.sign(v) * np.floor(np.abs(v) + 0.5)).astype(np.int64)
    _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    inv = inv.reshape(-1)
    # number the welded vertices in order of first appearance
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    idx = rank[inv].reshape(-1, 3)
    positions = pts[first[order]]
    keep = (idx[:, 0] != idx[:, 1]) & (idx[:, 1] != idx[:, 2]) & (idx[:, 0] != idx[:, 2])
    return TriMesh(positions.tolist(), idx[keep].tolist())


def read_stl(data: bytes) -> TriMesh:
    """stl.rs:8-15."""
    tris = _read_binary(data) if _is_binary(data) else _read_ascii(data)
    if len(tris) == 0:
        raise ValueError("the STL has no triangles")
    return weld(tris)


def write_binary_stl(m: TriMesh) -> bytes:
    """stl.rs:94-107: 80-byte header, count, per triangle its unit normal,
    three corners (f32) and a zero attribute word."""
    P = np.asarray(m.positions, np.float64)
    T = np.asarray(m.triangles, np.int64).reshape(-1, 3)
    rec = np.zeros(len(T), dtype=np.dtype([("n", "<f4", 3), ("v", "<f4", (3, 3)), ("a", "<u2")]))
    if len(T):
        a, b, c = P[T[:, 0]], P[T[:, 1]], P[T[:, 2]]
        n = np.cross(b - a, c - a)
        ln = np.linalg.norm(n, axis=1, keepdims=True)
        rec["n"] = n / np.where(ln > 0, ln, 1.0)
        rec["v"] = np.stack([a, b, c], axis=1)
    return bytes(80) + struct.pack("<I", len(T)) + rec.tobytes()
