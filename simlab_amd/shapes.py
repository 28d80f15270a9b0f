"""Test geometry as triangle meshes, restating crates/cfd/src/shapes.rs with
the same double-precision arithmetic (so voxelizations match): `cuboid`
(shapes.rs:8-20), `disk` (:22-45), icosphere `sphere` (:47-111),
`fix_orientation` (:245-280), plus `merge` (crates/cfd/tests/channel.rs:13-20)
and `cylinder` (a closed, outward-facing cylinder of any axis, used by the
drone-arm workload)."""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List


@dataclass
class TriMesh:
    """crates/core/src/mesh.rs TriMesh: positions and triangles."""
    positions: List[list] = field(default_factory=list)
    triangles: List[list] = field(default_factory=list)

    def translated(self, d) -> "TriMesh":
        return TriMesh([[p[0] + d[0], p[1] + d[1], p[2] + d[2]] for p in self.positions], [list(t) for t in self.triangles])

    def scaled(self, s: float) -> "TriMesh":
        return TriMesh([[p[0] * s, p[1] * s, p[2] * s] for p in self.positions], [list(t) for t in self.triangles])


def _add(a, b):
    return [a[0] + b[0], a[1] + b[1], a[2] + b[2]]


def _scale(a, s):
    return [a[0] * s, a[1] * s, a[2] * s]


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def _normalize(a):
    # crates/core/src/lib.rs normalize: scale by 1/|a|
    n = math.sqrt(_dot(a, a))
    return _scale(a, 1.0 / n) if n > 0.0 else a


def cuboid(lo, hi) -> TriMesh:
    """Axis-aligned box, outward-facing (shapes.rs:8-20)."""
    def p(i):
        return [hi[0] if i & 1 else lo[0], hi[1] if i & 2 else lo[1], hi[2] if i & 4 else lo[2]]
    positions = [p(i) for i in range(8)]
    quads = [[0, 2, 3, 1], [4, 5, 7, 6], [0, 1, 5, 4], [2, 6, 7, 3], [0, 4, 6, 2], [1, 3, 7, 5]]
    tris = []
    for q in quads:
        tris.append([q[0], q[1], q[2]])
        tris.append([q[0], q[2], q[3]])
    return TriMesh(positions, tris)


def sphere(center, radius: float, subdiv: int) -> TriMesh:
    """Geodesic sphere (subdivided icosahedron), outward-facing (shapes.rs:47-111)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = [[-1.0, t, 0.0], [1.0, t, 0.0], [-1.0, -t, 0.0], [1.0, -t, 0.0], [0.0, -1.0, t], [0.0, 1.0, t],
         [0.0, -1.0, -t], [0.0, 1.0, -t], [t, 0.0, -1.0], [t, 0.0, 1.0], [-t, 0.0, -1.0], [-t, 0.0, 1.0]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6],
         [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
         [8, 6, 7], [9, 8, 1]]
    v = [_normalize(p) for p in v]
    for _ in range(subdiv):
        mid = {}

        def get(a, b):
            k = (min(a, b), max(a, b))
            r = mid.get(k)
            if r is None:
                v.append(_normalize(_add(v[a], v[b])))
                r = len(v) - 1
                mid[k] = r
            return r

        nf = []
        for a, b, c in f:
            ab = get(a, b)
            bc = get(b, c)
            ca = get(c, a)
            nf.extend([[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]])
        f = nf
    positions = [_add(list(center), _scale(p, radius)) for p in v]
    return TriMesh(positions, f)


def disk(center, d: float, t: float) -> TriMesh:
    """Thin circular disk of diameter d and thickness t, axis along x (shapes.rs:22-45)."""
    n = 96
    m = TriMesh()
    for side in (-0.5, 0.5):
        for k in range(n):
            a = 2.0 * math.pi * k / n
            m.positions.append([center[0] + side * t, center[1] + 0.5 * d * math.cos(a), center[2] + 0.5 * d * math.sin(a)])
        m.positions.append([center[0] + side * t, center[1], center[2]])
    c0, c1 = n, 2 * n + 1
    for k in range(n):
        k1 = (k + 1) % n
        a2, b2 = k + n + 1, k1 + n + 1
        m.triangles.append([c0, k1, k])
        m.triangles.append([c1, a2, b2])
        m.triangles.append([k, k1, b2])
        m.triangles.append([k, b2, a2])
    fix_orientation(m)
    return m


def cylinder(p0, p1, radius: float, segments: int = 48) -> TriMesh:
    """Closed cylinder from p0 to p1 (same construction as `disk`, any axis)."""
    ax = [p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]]
    a = _normalize(ax)
    ref = [1.0, 0.0, 0.0] if abs(a[0]) < 0.9 else [0.0, 1.0, 0.0]
    u = _normalize(_cross(a, ref))
    w = _cross(a, u)
    m = TriMesh()
    for end in (p0, p1):
        for k in range(segments):
            ang = 2.0 * math.pi * k / segments
            c, s = math.cos(ang) * radius, math.sin(ang) * radius
            m.positions.append([end[0] + c * u[0] + s * w[0], end[1] + c * u[1] + s * w[1], end[2] + c * u[2] + s * w[2]])
        m.positions.append(list(end))
    n = segments
    c0, c1 = n, 2 * n + 1
    for k in range(n):
        k1 = (k + 1) % n
        a2, b2 = k + n + 1, k1 + n + 1
        m.triangles.append([c0, k1, k])
        m.triangles.append([c1, a2, b2])
        m.triangles.append([k, k1, b2])
        m.triangles.append([k, b2, a2])
    fix_orientation(m)
    return m


def fix_orientation(m: TriMesh) -> None:
    """Flips connected components whose signed volume is negative (shapes.rs:245-280)."""
    n = len(m.positions)
    parent = list(range(n))

    def find(x):
        r = x
        while parent[r] != r:
            r = parent[r]
        y = x
        while parent[y] != r:
            nx = parent[y]
            parent[y] = r
            y = nx
        return r

    for t in m.triangles:
        a = find(t[0])
        for v in t[1:]:
            b = find(v)
            parent[b] = a
    vol = {}
    roots = [find(t[0]) for t in m.triangles]
    for t, r in zip(m.triangles, roots):
        a, b, c = m.positions[t[0]], m.positions[t[1]], m.positions[t[2]]
        vol[r] = vol.get(r, 0.0) + _dot(a, _cross(b, c)) / 6.0
    for t, r in zip(m.triangles, roots):
        if vol[r] < 0.0:
            t[1], t[2] = t[2], t[1]


def merge(*meshes: TriMesh) -> TriMesh:
    """Concatenates meshes (crates/cfd/tests/channel.rs:13-20)."""
    out = TriMesh()
    for m in meshes:
        off = len(out.positions)
        out.positions.extend([list(p) for p in m.positions])
        out.triangles.extend([[t[0] + off, t[1] + off, t[2] + off] for t in m.triangles])
    return out


def volume(m: TriMesh) -> float:
    """Signed volume of a closed mesh."""
    s = 0.0
    for t in m.triangles:
        s += _dot(m.positions[t[0]], _cross(m.positions[t[1]], m.positions[t[2]])) / 6.0
    return s


@dataclass
class PropSpec:
    """shapes.rs:119-147: a propeller by radius, hub, blade count, geometric
    pitch (advance per revolution), chord at 75 % radius and section
    thickness / chord; same length unit throughout."""
    radius: float
    hub_radius: float
    hub_height: float
    blades: int
    pitch: float
    chord: float
    thick: float

    @staticmethod
    def five_inch() -> "PropSpec":
        """A 5-inch-class prop (5x4.3 pitch) in metres (shapes.rs:133-136)."""
        return PropSpec(0.0635, 0.0075, 0.008, 2, 0.109, 0.0145, 0.10)

    def chord_at(self, r: float) -> float:
        x = min(max(r / self.radius, 0.0), 1.0)
        shape = math.sqrt(max(1.0 - x * x, 0.0)) * (0.55 + 0.9 * x - 0.4 * x * x)
        s75 = math.sqrt(1.0 - 0.5625) * (0.55 + 0.9 * 0.75 - 0.4 * 0.5625)
        return max(self.chord * shape / s75, 0.25 * self.chord)

    def pitch_angle(self, r: float) -> float:
        return math.atan(self.pitch / (2.0 * math.pi * r))

    def mesh(self) -> TriMesh:
        """Hub cylinder plus blades lofted from sections (shapes.rs:149-230)."""
        return _prop_mesh(self.radius, self.hub_radius, self.hub_height, self.blades, self.chord_at,
                          self.pitch_angle, self.thick)


def _section_point(i: int, nc: int, thick: float):
    """shapes.rs:233-246: (chordwise s from the leading edge, thickness
    offset / chord) around a flat-bottomed, rounded-top section."""
    half = nc // 2
    if i <= half:
        s = 0.5 - 0.5 * math.cos(math.pi * i / half)
        return s, thick * 2.6 * math.sqrt(s) * (1.0 - s) * 0.77 + 0.004
    k = i - half
    return 0.5 + 0.5 * math.cos(math.pi * k / half), -0.004


def _prop_mesh(radius, hub_radius, hub_height, blades, chord_at, angle_at, thick, nr: int = 24, nc: int = 16) -> TriMesh:
    m = TriMesh()
    ns = 24
    h0, h1 = -hub_height / 2.0, hub_height / 2.0
    base = 0
    for k in range(ns):
        a = 2.0 * math.pi * k / ns
        c, s = math.cos(a) * hub_radius, math.sin(a) * hub_radius
        m.positions.append([c, s, h0])
        m.positions.append([c, s, h1])
    cb = len(m.positions)
    m.positions.append([0.0, 0.0, h0])
    m.positions.append([0.0, 0.0, h1])
    for k in range(ns):
        a0 = base + 2 * k
        a1 = base + 2 * ((k + 1) % ns)
        m.triangles += [[a0, a1, a1 + 1], [a0, a1 + 1, a0 + 1], [cb, a1, a0], [cb + 1, a0 + 1, a1 + 1]]
    for b in range(blades):
        phase = 2.0 * math.pi * b / blades
        first = len(m.positions)
        r0 = hub_radius * 0.8
        for j in range(nr + 1):
            tt = j / nr
            r = radius * 0.999 if j == nr else r0 + (radius - r0) * tt
            c = chord_at(r)
            beta = angle_at(r)
            for i in range(nc):
                s, t = _section_point(i, nc, thick)
                sc = (0.35 - s) * c
                tc = t * c
                along = sc * math.cos(beta) - tc * math.sin(beta)
                up = sc * math.sin(beta) + tc * math.cos(beta)
                phi = phase + along / r
                m.positions.append([r * math.cos(phi), r * math.sin(phi), up])
        for j in range(nr):
            for i in range(nc):
                a = first + j * nc + i
                b2 = first + j * nc + (i + 1) % nc
                m.triangles += [[a, a + nc, b2 + nc], [a, b2 + nc, b2]]
        for ring, flip in ((0, True), (nr, False)):
            r0i = first + ring * nc
            for i in range(1, nc - 1):
                m.triangles.append([r0i, r0i + i, r0i + i + 1] if flip else [r0i, r0i + i + 1, r0i + i])
    fix_orientation(m)
    return m


def tapered_prop(radius: float = 120.0, chord_root: float = 30.0, chord_tip: float = 12.0, twist_root: float = 35.0,
                 twist_tip: float = 10.0, thick: float = 0.08, blades: int = 2, hub_radius: float = None,
                 hub_height: float = None, seed: int = 20261203) -> TriMesh:
    """A prop in lattice cells about the z axis: chord tapering linearly
    from chord_root at the hub to chord_tip at the tip, twist (section angle
    to the rotation plane) from twist_root to twist_tip degrees, thickness a
    fraction of the chord. `seed` jitters chord (2 %) and twist (1 degree)
    along the span so the blades are not exactly alike."""
    import numpy as np
    rng = np.random.default_rng(seed)
    hub_radius = hub_radius if hub_radius is not None else 0.12 * radius
    hub_height = hub_height if hub_height is not None else 0.9 * chord_root
    jc = rng.uniform(-0.02, 0.02, 4)
    jt = rng.uniform(-1.0, 1.0, 4)

    def frac(r):
        return min(max((r - hub_radius) / (radius - hub_radius), 0.0), 1.0)

    def wobble(j, x):
        return j[0] + j[1] * math.sin(math.pi * x) + j[2] * math.sin(2 * math.pi * x) + j[3] * math.sin(3 * math.pi * x)

    def chord_at(r):
        x = frac(r)
        return (chord_root + (chord_tip - chord_root) * x) * (1.0 + wobble(jc, x) / 4.0)

    def angle_at(r):
        x = frac(r)
        return math.radians(twist_root + (twist_tip - twist_root) * x + wobble(jt, x) / 4.0)

    return _prop_mesh(radius, hub_radius, hub_height, blades, chord_at, angle_at, thick)
