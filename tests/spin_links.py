"""Shared by the spinning-part tests: the five-inch prop case of the
reference's crates/cfd/tests/geometry.rs:79-160 (gpu_revoxelizer_matches_cpu)
and the link comparison it makes between the re-voxelized records of a
spinning part and the static voxelizer's wall cells."""
import math

import numpy as np

import simlab_amd as sa
from simlab_amd import workloads as W
from simlab_amd.shapes import PropSpec

N = 64
CENTER = (31.0, 32.0, 30.0)


def prop_case(storage=sa.Storage.FP32):
    """geometry.rs:81-111: the five-inch prop in 2.5 mm cells about +z at
    CENTER of a periodic 64^3 box, occupancy at h = 0.125, swept box
    radius + 3 across and 13 cells along the axis."""
    m = PropSpec.five_inch().mesh().scaled(1.0 / 0.0025)
    vox = sa.voxelize_spin(m.positions, m.triangles, (0.0, 0.0, 1.0), 0.125)
    r = int(math.ceil(vox.radius)) + 3
    c = CENTER
    box_lo = (int(c[0]) - r, int(c[1]) - r, int(c[2]) - 6)
    box_n = (2 * r + 1, 2 * r + 1, 13)
    setup = sa.LatticeSetup(dims=(N, N, N), faces=(sa.FaceBc.PERIODIC,) * 6, nu=0.05, storage=storage, center=c)
    return m, setup, sa.Geometry.spin(vox, box_lo, box_n)


def static_walls(mesh, rot=None):
    """The static voxelizer's view of the prop turned by `rot` (lattice =
    CENTER + rot . body)."""
    R = np.eye(3) if rot is None else np.asarray(rot, np.float64)
    p = np.asarray(mesh.positions, np.float64) @ R.T + np.asarray(CENTER)
    return sa.voxelize_static(p, mesh.triangles, (N, N, N))


def half(w):
    return np.asarray([w], np.uint16).view(np.float16).astype(np.float64)[0]


def compare(recs, st):
    """geometry.rs:112-159: links both sides find, links only one finds, and
    common links whose q differs by more than 0.05."""
    spin = {}
    for r in recs:
        if r[0] != 0:
            q = [0.0] * 19
            for k in range(9):
                q[2 * k + 1] = half(int(r[2 + k]) & 0xFFFF)
                q[2 * k + 2] = half(int(r[2 + k]) >> 16)
            spin[int(r[1])] = (int(r[0]), q)
    same = diff = qbad = 0
    qerr = 0.0
    for w in st.walls:
        got = spin.get(int(w["cell"]))
        wm = int(w["mask"])
        if got is None:
            diff += bin(wm).count("1")
            continue
        mask, q = got
        common = mask & wm
        diff += bin(mask ^ wm).count("1")
        same += bin(common).count("1")
        for d in range(1, 19):
            if common & (1 << d):
                e = abs(q[d] - float(w["q"][d]))
                qbad += e > 0.05
                qerr = max(qerr, e)
    return same, diff, qbad, qerr
