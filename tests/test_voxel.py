"""Voxelizer pins: crates/cfd/tests/geometry.rs:13-76 restated, plus a brute
force inside test against the analytic sphere."""
import math

import numpy as np
import pytest

import simlab_amd as sa
from simlab_amd.shapes import cuboid, merge, sphere, volume

C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0], [-1, -1, 0],
              [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1],
              [0, -1, 1]])


def test_sphere_volume_and_links():
    # geometry.rs:13-41
    r = 10.3
    m = sphere([20.2, 19.7, 20.4], r, 5)
    v = sa.voxelize_static(m.positions, m.triangles, (41, 41, 41))
    exact = 4.0 / 3.0 * math.pi * r ** 3
    assert abs(len(v.solid) / exact - 1.0) < 0.03
    net = np.zeros(3, np.int64)
    for w in v.walls:
        for d in range(1, 19):
            if w["mask"] & (1 << d):
                assert 0.0 < w["q"][d] <= 1.0
                net += C[d]
    assert list(net) == [0, 0, 0]


def test_holes_and_overlaps_do_not_flood():
    # geometry.rs:43-64
    m = cuboid([3.3, 3.3, 3.3], [16.7, 12.6, 14.2])
    m.triangles = m.triangles[2:]
    ins = sa.inside_grid(m.positions, m.triangles, (0.0, 0.0, 0.0), 1.0, (20, 16, 18))
    n = int(ins.sum())
    exact = (16 - 4 + 1) * (12 - 4 + 1) * (14 - 4 + 1)
    assert abs(n - exact) <= exact // 20
    a = cuboid([2.3, 2.3, 2.3], [10.6, 10.6, 10.6])
    b = cuboid([6.4, 6.4, 6.4], [14.7, 14.7, 14.7])
    u = merge(a, b)
    ins = sa.inside_grid(u.positions, u.triangles, (0.0, 0.0, 0.0), 1.0, (17, 17, 17))
    assert ins[(8 * 17 + 8) * 17 + 8]


def test_thin_plate_blocks_links():
    # geometry.rs:66-76
    m = cuboid([2.0, 2.0, 5.4], [12.0, 12.0, 5.6])
    v = sa.voxelize_static(m.positions, m.triangles, (15, 15, 12))
    assert len(v.solid) == 0
    up = int(np.sum((v.walls["mask"] & (1 << 5)) != 0))
    down = int(np.sum((v.walls["mask"] & (1 << 6)) != 0))
    assert up > 50 and up == down


def test_inside_matches_analytic_sphere_away_from_surface():
    # brute force: every grid point farther than 0.3 cells from the surface
    # is classified as the analytic sphere says
    c, r = np.array([7.3, 6.8, 7.1]), 4.3
    m = sphere(list(c), r, 5)
    dims = (15, 14, 15)
    ins = sa.inside_grid(m.positions, m.triangles, (0.0, 0.0, 0.0), 1.0, dims).reshape(dims[2], dims[1], dims[0])
    z, y, x = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
    d = np.sqrt((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2)
    far = np.abs(d - r) > 0.3
    assert np.array_equal(ins[far], (d < r)[far])


def test_halfway_slab_links_have_q_one_half():
    # channel walls at y = 0.5: every link from y=1 into y=0 is cut at q = 1/2
    dims = (8, 10, 6)
    m = cuboid([-5.0, -10.0, -5.0], [13.0, 0.5, 11.0])
    v = sa.voxelize_static(m.positions, m.triangles, dims)
    ys = (v.walls["cell"] // dims[0]) % dims[1]
    assert set(ys.tolist()) == {1}
    for w in v.walls:
        for d in range(1, 19):
            if w["mask"] & (1 << d):
                assert C[d][1] == -1
                assert w["q"][d] == pytest.approx(0.5, abs=1e-6)
    # every fluid cell of the y=1 layer has the 5 links with c_y = -1
    assert len(v.walls) == dims[0] * dims[2]
    assert all(bin(int(w["mask"])).count("1") == 5 for w in v.walls)


def test_sphere_mesh_is_closed_and_outward():
    m = sphere([0.0, 0.0, 0.0], 2.0, 3)
    assert volume(m) > 0
    assert abs(volume(m) / (4 / 3 * math.pi * 8) - 1) < 0.01
