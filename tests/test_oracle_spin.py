"""The oracle's spinning-part path (tri_scatter, voxelize, qclear,
shader.rs:865-1104) pinned on the CPU against the static voxelizer
(voxel.rs:348-490, itself pinned by the reference's voxel tests):

* at angle zero the re-voxelized walls of the five-inch prop match the
  static ones as geometry.rs:79-160 requires (at most 6 % of links differ,
  at most 0.2 % of common links have q more than 0.05 apart);
* turned by 0.4 rad the same holds against the static voxelization of the
  turned mesh, and fails against the mesh turned the other way (a
  transposed rotation would pass at angle zero only).
"""
import numpy as np

import simlab_amd as sa
from simlab_amd import workloads as W
from oracle import pyoracle as O

import spin_links as S


def oracle_records(rot):
    m, setup, geom = S.prop_case()
    ora = O.OracleLattice(setup, geom, store=O.F32)
    ora.run([sa.StepIn(rot, rot, (0.0, 0.0, 0.0), False)])
    return m, ora.records()


def test_oracle_revoxelizer_matches_static_at_angle_zero():
    m, recs = oracle_records(sa.IDENTITY)
    same, diff, qbad, qerr = S.compare(recs, S.static_walls(m))
    print(f"links: {same} agree, {diff} differ, {qbad} with |dq| > 0.05, max dq {qerr:.4f}")
    assert same > 1000
    assert diff <= 0.06 * same
    assert qbad <= 0.002 * same


def test_oracle_revoxelizer_follows_the_rotation():
    th = 0.4
    rot = W.rotation((0.0, 0.0, 1.0), th)
    m, recs = oracle_records(rot)
    same, diff, qbad, qerr = S.compare(recs, S.static_walls(m, rot))
    print(f"turned: {same} agree, {diff} differ, {qbad} with |dq| > 0.05")
    assert same > 1000 and diff <= 0.06 * same and qbad <= 0.002 * same
    wrong = W.rotation((0.0, 0.0, 1.0), -th)
    s2, d2, _, _ = S.compare(recs, S.static_walls(m, wrong))
    print(f"turned the other way: {s2} agree, {d2} differ")
    assert d2 > 0.5 * s2


def test_rotation_matrix_and_spin_driver():
    r = np.asarray(W.rotation((0.0, 0.0, 1.0), 0.3))
    assert np.allclose(r @ r.T, np.eye(3), atol=1e-15)
    assert np.allclose(r @ [1.0, 0.0, 0.0], [np.cos(0.3), np.sin(0.3), 0.0])   # counter-clockwise about +z
    d = W.SpinDriver((0.0, 0.0, 2.0), 0.01, ramp_revs=0.5)
    assert d.ramp == int(np.ceil(np.pi / 0.01))
    steps = d.next(d.ramp + 3)
    f = [s.omega[2] / 0.01 for s in steps]
    assert 0.0 < f[0] < 1e-4 and all(b >= a for a, b in zip(f, f[1:])) and f[-1] == 1.0
    assert np.allclose(steps[1].rot_prev, steps[0].rot)
    assert np.isclose(d.theta, sum(s.omega[2] for s in steps))
