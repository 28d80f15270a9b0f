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


CV = np.array([(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
               (1, 0, 1), (-1, 0, -1), (0, 1, 1), (0, -1, -1), (1, -1, 0), (-1, 1, 0), (1, 0, -1), (-1, 0, 1),
               (0, 1, -1), (0, -1, 1)], np.float64)
WQ = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


def test_refill_and_momentum_of_uncovered_and_covered_cells():
    # shader.rs:1040-1100 on a plate turned by 0.35 rad in one step through
    # fluid in uniform motion u0 (records at equilibrium, rho = 1): a cell
    # the plate uncovers gets feq(rho = 1, u_w = omega x r) written as its
    # previous outputs and books F = -rho u_w; a cell it covers books the
    # momentum of its fluid, F = rho u0; both with T = r x F
    from simlab_amd.shapes import cuboid
    n, u0, om, th = 32, 0.02, 0.01, 0.35
    c = (15.5, 16.0, 15.5)
    m = cuboid([-6.0, -1.5, -1.5], [6.0, 1.5, 1.5])
    vox = sa.voxelize_spin(m.positions, m.triangles, (0.0, 0.0, 1.0), 0.125)
    geom = sa.Geometry.spin(vox, (5, 5, 11), (22, 22, 10))
    setup = sa.LatticeSetup(dims=(n, n, n), faces=(sa.FaceBc.PERIODIC,) * 6, nu=0.05, u_inf=(u0, 0.0, 0.0),
                            u_init=(u0, 0.0, 0.0), storage=sa.Storage.FP64, center=c)
    ora = O.OracleLattice(setup, geom, store=O.F64)
    ora.spin_stage(sa.StepIn(sa.IDENTITY, sa.IDENTITY, (0.0, 0.0, 0.0), False))
    s0 = set(np.nonzero((ora.flags() & 3) == 1)[0].tolist())
    rot = W.rotation((0.0, 0.0, 1.0), th)
    ora.spin_stage(sa.StepIn(rot, sa.IDENTITY, (0.0, 0.0, om), False))
    s1 = set(np.nonzero((ora.flags() & 3) == 1)[0].tolist())
    unc, cov = sorted(s0 - s1), sorted(s1 - s0)
    print(f"solid {len(s0)} -> {len(s1)}: {len(unc)} uncovered, {len(cov)} covered")
    assert len(unc) > 10 and len(cov) > 10
    recs = ora.records()
    by_cell = {int(r[1]): r[33:39].view(np.float32).astype(np.float64) for r in recs}
    pops = ora.get_pops().astype(np.float64)
    _, st = O.aa_maps(setup.dims, 1)   # the refill goes where a parity-1 step would have put its outputs

    def pc_of(cell):
        x, y, z = cell % n, (cell // n) % n, cell // (n * n)
        return np.array([np.float32(x - np.float32(c[0])), np.float32(y - np.float32(c[1])), np.float32(z - np.float32(c[2]))],
                        np.float64)

    for cell in cov:
        ft, pc = by_cell[cell], pc_of(cell)
        F = np.array([u0, 0.0, 0.0])
        assert np.allclose(ft[:3], F, atol=1e-7) and np.allclose(ft[3:], np.cross(pc, F), atol=1e-6)
    for cell in unc:
        pc = pc_of(cell)
        uw = np.array([np.float32(-np.float32(om) * np.float32(pc[1])), np.float32(np.float32(om) * np.float32(pc[0])), 0.0],
                      np.float64)
        ft = by_cell[cell]
        assert np.allclose(ft[:3], -uw, atol=1e-7) and np.allclose(ft[3:], np.cross(pc, -uw), atol=1e-6)
        cu = CV @ uw
        feq = WQ * (3.0 * cu + 4.5 * cu * cu - 1.5 * uw @ uw)
        got = pops[st[cell].astype(np.int64)]
        assert np.max(np.abs(got - feq)) < 1e-12, (cell, got - feq)


def test_refill_mean_stress_is_the_neighbours_traceless_stress():
    # shader.rs:1062-1082: a refilled cell takes the mean non-equilibrium
    # stress of its fluid neighbours' records. A record holding
    # feq(rho, u) + fneq(Pi) (the second moment of the D3Q19 equilibrium is
    # (rho - 1)/3 I + rho u u in deviations, that of fneq(Pi) is Pi) gives back
    # Pi less a third of its trace on the diagonal, whatever rho and u; the
    # rest population, which records do not keep, does not enter.
    rng = np.random.default_rng(7)
    n = 5
    fpost, mac, want = [], [], np.zeros(6)
    for _ in range(n):
        rho = 1.0 + 0.02 * rng.standard_normal()
        u = 0.05 * rng.standard_normal(3)
        P = 1e-3 * rng.standard_normal((3, 3))
        P = P + P.T
        cu = CV @ u
        f = WQ * ((rho - 1.0) + rho * (3.0 * cu + 4.5 * cu * cu - 1.5 * u @ u))
        f = f + 4.5 * WQ * (np.einsum("ia,ib,ab->i", CV, CV, P) - np.trace(P) / 3.0)
        fpost.append(f[1:])
        mac.append([rho, *u])
        D = P - np.trace(P) / 3.0 * np.eye(3)
        want += np.array([D[0, 0], D[1, 1], D[2, 2], D[0, 1], D[0, 2], D[1, 2]]) / n
    got = O.refill_stress_f64(np.array(fpost), np.array(mac))
    assert np.allclose(got, want, rtol=1e-10, atol=1e-15)
    # one neighbour at equilibrium: no stress
    rho, u = 1.01, np.array([0.03, -0.01, 0.02])
    cu = CV @ u
    f = WQ * ((rho - 1.0) + rho * (3.0 * cu + 4.5 * cu * cu - 1.5 * u @ u))
    assert np.allclose(O.refill_stress_f64(f[None, 1:], np.array([[rho, *u]])), 0.0, atol=1e-17)
