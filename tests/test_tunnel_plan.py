"""Planning of a study (simlab_amd.tunnel, restating crates/cfd/src/tunnel.rs
`plan`, `estimate`, `finest_fit`, `sig`) on the CPU: crates/cfd/tests/tunnel.rs
`estimate_and_budget`, `headline_numbers_use_significant_figures` and the
parts of `slender_parts_get_cells_across_their_thickness` that hold on a
uniform grid, plus hand-derived grids of parts with dyadic sizes (every
length an exact binary fraction, so the cell counts are exact)."""
import math

import numpy as np
import pytest

from simlab_amd import FaceBc, Storage
from simlab_amd import tunnel as T
from simlab_amd.shapes import PropSpec, cuboid, sphere


def test_sig_matches_the_reference_strings():
    # tunnel.rs tests: headline_numbers_use_significant_figures
    assert T.sig(0.000123456, 3) == "0.000123"
    assert T.sig(1234.5, 3) == "1230"
    assert T.sig(2.0, 3) == "2.00"
    assert T.sig(-0.0456, 2) == "-0.046"
    # halves round away from zero (f64::round), not to even
    assert T.sig(0.125, 2) == "0.13"
    assert T.sig(-2.5, 1) == "-3"
    assert T.sig(0.0, 3) == "0"


def test_cells_for_thin():
    assert T.cells_for_thin(64, 16, 0.15, 0.012) == 200
    assert T.cells_for_thin(64, 16, 0.02, 0.02) == 64
    assert T.cells_for_thin(64, 16, 1.0, 0.001) == 64 * T.THIN_MAX_BOOST
    assert T.cells_for_thin(64, 16, 1.0, 0.0) == 64


def test_quality_tables():
    assert [T.cells_across(q) for q in ("draft", "normal", "fine")] == [64, 128, 192]
    assert T.cells_across(3) == 8 and T.cells_across(9000) == 4096
    assert T.default_duration("draft", False) == 20.0 and T.default_duration(48, False) == 40.0
    assert T.default_duration("fine", True) == 16.0 and T.default_duration(48, True) == 12.0


def test_frame_is_orthonormal_and_right_handed():
    for d in [(1, 0, 0), (0, 0, 1), (0.3, -0.4, 0.2), (0, 0, -1), (0.1, 0.0, 0.99)]:
        r = T.frame(d)
        assert np.allclose(r @ r.T, np.eye(3), atol=1e-12)
        assert np.linalg.det(r) == pytest.approx(1.0)
        assert np.allclose(r[0], np.asarray(d) / np.linalg.norm(d))
    # the wind along +x: lattice axes are the model axes
    assert np.allclose(T.frame((1, 0, 0)), np.eye(3))


def test_wind_plan_of_a_dyadic_box():
    # 125 x 62.5 x 31.25 mm box, 16 cells across, 10 m/s along +x:
    # dx 7.8125 mm; frontal size 62.5 mm, thinner size 31.25 mm, so the
    # sides sit max(3 * min(62.5, 2 * 31.25), 1.5 * 62.5) = 187.5 mm out
    m = cuboid([-0.0625, -0.03125, -0.015625], [0.0625, 0.03125, 0.015625])
    p = T.plan(T.WindConfig(speed_m_s=10.0, quality=16), m, 1e-3, T.SolverOptions())
    assert p.dx == 0.0078125
    # x: 2.5 + 5 frontal sizes past the 125 mm -> 76 cells + 1; y: 437.5 mm
    # -> 56 + 1; z: 406.25 mm -> 52 + 1
    assert p.dims == (77, 57, 53)
    assert p.center == (28.0, 28.0, 26.0)
    # Mach limit: u_lat = 0.075
    assert p.u_lat == pytest.approx(0.075) and p.dt == pytest.approx(0.075 * 0.0078125 / 10.0)
    assert p.nu_lat == pytest.approx(1.516e-5 * p.dt / p.dx ** 2)
    assert p.steps_unit == pytest.approx(16 / 0.075)
    # sponges: 0.4 frontal sizes (3.2 cells, at least 6) inlet and sides, 1.0 (8 cells) outlet
    assert p.sponge == (6, 8, 6, 6, 6, 6)
    assert p.faces == (FaceBc.VELOCITY,) + (FaceBc.OPEN,) * 5
    # the rasterized frontal area of a box face is exact
    assert p.frontal_area == 0.0625 * 0.03125
    assert p.blockage == pytest.approx(0.0625 * 0.03125 / (0.4375 * 0.40625))
    assert p.storage == Storage.FP32 and not p.spin


def test_velocity_side_faces_and_a_viscosity_limited_step():
    m = cuboid([-0.0625, -0.03125, -0.015625], [0.0625, 0.03125, 0.015625])
    cfg = T.WindConfig(speed_m_s=10.0, quality=16)
    cfg.air = T.Air(density=1.2, kinematic_viscosity=0.5)
    p = T.plan(cfg, m, 1e-3, T.SolverOptions(side_bc="velocity"))
    assert p.faces == (FaceBc.VELOCITY, FaceBc.OPEN) + (FaceBc.VELOCITY,) * 4
    # nu * dt / dx^2 capped at (max_tau - 0.5) / 3
    assert p.nu_lat == pytest.approx(0.5 / 3.0)
    assert p.u_lat < 0.075


def test_spinning_plan_in_still_air():
    spec = PropSpec.five_inch()
    cfg = T.WindConfig(wind_dir=(0, 0, 1), speed_m_s=0.0, spin=T.Spin(20000.0, (0, 0, 1)), quality=24)
    p = T.plan(cfg, spec.mesh(), 1e-3, T.SolverOptions())
    assert p.spin and p.still
    # the lattice x axis is the spin axis
    assert np.allclose(p.axis_lat, (1, 0, 0))
    assert p.rotor_radius == pytest.approx(0.0635 * 0.999, rel=1e-3)
    # the tip moves max_tip_speed cells per step
    assert np.linalg.norm(p.omega_lat) * p.rotor_radius / p.dx == pytest.approx(0.1)
    # one revolution in steps
    assert p.steps_unit == pytest.approx(2 * math.pi / np.linalg.norm(p.omega_lat))
    assert p.faces == (FaceBc.OPEN,) * 6 and len(set(p.sponge)) == 1 and p.sponge[0] >= 8
    # the swept box holds the rotor plus 3 cells, inside the domain
    rr = p.rotor_radius / p.dx
    for k in (1, 2):
        assert p.box_n[k] >= 2 * rr + 6 - 1
        assert p.box_lo[k] >= 1 and p.box_lo[k] + p.box_n[k] <= p.dims[k] - 1
    assert p.frontal_area == pytest.approx(math.pi * p.rotor_radius ** 2)
    # the reverse spin plans the same grid with omega negated
    q = T.plan(T.WindConfig(wind_dir=(0, 0, 1), speed_m_s=0.0, spin=T.Spin(-20000.0, (0, 0, 1)), quality=24),
               spec.mesh(), 1e-3, T.SolverOptions())
    assert q.dims == p.dims and np.allclose(q.omega_lat, -p.omega_lat)


def test_plan_rejects_what_the_reference_rejects():
    m = cuboid([0, 0, 0], [0.01, 0.01, 0.01])
    with pytest.raises(ValueError, match="still air"):
        T.plan(T.WindConfig(speed_m_s=0.0), m, 1e-3, T.SolverOptions())
    with pytest.raises(ValueError, match="zero or positive"):
        T.plan(T.WindConfig(speed_m_s=-1.0), m, 1e-3, T.SolverOptions())
    with pytest.raises(ValueError, match="direction"):
        T.plan(T.WindConfig(wind_dir=(0, 0, 0)), m, 1e-3, T.SolverOptions())
    with pytest.raises(ValueError, match="no surface"):
        T.plan(T.WindConfig(), T.TriMesh(), 1e-3, T.SolverOptions())


def test_estimate_and_budget():
    # tunnel.rs tests: estimate_and_budget
    cfg, m = T.sphere_case(1000.0, 32)
    e = T.estimate(cfg, m, 1e-3, 8_000_000_000)
    assert e.fits and e.cells > 100_000 and e.seconds > 0.0
    c8 = T.finest_fit(cfg, m, 1e-3, 8_000_000_000)
    c32 = T.finest_fit(cfg, m, 1e-3, 32_000_000_000)
    assert c32 > c8
    # the finest fit fits, one more cell across does not
    o = T.SolverOptions(cells_across=c8)
    assert T.plan(cfg, m, 1e-3, o, 8_000_000_000).memory <= 8_000_000_000
    o = T.SolverOptions(cells_across=c8 + 1)
    assert T.plan(cfg, m, 1e-3, o, 8_000_000_000).memory > 8_000_000_000
    spin = T.WindConfig(wind_dir=(0, 0, 1), speed_m_s=0.0, spin=T.Spin(20000.0, (0, 0, 1)), quality="normal")
    e = T.estimate(spin, PropSpec.five_inch().mesh(), 1e-3, 32_000_000_000)
    assert e.steps > 1000 and "revolutions" in e.note


def test_memory_budget_selects_fp16():
    cfg, m = T.sphere_case(1000.0, 64)
    big = T.plan(cfg, m, 1e-3, T.SolverOptions(), None)
    assert big.storage == Storage.FP32
    tight = T.plan(cfg, m, 1e-3, T.SolverOptions(), int(big.memory * 0.8))
    assert tight.storage == Storage.FP16 and tight.memory < big.memory


def test_slender_parts_get_cells_across_their_thickness():
    # tunnel.rs tests: slender_parts_get_cells_across_their_thickness (tests/tunnel.rs:146-174)
    arm = cuboid([-0.003, -0.075, -0.006], [0.003, 0.075, 0.006])
    for q in ("draft", "normal"):
        e = T.estimate(T.WindConfig(speed_m_s=15.0, quality=q), arm, 1e-3, 32_000_000_000)
        assert abs(e.thin_mm - 12.0) < 0.01
        assert e.cells_thin >= T.thin_cells_target(q) - 0.5, (q, e.cells_thin)
        assert e.thin_warning is None
        assert "grid levels" in e.note
    # a compact part keeps the quality's grid
    cfg = T.WindConfig(speed_m_s=15.0, quality="draft")
    ball = sphere([0.0, 0.0, 0.0], 0.01, 4)
    assert T.estimate(cfg, ball, 1e-3, 32_000_000_000).cells_across == 64
    # a small memory budget caps the refinement, and the estimate says so
    e = T.estimate(T.WindConfig(speed_m_s=15.0, quality="normal"), arm, 1e-3, 600_000_000)
    assert e.cells_thin < 24.0 and e.thin_warning is not None and e.memory_bytes <= 600_000_000
    # a custom grid is kept as asked, with the warning
    e = T.estimate(T.WindConfig(speed_m_s=15.0, quality=64), arm, 1e-3, 32_000_000_000)
    assert e.cells_across == 64 and e.thin_warning is not None
    assert T.cells_for_thin(64, 16, 0.15, 0.012) == 200
    assert T.cells_for_thin(64, 16, 0.02, 0.02) == 64
    assert T.cells_for_thin(64, 16, 1.0, 0.001) == 64 * T.THIN_MAX_BOOST


def test_slender_part_on_one_uniform_grid():
    # refine = false: one uniform D3Q19 grid, where the time cap holds the
    # arm near the quality's grid and the estimate says its thickness is
    # under-resolved
    opt = T.SolverOptions(refine=False)
    arm = cuboid([-0.003, -0.075, -0.006], [0.003, 0.075, 0.006])
    for q, target, want in (("draft", 16, 200), ("normal", 24, 300)):
        e = T.estimate(T.WindConfig(speed_m_s=15.0, quality=q), arm, 1e-3, 32_000_000_000, opt)
        assert abs(e.thin_mm - 12.0) < 0.01
        assert e.cells_across >= T.cells_across(q) and "grid levels" not in e.note
        if e.cells_thin < target - 0.5:
            assert e.thin_warning is not None and f"{want} cells" in e.thin_warning


def test_refined_plan_nests_boxes_round_the_part():
    """tunnel.rs:467-475 and 643-687: levels from the cells across (the
    coarsest keeps 24 or more), coarse dims that make the finest spacing
    dx, each box inside its parent with the solver's margins, cell updates
    per finest step summed over levels."""
    assert [T.still_levels(n) for n in (16, 47, 48, 95, 96, 192, 400)] == [1, 1, 2, 2, 3, 4, 4]
    cfg, m = T.sphere_case(100.0, 48)
    p = T.plan(cfg, m, 1e-3, T.SolverOptions())
    mp = p.multi
    assert mp is not None and len(mp.cells) == 2
    assert all(p.dims[k] == (mp.dims0[k] - 1) * 2 + 1 for k in range(3))
    dims = [tuple(mp.dims0)]
    for l, b in enumerate(mp.boxes, start=1):
        margin = 2 if l == 1 else 3
        for k in range(3):
            assert b.lo[k] >= margin and b.lo[k] + b.n[k] - 1 <= dims[-1][k] - 1 - margin and b.n[k] >= 8
        dims.append(tuple(2 * (b.n[k] - 1) + 1 for k in range(3)))
    assert mp.cells == [int(np.prod(d)) for d in dims]
    assert p.updates_per_step == pytest.approx(sum(c * 2 ** l for l, c in enumerate(mp.cells)) / 2)
    # the part lies inside the finest box: its centre, in finest cells from org_f
    c = [p.center[k] - mp.org_f[k] for k in range(3)]
    r = 0.025 / p.dx
    assert all(r + 2 < c[k] < dims[-1][k] - r - 2 for k in range(3))
    # refine = false plans the uniform grid
    assert T.plan(cfg, m, 1e-3, T.SolverOptions(refine=False)).multi is None
    # spinning parts never refine
    spin = T.WindConfig(wind_dir=(0, 0, 1), speed_m_s=0.0, spin=T.Spin(20000.0, (0, 0, 1)), quality=64)
    assert T.plan(spin, PropSpec.five_inch().mesh(), 1e-3, T.SolverOptions()).multi is None


def test_result_files_round_trip(tmp_path):
    # engine wind.rs:465-497 file set, written from a synthetic result
    nu, nv = 5, 3
    data = np.arange(nu * nv, dtype=np.float32)
    data[4] = np.nan
    sl = T.FlowSlice("along", "Along the wind", np.array([1.0, 2.0, 3.0]), np.array([50.0, 0, 0]),
                     np.array([0, 0, 30.0]), nu, nv,
                     [T.SliceField("speed", "Air speed", "m/s", 0.0, 14.0, data),
                      T.SliceField("../bad", "x", "", 0.0, 0.0, data)])
    pres = np.array([-3.0, 1.0, np.nan, 2.0], np.float32)
    r = T.WindResult(None, pres, pres / 2.0, [sl], [])
    m = cuboid([0, 0, 0], [1, 2, 3])
    fields, slices = T.write_result(str(tmp_path), r, m)
    assert [f["name"] for f in fields] == ["pressure", "cp"]
    # pressure range symmetric about zero (white is ambient), data range kept
    assert fields[0]["min"] == -3.0 and fields[0]["max"] == 3.0
    assert fields[0]["data_min"] == -3.0 and fields[0]["data_max"] == 2.0
    got = np.fromfile(tmp_path / "field.pressure.f32", "<f4")
    assert np.array_equal(got, pres, equal_nan=True)
    assert slices == [{"name": "along", "label": "Along the wind"}]
    meta, d = T.read_slice(str(tmp_path), "along")
    assert meta["nu"] == nu and meta["nv"] == nv and meta["origin_mm"] == [1.0, 2.0, 3.0]
    # unsafe field names are skipped
    assert [f["name"] for f in meta["fields"]] == ["speed"]
    assert meta["fields"][0]["min"] == 0.0 and meta["fields"][0]["max"] == 14.0
    assert np.array_equal(d["speed"].reshape(-1), data, equal_nan=True)
    assert not (tmp_path / "slice.along.../bad.f32").exists()
    # display mesh: positions, unit normals, indices, one patch per triangle
    pos = np.fromfile(tmp_path / "mesh.positions.f32", "<f4").reshape(-1, 3)
    nrm = np.fromfile(tmp_path / "mesh.normals.f32", "<f4").reshape(-1, 3)
    idx = np.fromfile(tmp_path / "mesh.indices.u32", "<u4").reshape(-1, 3)
    patch = np.fromfile(tmp_path / "mesh.patch.u32", "<u4")
    assert np.array_equal(pos, np.asarray(m.positions, np.float32))
    assert np.allclose(np.linalg.norm(nrm, axis=1), 1.0, atol=1e-6)
    assert np.array_equal(idx, np.asarray(m.triangles)) and patch.size == len(m.triangles)
    # outward normals of a closed box point away from its centre
    assert np.all(((pos - pos.mean(axis=0)) * nrm).sum(axis=1) > 0)


def test_safe_names():
    assert T.safe_name("speed") and T.safe_name("slice_1.v-2")
    assert not T.safe_name("") and not T.safe_name("../x") and not T.safe_name("a b") and not T.safe_name("é")


def test_layout_limit_is_the_lattices():
    import simlab_amd as sa
    for st in (Storage.FP32, Storage.FP16):
        assert T.max_cells_layout(st) == sa.max_cells(st)
    # a sphere at 128 cells across (876 M cells) keeps FP32 without a budget
    cfg, m = T.sphere_case(1000.0, 128)
    p = T.plan(cfg, m, 1e-3, T.SolverOptions())
    assert int(np.prod(p.dims)) > 537_000_000 and p.storage == Storage.FP32


def test_a_cancelled_run_stops_before_the_lattice():
    # tunnel.rs:947-949: cancellation is asked right after the voxelization,
    # before any GPU work; the error reads "stopped" (core progress.rs:44-50)
    class Stop(T.Progress):
        asked = 0

        def cancelled(self):
            Stop.asked += 1
            return True

    cfg, m = T.sphere_case(100.0, 8)
    with pytest.raises(T.Cancelled, match="stopped"):
        T.run_with(cfg, m, 1e-3, None, Stop())
    assert Stop.asked == 1
