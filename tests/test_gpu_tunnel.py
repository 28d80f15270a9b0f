"""Whole studies through the study driver (simlab_amd.tunnel.run_with), as
crates/cfd/tests/tunnel.rs runs them through `run_with`: the sphere at
Re 100 on 16 cells across against the drag correlations, the test propeller
in hover (thrust along +axis, torque against the rotation, shaft power
positive, for either spin direction) and a slowly spinning sphere against
Stokes' torque on the uniform D3Q19 grid (under 48 cells across, and for
every spinning part, as in the reference), and the spheres of 48 cells
across on the refined D3Q27 grid, as the reference plans them."""
import math

import numpy as np
import pytest

from simlab_amd import tunnel as T
from simlab_amd.shapes import PropSpec

pytestmark = pytest.mark.gpu

# results the reference publishes for these studies (docs/VALIDATION.md)
REF_CD_RE20 = 2.689
REF_SPIN_TORQUE = 1.146e-3   # N m


def _say(msg):
    print(msg, flush=True)


def test_sphere_re100_quick(tmp_path):
    # tunnel.rs tests: sphere_re100_quick = sphere_check(100, 16, 15, 0.05)
    _say("study: sphere Re 100, 16 cells across, 15 passes")
    cd, r = T.sphere_drag(100.0, 16, 15.0)
    s = r.summary
    refc = 0.5 * (T.sphere_cd_cgw(100.0) + T.sphere_cd_sn(100.0))
    _say(f"  Cd {cd:.4f} vs {refc:.4f} ({(cd / refc - 1) * 100:+.2f} %), converged {s.converged}, "
         f"{s.simulated:.1f} passes, {s.mlups:.0f} MLUPS, {s.seconds:.1f} s")
    assert abs(cd / refc - 1.0) < 0.05
    side = math.hypot(s.lift_n, s.side_n) / abs(s.drag_n)
    assert side < 0.05
    # the summary's own Cd uses the rasterized frontal area (within 0.1 % of pi D^2 / 4)
    assert s.cd == pytest.approx(cd, rel=2e-3)
    assert s.headline.startswith("Drag ") and "Cd" in s.headline
    assert s.grid == (137, 113, 113) and s.cells == 137 * 113 * 113
    # the plan's memory estimate (the reference's formula: populations,
    # flags, records, faces and a 64 MiB allowance) is what the lattice holds
    cfg, m = T.sphere_case(100.0, 16)
    est = T.plan(cfg, m, 1e-3, T.SolverOptions()).memory - (64 << 20)
    assert abs(s.memory_bytes / est - 1.0) < 0.1
    # surface pressure: stagnation near +q at the front, finite on most vertices
    cp = r.surface_cp[np.isfinite(r.surface_cp)]
    assert cp.size > 0.9 * r.surface_cp.size
    assert 0.7 < cp.max() < 1.3
    along = next(f for f in r.slices if f.name == "along")
    assert (along.nu, along.nv) == (135, 111)
    sp = along.field("speed").data
    assert np.isfinite(sp).sum() > 0.9 * sp.size
    # far-field air moves at about the wind speed
    assert abs(np.nanmedian(sp) / 10.0 - 1.0) < 0.1
    # result files as the engine writes them; the slice through the centre
    # is NaN inside the sphere (a disk of radius 8 cells: ~200 pixels)
    _, m = T.sphere_case(100.0, 16)
    T.write_result(str(tmp_path), r, m.scaled(1e3))
    meta, d = T.read_slice(str(tmp_path), "along")
    assert (meta["nv"], meta["nu"]) == d["speed"].shape == (111, 135)
    assert 120 < np.isnan(d["speed"]).sum() < 300
    # the slice spans the domain: nu cells of dx (3.125 mm) along the wind
    assert meta["u_mm"][0] == pytest.approx(135 * 3.125) and meta["v_mm"][2] == pytest.approx(111 * 3.125)
    assert np.array_equal(np.fromfile(tmp_path / "field.pressure.f32", "<f4"), r.surface_pressure_pa, equal_nan=True)


def test_sphere_re20():
    # tunnel.rs tests: sphere_re20 = sphere_check(20, 32, 30, 0.05), a uniform
    # grid in the reference too (under 48 cells across)
    _say("study: sphere Re 20, 32 cells across, 30 passes")
    cd, r = T.sphere_drag(20.0, 32, 30.0)
    s = r.summary
    refc = 0.5 * (T.sphere_cd_cgw(20.0) + T.sphere_cd_sn(20.0))
    _say(f"  Cd {cd:.4f} vs {refc:.4f} ({(cd / refc - 1) * 100:+.2f} %), converged {s.converged}, "
         f"{s.simulated:.1f} passes, {s.mlups:.0f} MLUPS, {s.seconds:.1f} s")
    assert abs(cd / refc - 1.0) < 0.05
    assert math.hypot(s.lift_n, s.side_n) / abs(s.drag_n) < 0.05
    # the reference's own result for this study (docs/VALIDATION.md:95):
    # Cd 2.689 on its uniform D3Q19 grid
    assert abs(cd / REF_CD_RE20 - 1.0) < 0.005


def test_refined_sphere_re100_quick():
    # crates/cfd/tests/refined.rs:122-131: 48 cells across, two levels (24
    # on the coarse grid), 12 passes, within 3 % of the correlations
    _say("study: refined sphere Re 100, 48 cells across, 12 passes")
    cd, r = T.sphere_drag(100.0, 48, 12.0)
    s = r.summary
    refc = 0.5 * (T.sphere_cd_cgw(100.0) + T.sphere_cd_sn(100.0))
    _say(f"  Cd {cd:.4f} vs {refc:.4f} ({(cd / refc - 1) * 100:+.2f} %), {s.cells} cells, {s.mlups:.0f} MLUPS, "
         f"{s.seconds:.1f} s")
    assert abs(cd / refc - 1.0) < 0.03
    assert np.isfinite(r.surface_pressure_pa).sum() > 0.9 * r.surface_pressure_pa.size
    sl = r.slices[0]
    assert np.isfinite(sl.fields[0].data).sum() > sl.fields[0].data.size / 2
    cfg, m = T.sphere_case(100.0, 48)
    p = T.plan(cfg, m, 1e-3, T.SolverOptions())
    assert p.multi is not None and s.cells == sum(p.multi.cells) and s.grid == p.dims


@pytest.mark.parametrize("re,d,passes,tol,refine", [(100.0, 48, 30.0, 0.03, True), (300.0, 48, 40.0, 0.06, True),
                                                     (100.0, 48, 30.0, 0.03, False)])
def test_sphere_on_48_cells(re, d, passes, tol, refine):
    # tunnel.rs tests: sphere_re100 and sphere_re300 (sphere_check with the
    # reference's bounds), on the refined D3Q27 grid as the reference plans
    # them, and on one uniform D3Q19 grid (refine = false) at Re 100
    _say(f"study: sphere Re {re:.0f}, {d} cells across, {passes:.0f} passes, refine {refine}")
    cd, r = T.sphere_drag(re, d, passes, T.SolverOptions(refine=refine))
    s = r.summary
    refc = T.sphere_cd_cgw(re) if re > 250.0 else 0.5 * (T.sphere_cd_cgw(re) + T.sphere_cd_sn(re))
    _say(f"  Cd {cd:.4f} vs {refc:.4f} ({(cd / refc - 1) * 100:+.2f} %), converged {s.converged}, "
         f"{s.simulated:.1f} passes, {s.mlups:.0f} MLUPS, {s.seconds:.1f} s")
    assert abs(cd / refc - 1.0) < tol
    assert math.hypot(s.lift_n, s.side_n) / abs(s.drag_n) < 0.05


def _hover(rpm, cells, revs):
    _say(f"study: 5-inch prop hover, {rpm:+.0f} rpm, {cells} cells across, {revs} revolutions")
    r = T.prop_hover(PropSpec.five_inch(), rpm, cells, revs)
    s = r.summary
    _say(f"  thrust {s.thrust_n:.4f} N, torque {s.torque_nm:.5f} N m, power {s.power_w:.2f} W, {s.steps} steps, "
         f"{s.seconds:.1f} s")
    return r


def test_reversed_spin_mirrors_torque():
    # tunnel.rs tests: reversed_spin_mirrors_torque
    a = _hover(20000.0, 24, 1.5).summary
    b = _hover(-20000.0, 24, 1.5).summary
    assert a.torque_nm < 0.0 and b.torque_nm > 0.0
    assert a.power_w > 0.0 and b.power_w > 0.0


def test_spinning_prop_smoke():
    # tunnel.rs tests: spinning_prop_smoke
    r = _hover(20000.0, 32, 2.0)
    s = r.summary
    assert s.thrust_n > 0.0
    assert s.torque_nm < 0.0
    assert s.power_w > 0.0
    assert np.isfinite(r.surface_pressure_pa).sum() > r.surface_pressure_pa.size // 2
    assert s.headline.startswith("Thrust ") and "rpm" in s.headline


def test_rotating_sphere_stokes_torque():
    # tunnel.rs tests: rotating_sphere_stokes_torque (side margin 2 rotor diameters)
    _say("study: sphere spinning at Re 1, 24 cells across, 6 revolutions")
    o = T.SolverOptions()
    o.still_margins = T.Margins(2.5, 2.5, 2.0)
    t, stokes, r = T.rotating_sphere(1.0, 24, 6.0, o)
    _say(f"  torque {-t:.4e} vs 8 pi mu R^3 omega {stokes:.4e} ({(-t / stokes - 1) * 100:+.2f} %), "
         f"thrust {r.summary.thrust_n:.2e}, {r.summary.steps} steps, {r.summary.seconds:.1f} s")
    assert abs(-t / stokes - 1.0) < 0.05
    # the reference's own result (docs/VALIDATION.md:102): 1.146e-3 N m
    assert abs(-t / REF_SPIN_TORQUE - 1.0) < 0.005


def test_prop_in_forward_flight_loses_thrust():
    # a spinning part in a wind (tunnel.rs plans it in the wind's frame): the
    # 5-inch prop at 20000 rpm flying at 10 m/s along +z (advance ratio
    # 0.24 against a pitch/diameter of 0.86) makes less thrust than in hover
    # and still needs driving. On 24 cells across the preview's thrust is
    # small (the reference's own note: previews read thrust 30-90 % off), and
    # in this wind the blades' drag outweighs it, so its sign is not checked.
    hover = _hover(20000.0, 24, 3.0).summary
    _say("study: 5-inch prop, 20000 rpm, 10 m/s axial wind, 24 cells across, 2 passes")
    cfg = T.WindConfig(wind_dir=(0.0, 0.0, -1.0), speed_m_s=10.0, air=T.Air(), spin=T.Spin(20000.0, (0, 0, 1)),
                       quality=24, duration=2.0)
    r = T.run_with(cfg, PropSpec.five_inch().mesh(), 1e-3)
    s = r.summary
    _say(f"  thrust {s.thrust_n:.4f} N (hover {hover.thrust_n:.4f}), torque {s.torque_nm:.5f} N m, "
         f"{s.steps} steps, {s.seconds:.1f} s")
    assert s.cd is None and s.torque_nm < 0.0 and s.power_w > 0.0
    assert s.thrust_n < hover.thrust_n
    # the drag axis is the wind: drag is the force along -z
    assert s.drag_n == pytest.approx(-s.force_n[2])


def test_run_stl_command_line(tmp_path):
    # crates/cfd/examples/run_stl.rs: an STL in mm, a WindConfig, --out files
    import json
    from simlab_amd import run_stl
    from simlab_amd.shapes import sphere
    from simlab_amd.stl import write_binary_stl
    _say("study: run_stl on a 50 mm sphere STL, 16 cells across, 3 passes")
    stl = tmp_path / "ball.stl"
    m = sphere([0.0, 0.0, 0.0], 25.0, 4)
    stl.write_bytes(write_binary_stl(m))
    out = tmp_path / "out"
    cfg = '{"wind_dir":[1,0,0],"speed_m_s":10,"quality":{"type":"normal"},"duration":3}'
    import contextlib
    import io
    buf = io.StringIO()
    with contextlib.redirect_stdout(buf):
        assert run_stl.main([str(stl), cfg, "--cells", "16", "--out", str(out), "--history"]) == 0
    text = buf.getvalue()
    s = json.loads(text[:text.index("\n}") + 2])
    _say(f"  {s['headline']}, {s['steps']} steps")
    assert s["cells_across"] == 16 and s["grid"] == [137, 113, 113] and s["storage"] == "Fp32"
    assert s["cd"] is not None and s["cd"] > 0 and s["thrust_n"] is None
    assert list(s) == list(T._SUMMARY_FIELDS)
    nv = len(m.positions)
    assert np.fromfile(out / "pressure_pa.f32", "<f4").size == nv
    assert np.fromfile(out / "cp.f32", "<f4").size == nv
    js = json.loads((out / "slice.along.json").read_text())
    assert js["nu"] * js["nv"] == len(js["fields"][0]["data"]) == np.fromfile(out / "slice.along.speed.f32", "<f4").size
    hist = (out / "history.csv").read_text().splitlines()
    assert hist[0] == "t,fx,fy,fz,mx,my,mz" and len(hist) > 1
    assert "t=" in text.splitlines()[-1]


def test_validate_command_line(capsys):
    # crates/cfd/examples/validate.rs: `validate sphere 100 16 15`
    from simlab_amd import validate
    _say("study: validate sphere 100 16 15")
    assert validate.main(["sphere", "100", "16", "15"]) == 0
    out = capsys.readouterr().out
    line = next(l for l in out.splitlines() if l.startswith("SPHERE"))
    assert line.startswith("SPHERE Re=100 D=16 cells: Cd=") and "grid=[137, 113, 113]" in line
    err = float(line.split("err=")[1].split("%")[0])
    assert abs(err) < 5.0


def test_sphere_study_in_fp16_storage():
    # the same study with FP16 population storage (as a memory budget would
    # choose it): the drag coefficient within 0.5 % of the FP32 study's
    from simlab_amd import Storage
    _say("study: sphere Re 100, 16 cells across, 15 passes, FP16 and FP32 storage")
    cd16, r16 = T.sphere_drag(100.0, 16, 15.0, T.SolverOptions(storage=Storage.FP16))
    cd32, r32 = T.sphere_drag(100.0, 16, 15.0, T.SolverOptions(storage=Storage.FP32))
    _say(f"  Cd FP16 {cd16:.5f} ({r16.summary.simulated:.1f} passes, {r16.summary.mlups:.0f} MLUPS), "
         f"FP32 {cd32:.5f} ({r32.summary.simulated:.1f} passes), {(cd16 / cd32 - 1) * 100:+.3f} %")
    assert r16.summary.storage == Storage.FP16
    assert any("FP16 storage" in w for w in r16.summary.warnings)
    assert abs(cd16 / cd32 - 1.0) < 0.005


def test_cancel_after_the_first_batch():
    # tunnel.rs:1139-1141: asked after every batch, after the report and the
    # live metrics (drag_n, lift_n in a wind)
    class StopAfterOne(T.Progress):
        def __init__(self):
            self.reports, self.metrics = [], []

        def report(self, fraction, message):
            self.reports.append((fraction, message))

        def metric(self, name, step, value):
            self.metrics.append((name, step, value))

        def cancelled(self):
            return len(self.reports) >= 1

    pr = StopAfterOne()
    cfg, m = T.sphere_case(100.0, 16)
    with pytest.raises(T.Cancelled):
        T.run_with(cfg, m, 1e-3, None, pr)
    assert len(pr.reports) == 1 and pr.reports[0][1].startswith("Solving: ")
    assert [n for n, _, _ in pr.metrics] == ["drag_n", "lift_n"]
    assert pr.metrics[0][2] > 0.0
