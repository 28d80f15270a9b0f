"""STL reading and writing (simlab_amd.stl, restating crates/core/src/stl.rs)
and the run_stl command line (crates/cfd/examples/run_stl.rs) up to its
estimate, on the CPU."""
import json
import os
import struct
import subprocess
import sys

import numpy as np
import pytest

from simlab_amd import tunnel as T
from simlab_amd.shapes import TriMesh, cuboid, sphere
from simlab_amd.stl import read_stl, weld, write_binary_stl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _corners(m):
    return np.asarray(m.positions, np.float64)[np.asarray(m.triangles)]


def test_binary_round_trip_keeps_every_triangle():
    m = sphere([0.1, -0.2, 0.3], 1.0, 2)
    r = read_stl(write_binary_stl(m))
    assert len(r.positions) == len(m.positions) and len(r.triangles) == len(m.triangles)
    assert np.array_equal(_corners(r), _corners(m).astype(np.float32))


def test_binary_header_and_normals():
    m = cuboid([0, 0, 0], [1, 2, 3])
    b = write_binary_stl(m)
    assert len(b) == 84 + 50 * len(m.triangles)
    assert struct.unpack_from("<I", b, 80)[0] == len(m.triangles)
    n = np.frombuffer(b, "<f4", count=3, offset=84)
    c = _corners(m)[0]
    want = np.cross(c[1] - c[0], c[2] - c[0])
    assert np.allclose(n, want / np.linalg.norm(want))


def test_binary_that_starts_with_solid_is_read_as_binary():
    m = cuboid([0, 0, 0], [1, 1, 1])
    b = bytearray(write_binary_stl(m))
    b[:5] = b"solid"
    assert len(read_stl(bytes(b)).triangles) == 12


def test_ascii():
    txt = """solid t
  facet normal 0 0 -1
    outer loop
      vertex 0 0 0
      vertex 0 1 0
      vertex 1 0 0
    endloop
  endfacet
  facet normal 0 -1 0
    outer loop
      vertex 0 0 0
      vertex 1 0 0
      vertex 0 0 1
    endloop
  endfacet
endsolid t
"""
    m = read_stl(txt.encode())
    assert len(m.triangles) == 2 and len(m.positions) == 4
    assert m.positions[0] == [0.0, 0.0, 0.0] and m.triangles == [[0, 1, 2], [0, 2, 3]]
    with pytest.raises(ValueError, match="bad vertex line"):
        read_stl(b"solid x\n vertex 1 2\n")
    with pytest.raises(ValueError, match="no triangles"):
        read_stl(b"solid empty\nendsolid empty\n")
    with pytest.raises(ValueError, match="not a valid STL"):
        read_stl(b"\xff\xfe\x00 not text")


def test_weld_merges_close_vertices_and_drops_degenerate_triangles():
    s = 1e-7   # 1e-7 of a unit-size part: within the 1e-6 weld grid
    tris = np.array([
        [[0, 0, 0], [1, 0, 0], [0, 1, 0]],
        [[1 + s, 0, 0], [0, s, 0], [0, 0, 1]],
        [[0, 0, 0], [s, 0, 0], [0, 0, 1]],     # collapses to a segment
    ], np.float64)
    m = weld(tris)
    assert len(m.positions) == 4
    assert m.triangles == [[0, 1, 2], [1, 0, 3]]


def test_config_json_forms():
    # the two configurations of run_stl.rs's usage text
    a = T.config_from_json({"wind_dir": [1, 0, 0], "speed_m_s": 15, "quality": {"type": "normal"}})
    assert a.quality == "normal" and a.spin is None and a.air.density == 1.204 and a.duration is None
    b = T.config_from_json('{"wind_dir":[0,0,1],"speed_m_s":0,"spin":{"rpm":17000,"axis":[0,0,1],'
                           '"center_mm":[8.58,-0.22,0]},"quality":{"type":"custom","cells_across":128}}')
    assert b.quality == 128 and b.spin.rpm == 17000.0 and b.spin.center_mm == (8.58, -0.22, 0.0)
    assert T.config_from_json(T.config_to_json(b)) == b
    # quality defaults to normal; spin needs its centre
    assert T.config_from_json({"wind_dir": [1, 0, 0], "speed_m_s": 3}).quality == "normal"
    with pytest.raises(ValueError, match="center_mm"):
        T.config_from_json({"wind_dir": [1, 0, 0], "speed_m_s": 0, "spin": {"rpm": 1, "axis": [0, 0, 1]}})
    with pytest.raises(ValueError, match="speed_m_s"):
        T.config_from_json({"wind_dir": [1, 0, 0]})


def test_run_stl_estimate(tmp_path):
    stl = tmp_path / "ball.stl"
    stl.write_bytes(write_binary_stl(sphere([0, 0, 0], 25.0, 4)))     # mm
    cfg = '{"wind_dir":[1,0,0],"speed_m_s":10,"quality":{"type":"draft"}}'
    r = subprocess.run([sys.executable, "-m", "simlab_amd.run_stl", str(stl), cfg, "--estimate", "--cells", "16"],
                       cwd=ROOT, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    e = json.loads(r.stdout)
    assert e["cells_across"] == 16 and e["fits"] and e["storage"] == "Fp32"
    # 50 mm across at 16 cells: the sphere_re100_quick grid
    assert e["grid"] == [137, 113, 113] and e["cells"] == 137 * 113 * 113
    assert set(e) == {"fits", "cells", "grid", "cells_across", "memory_bytes", "budget_bytes", "storage", "steps",
                      "seconds", "note", "thin_mm", "cells_thin", "thin_warning"}
    r = subprocess.run([sys.executable, "-m", "simlab_amd.run_stl", str(stl), cfg, "--estimate", "--fp16",
                        "--wind-margins", "2,4,2"], cwd=ROOT, capture_output=True, text=True, timeout=300)
    e = json.loads(r.stdout)
    assert e["storage"] == "Fp16" and e["cells_across"] == 64
    r = subprocess.run([sys.executable, "-m", "simlab_amd.run_stl", str(stl), cfg, "--bogus"], cwd=ROOT,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode != 0 and "unknown option --bogus" in r.stderr


def test_validate_options():
    # crates/cfd/examples/validate.rs option words and key=value pairs
    from simlab_amd import Storage
    from simlab_amd.validate import options
    o = options(["sphere", "100", "32", "20", "fp16", "side=4", "up=3", "down=6", "smag=0.2", "u=0.05", "tol=0.01",
                 "tau=0.9", "radial=2", "axial=3", "voff=0.5", "sigma=0.2", "spin=0.5", "spout=1.5", "spside=0.6",
                 "velsides", "absu", "p0"])
    assert o.storage == Storage.FP16 and o.side_bc == "velocity" and o.absorb_velocity and o.total_pressure_inlet
    assert (o.wind_margins.upstream, o.wind_margins.downstream, o.wind_margins.side) == (3.0, 6.0, 4.0)
    assert (o.still_margins.upstream, o.still_margins.downstream, o.still_margins.side) == (3.0, 3.0, 2.0)
    assert o.smagorinsky == 0.2 and o.max_lattice_speed == 0.05 and o.max_tip_speed == 0.05
    assert o.tolerance == 0.01 and o.max_tau == 0.9 and o.vertex_offset == 0.5 and o.sponge_sigma == 0.2
    assert o.sponge_widths == (0.5, 1.5, 0.6)
    d = options(["rot", "1", "24", "4"])
    assert d.storage is None and d.sponge_widths == (0.4, 1.0, 0.4) and d.wind_margins.side == 3.0
