"""The C-ABI library loads, exports every function include/*.h declares, and
rejects invalid setups with the reference's messages (lattice.rs:233-257)
before any device work, so this runs without a GPU."""
import ctypes as C
import glob
import os
import re
import subprocess

import pytest

import simlab_amd as sa

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w \*]*?\b(simlab_cfd_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 20
    out = subprocess.run(["nm", "-D", "--defined-only", sa.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = sa.load_library()
    for n in names:
        getattr(lib, n)


@pytest.mark.parametrize("dims,faces,msg", [
    ((2, 8, 8), (2,) * 6, "grid must be at least 3 cells in every direction"),
    ((8, 8, 8), (2, 1, 2, 2, 2, 2), "periodic faces must come in pairs"),
    ((0, 8, 8), (2,) * 6, "grid of 0 cells is out of range"),
])
def test_setup_errors_carry_the_reference_messages(dims, faces, msg):
    setup = sa.LatticeSetup(dims=dims, faces=[sa.FaceBc(f) for f in faces])
    with pytest.raises(sa.SimlabError, match=re.escape(msg)):
        sa.Lattice(setup)


def test_struct_sizes_match_the_header():
    # the ctypes mirror of the C structs has the header's layout
    assert C.sizeof(sa._WallCell) == 116
    assert C.sizeof(sa._StepIn) == 9 * 8 * 2 + 3 * 8 + 8
    assert C.sizeof(sa._Setup) % 8 == 0


def test_max_cells_is_positive():
    assert sa.max_cells(sa.Storage.FP16) >= 1 << 30


def test_refined_setup_errors_carry_the_reference_messages():
    """mlattice.rs:192-209 and 713-717, checked before any device work."""
    from simlab_amd import mlattice as ml
    F = sa.FaceBc
    s = ml.MultiSetup(dims0=(32, 24, 24), faces=(F.PERIODIC,) * 6, boxes=[ml.LevelBox((8, 6, 6), (16, 12, 12))])
    with pytest.raises(sa.SimlabError, match="periodic faces only work without refinement"):
        ml.MultiLattice(s)
    s.faces = (F.VELOCITY,) + (F.OPEN,) * 5
    s.boxes = [ml.LevelBox((1, 6, 6), (16, 12, 12))]
    with pytest.raises(sa.SimlabError, match="refinement box 1 does not fit inside its parent"):
        ml.MultiLattice(s)
    s.boxes = [ml.LevelBox((8, 6, 6), (16, 12, 7))]
    with pytest.raises(sa.SimlabError, match="refinement box 1 does not fit inside its parent"):
        ml.MultiLattice(s)
    s.boxes = []
    s.storage = sa.Storage.BF16
    with pytest.raises(sa.SimlabError, match="FP32, FP16 or FP64"):
        ml.MultiLattice(s)


def test_refined_struct_layout():
    from simlab_amd import mlattice as ml
    assert C.sizeof(ml._LevelBox) == 24
    assert C.sizeof(ml._MultiSetup) == 16 + 7 * 24 + 24 + 8 * 2 + 24 * 3 + 24 + 16 + 16 + 24
