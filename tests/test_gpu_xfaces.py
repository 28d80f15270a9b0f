"""In 16-bit storage the x-face cells (velocity inlet, open outlet) are
written after the step by k_xfaces from the states their sources left
(tagged with the step), instead of in the step by the source's wave: the
same rule, values, slots, dither and rounding (the face values have no fused
multiply-adds), so the populations and forces are bit-identical to the
in-step writes."""
import numpy as np
import pytest

import simlab_amd as sa
from simlab_amd import workloads as W

pytestmark = pytest.mark.gpu


def run(storage, defer, dims, monkeypatch):
    monkeypatch.setenv("SIMLAB_CFD_XDEFER", "1" if defer else "0")
    setup, geom, meta = W.sphere_channel(dims, 6.0, 30.0, 200.0, 0.05, 0.1, storage, subdiv=3)
    lat = sa.Lattice(setup, geom)
    h = np.concatenate([lat.run_still(17), lat.run_still(1), lat.run_still(22)])
    return h, lat.get_pops()


@pytest.mark.parametrize("storage", [sa.Storage.FP16, sa.Storage.BF16])
@pytest.mark.parametrize("dims", [(96, 48, 40), (64, 48, 40)])
def test_deferred_x_faces_match_the_in_step_writes(storage, dims, monkeypatch):
    # nx = 64: the x = 1 and x = 62 readers share one wave
    print(f"case x faces deferred {storage.name} {dims}", flush=True)
    h1, p1 = run(storage, True, dims, monkeypatch)
    h0, p0 = run(storage, False, dims, monkeypatch)
    print(f"  identical stored values {np.mean(p1 == p0):.6f}", flush=True)
    assert np.array_equal(p1, p0) and np.array_equal(h1, h0)


def test_deferred_x_faces_spinning_part_open_sides(monkeypatch):
    # open x and y faces: the x-face cells at the y and z edges stay with
    # their sources' in-step writes, the others are deferred
    print("case x faces deferred, spinning prop", flush=True)
    out = {}
    for defer in ("1", "0"):
        monkeypatch.setenv("SIMLAB_CFD_XDEFER", defer)
        setup, geom, meta = W.spinning_prop(dims=(64, 56, 48), radius=18.0, storage=sa.Storage.FP16)
        lat = sa.Lattice(setup, geom)
        drv = W.SpinDriver(meta["axis"], meta["omega"])
        out[defer] = (np.concatenate([lat.run(drv.next(11)), lat.run(drv.next(14))]), lat.get_pops())
    assert np.array_equal(out["1"][0], out["0"][0]) and np.array_equal(out["1"][1], out["0"][1])
