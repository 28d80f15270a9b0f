"""The environment switches change how a step is launched, not what it
computes: the spinning part's triangle scatter beside the step or before
it, and the one-kernel step against the split one."""
import numpy as np
import pytest

import simlab_amd as sa
from simlab_amd import workloads as W

pytestmark = pytest.mark.gpu


def _prop_run(monkeypatch, overlap, launch):
    monkeypatch.setenv("SIMLAB_CFD_SPIN_OVERLAP", "1" if overlap else "0")
    monkeypatch.setenv("SIMLAB_CFD_LAUNCH", launch)
    setup, geom, meta = W.spinning_prop(dims=(80, 80, 48), radius=24.0, storage=sa.Storage.FP32)
    lat = sa.Lattice(setup, geom)
    drv = W.SpinDriver(meta["axis"], meta["omega"])
    h = np.concatenate([lat.run(drv.next(13)), lat.run(drv.next(20))])
    return h, lat.get_pops()


@pytest.mark.parametrize("launch", ["fused", "split"])
def test_spin_scatter_overlap_changes_nothing(monkeypatch, launch):
    print(f"case spin overlap on/off, {launch} launch", flush=True)
    h1, p1 = _prop_run(monkeypatch, True, launch)
    h0, p0 = _prop_run(monkeypatch, False, launch)
    # the same kernels on the same data, only the stream each runs on differs
    assert np.array_equal(h1, h0) and np.array_equal(p1, p0)


@pytest.mark.parametrize("storage", [sa.Storage.FP32, sa.Storage.FP16])
def test_split_and_one_kernel_steps_agree(monkeypatch, storage):
    print(f"case split vs one-kernel step {storage.name}", flush=True)
    out = {}
    for mode in ("fused", "split"):
        monkeypatch.setenv("SIMLAB_CFD_LAUNCH", mode)
        monkeypatch.setenv("SIMLAB_CFD_XDEFER", "0")
        setup, geom, meta = W.sphere_channel((96, 48, 40), 6.0, 30.0, 200.0, 0.05, 0.1, storage, subdiv=3)
        lat = sa.Lattice(setup, geom)
        out[mode] = (lat.run_still(30), lat.get_pops())
    (hf, pf), (hs, ps) = out["fused"], out["split"]
    # separately compiled kernels: a multiply-add may round differently
    assert np.allclose(hf, hs, rtol=1e-3, atol=1e-3 * np.abs(hf).max())
    if storage == sa.Storage.FP32:
        assert np.max(np.abs(pf.astype(np.float64) - ps.astype(np.float64))) < 1e-5
    else:
        d = np.abs(pf.view(np.float16).astype(np.float64) - ps.view(np.float16).astype(np.float64)).max() / 32768.0
        assert d < 1e-4


@pytest.mark.parametrize("launch", ["fused", "split"])
def test_plain_groups_and_xcd_order_change_nothing(monkeypatch, launch):
    # skipping the flag loads of plain-fluid groups and taking the tiles in
    # XCD order change which loads run and in which order blocks start, not
    # a single arithmetic operation: bit-identical populations and forces
    print(f"case plain/swizzle on/off, {launch} launch", flush=True)
    monkeypatch.setenv("SIMLAB_CFD_LAUNCH", launch)
    out = {}
    for plain, swz in (("1", "1"), ("0", "0")):
        monkeypatch.setenv("SIMLAB_CFD_PLAIN", plain)
        monkeypatch.setenv("SIMLAB_CFD_SWIZZLE", swz)
        setup, geom, meta = W.sphere_channel((128, 64, 48), 6.0, 30.0, 200.0, 0.05, 0.1, sa.Storage.FP16, subdiv=3)
        lat = sa.Lattice(setup, geom)
        out[plain] = (lat.run_still(25), lat.get_pops())
    assert np.array_equal(out["1"][0], out["0"][0]) and np.array_equal(out["1"][1], out["0"][1])


@pytest.mark.parametrize("storage", [sa.Storage.FP16, sa.Storage.BF16, sa.Storage.FP32])
@pytest.mark.parametrize("launch,sponge", [("fused", False), ("split", False), ("split", True)])
def test_row_aligned_addressing_changes_nothing(monkeypatch, storage, launch, sponge):
    # rows of a multiple of 64 cells: the parity-1 step addresses its slots
    # from three per-lane offsets, one buffer base per z-plane and scalar
    # row shifts (buffer loads and stores); the same slots, values and
    # dither as the generic addressing, so bit-identical populations and
    # forces (the rows wrap at y, z = 0 and the far edge, the periodic sides
    # of the channel)
    print(f"case rows on/off {storage.name} {launch} sponge={sponge}", flush=True)
    monkeypatch.setenv("SIMLAB_CFD_LAUNCH", launch)
    out = {}
    for rows in ("1", "0"):
        monkeypatch.setenv("SIMLAB_CFD_ROWS", rows)
        setup, geom, meta = W.sphere_channel((128, 64, 48), 6.0, 30.0, 200.0, 0.05, 0.1, storage, subdiv=3)
        if sponge:
            setup.sponge = (0, 12, 4, 4, 4, 4)
            setup.sponge_tau, setup.sponge_sigma = 0.9, 0.05
        lat = sa.Lattice(setup, geom)
        out[rows] = (lat.run_still(25), lat.get_pops())
    assert np.array_equal(out["1"][0], out["0"][0]) and np.array_equal(out["1"][1], out["0"][1])


@pytest.mark.parametrize("storage", [sa.Storage.FP16, sa.Storage.FP32])
def test_special_kernel_stream_changes_nothing(monkeypatch, storage):
    # two-stream split step: the special kernel on the side stream (bulk on
    # the main one) or the other way round; the same kernels on the same
    # data, so bit-identical populations and forces
    print(f"case side special on/off {storage.name}", flush=True)
    monkeypatch.setenv("SIMLAB_CFD_LAUNCH", "split")
    monkeypatch.setenv("SIMLAB_CFD_SERIAL", "0")
    out = {}
    for side in ("1", "0"):
        monkeypatch.setenv("SIMLAB_CFD_SIDE_SPECIAL", side)
        setup, geom, meta = W.sphere_channel((128, 64, 48), 6.0, 30.0, 200.0, 0.05, 0.1, storage, subdiv=3)
        lat = sa.Lattice(setup, geom)
        out[side] = (np.concatenate([lat.run_still(7), lat.run_still(19)]), lat.get_pops())
    assert np.array_equal(out["1"][0], out["0"][0]) and np.array_equal(out["1"][1], out["0"][1])
