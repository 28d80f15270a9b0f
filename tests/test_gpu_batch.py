"""Batched launches and the launch-shape autotuner on the GPU, through the C
ABI: neither may change what a lattice computes."""
import numpy as np
import pytest

import simlab_amd as sa
from simlab_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def cases(storage):
    return [W.sphere_channel((48, 24, 20), 4.0, 14.0, 40.0, 0.05, 0.1, storage, subdiv=3),
            W.sphere_channel((64, 32, 24), 5.0, 20.0, 60.0, 0.05, 0.1, storage, subdiv=3),
            W.drone_arm(dims=(96, 64, 64), storage=storage)]


@pytest.fixture
def same_face_path(monkeypatch):
    # the batch kernels write the x-face cells in the step itself; single
    # lattices in 16-bit storage defer that to k_xfaces (same values up to
    # the rounding of a multiply-add), so the single runs compared here take
    # the batch's path
    monkeypatch.setenv("SIMLAB_CFD_XDEFER", "0")


@pytest.mark.parametrize("storage", [sa.Storage.FP32, sa.Storage.FP16])
def test_batch_equals_separate_runs(storage, same_face_path):
    cs = cases(storage)
    say("case batch", storage.name, [c[0].dims for c in cs])
    solo = [sa.Lattice(s, g) for s, g, _ in cs]
    ft_solo = [np.concatenate([l.run_still(12), l.run_still(18)]) for l in solo]
    members = [sa.Lattice(s, g) for s, g, _ in cs]
    b = sa.Batch(members)
    ft = np.concatenate([b.run_still(12), b.run_still(18)], axis=1)
    for k, (a, m) in enumerate(zip(solo, members)):
        assert m.steps_done == 30
        pa, pm = a.get_pops(), m.get_pops()
        same = np.mean(pa == pm)
        if storage == sa.Storage.FP32:
            dv = np.abs(pa.astype(np.float64) - pm.astype(np.float64)).max()
        else:
            dv = np.abs(pa.view(np.float16).astype(np.float64) - pm.view(np.float16).astype(np.float64)).max() / 32768.0
        say(f"  lattice {k} {cs[k][0].dims}: identical populations {same:.6f}, max diff {dv:.2e}")
        # the batched kernel runs the same per-cell code as the single one,
        # compiled in another function: fused multiply-adds may round a
        # value differently, nothing more
        assert same > (0.99 if storage == sa.Storage.FP32 else 0.9)
        assert dv < (1e-6 if storage == sa.Storage.FP32 else 0.02 * 2.0 ** -9)
        assert np.allclose(ft[k], ft_solo[k], rtol=1e-4, atol=1e-4 * np.abs(ft_solo[k]).max())


def test_batch_rejects_mixed_storage():
    a = sa.Lattice(cases(sa.Storage.FP32)[0][0], cases(sa.Storage.FP32)[0][1])
    s, g, _ = cases(sa.Storage.FP16)[0]
    b = sa.Lattice(s, g)
    with pytest.raises(sa.SimlabError, match="one population storage"):
        sa.Batch([a, b])


def test_autotune_keeps_state_and_results():
    s, g, _ = cases(sa.Storage.FP32)[2]
    say("case autotune", s.dims)
    ref = sa.Lattice(s, g)
    ref.run_still(10)
    lat = sa.Lattice(s, g)
    lat.run_still(10)
    before = lat.get_pops()
    tile, us = lat.autotune(steps=10, use_cache=False)
    say("  tile", tile, "step us", us)
    assert tile[0] % 64 == 0 and tile[0] * tile[1] <= 1024 and us > 0
    assert np.array_equal(lat.get_pops(), before) and lat.steps_done == 10
    h1 = lat.run_still(20)
    h0 = ref.run_still(20)
    pa, pb = lat.get_pops().astype(np.float64), ref.get_pops().astype(np.float64)
    say("  after 20 steps: identical", np.mean(pa == pb), "max diff", np.abs(pa - pb).max())
    # another tile changes which code path writes some face cells (per lane
    # or wave-cooperative), i.e. at most the rounding of fused multiply-adds
    assert np.abs(pa - pb).max() < 1e-6
    assert np.allclose(h1, h0, rtol=1e-4, atol=1e-4 * np.abs(h0).max())
    # the cache answers for a lattice of the same size without timing
    other = sa.Lattice(s, g)
    t2, us2 = other.autotune(use_cache=True)
    assert t2 == tile and us2 == 0.0


@pytest.mark.parametrize("mode", ["split", "fused"])
def test_batch_with_a_large_member(mode, monkeypatch):
    # two-kernel batches (SIMLAB_CFD_BATCH=split): a member of more than 4 Mi
    # cells steps with its own launches inside the batch call, bit-identical
    # to running it alone. The one-kernel batch (the default) steps it in
    # the batched launch: the same per-cell code in another kernel, equal up
    # to the rounding of fused multiply-adds (the single run takes the
    # batch's in-step x-face writes here)
    monkeypatch.setenv("SIMLAB_CFD_BATCH", mode)
    monkeypatch.setenv("SIMLAB_CFD_XDEFER", "0")
    big = W.sphere_channel((192, 160, 144), 12.0, 48.0, 200.0, 0.05, 0.1, sa.Storage.FP16, subdiv=3)
    small = cases(sa.Storage.FP16)[0]
    say("case batch large member", mode, big[0].dims, small[0].dims)
    solo = sa.Lattice(big[0], big[1])
    h_solo = solo.run_still(20)
    ms = [sa.Lattice(big[0], big[1]), sa.Lattice(small[0], small[1])]
    ft = sa.Batch(ms).run_still(20)
    assert ms[0].steps_done == 20 and ms[1].steps_done == 20
    if mode == "split":
        assert np.array_equal(ft[0], h_solo) and np.array_equal(ms[0].get_pops(), solo.get_pops())
    else:
        pa, pm = solo.get_pops(), ms[0].get_pops()
        dv = np.abs(pa.view(np.float16).astype(np.float64) - pm.view(np.float16).astype(np.float64)).max() / 32768
        say(f"  large member: identical {np.mean(pa == pm):.6f}, max diff {dv:.2e}")
        assert np.mean(pa == pm) > 0.9 and dv < 0.02 * 2.0 ** -9
        assert np.allclose(ft[0], h_solo, rtol=1e-4, atol=1e-4 * np.abs(h_solo).max())
    ref = sa.Lattice(small[0], small[1])
    h_ref = ref.run_still(20)
    assert np.allclose(ft[1], h_ref, rtol=1e-4, atol=1e-4 * np.abs(h_ref).max())
