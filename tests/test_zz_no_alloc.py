"""The allocation-free run (simlab_cfd_lattice_run_noalloc) through the C
ABI against simlab_cfd_lattice_run on twin lattices: the same force history
and populations, bit for bit (the same kernels in the same order; only where
the force partials live differs), on two calls in a row, the second with a
smaller lattice and fewer steps. The first lattice (256x128x128 FP16, about
520 wall blocks: 6.5 k record and 126 k halfway wall cells) holds about 500
steps per chunk in the 16 MiB scratch, fewer than the 512 of the lattice's
own buffers, so its 600 steps also cross a chunk boundary the plain run does
not have."""
import numpy as np
import pytest

import simlab_amd as sa
from simlab_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def twins(setup, geom):
    return sa.Lattice(setup, geom), sa.Lattice(setup, geom)


@pytest.mark.parametrize("case", ["large_then_small"])
def test_run_noalloc_matches_run(case):
    calls = [
        (W.sphere_channel_large(sa.Storage.FP16), 600),
        (W.sphere_channel((48, 24, 20), 4.0, 14.0, 40.0, 0.05, 0.1, sa.Storage.FP32, subdiv=3), 50),
    ]
    for (setup, geom, _), n in calls:
        say(f"case run_noalloc {setup.dims} {setup.storage.name}, {n} steps")
        a, b = twins(setup, geom)
        steps = [sa.StepIn.still()] * n
        ha = a.run(steps)
        hb = b.run_noalloc(steps)
        assert b.steps_done == n
        same_f = np.array_equal(ha, hb)
        same_p = np.array_equal(a.get_pops(), b.get_pops())
        say(f"  force history identical {same_f}, populations identical {same_p}, "
            f"max |dF| {np.abs(ha - hb).max():.3e}")
        assert same_f and same_p
        # a second call on the same lattice continues it as run does
        ha2, hb2 = a.run(steps[:7]), b.run_noalloc(steps[:7])
        assert np.array_equal(ha2, hb2)


def test_run_noalloc_spinning_part():
    # a spinning part (the re-voxelization kernels run too) in FP32
    setup, geom, meta = W.orbit_sphere(4.0, 14.0, 0.04, 10.0, storage=sa.Storage.FP32)
    say(f"case run_noalloc orbit {setup.dims}")
    a, b = twins(setup, geom)
    da, db = W.SpinDriver((0.0, 0.0, 1.0), meta["omega"], ramp_revs=0.0), W.SpinDriver((0.0, 0.0, 1.0), meta["omega"], ramp_revs=0.0)
    ha, hb = a.run(da.next(40)), b.run_noalloc(db.next(40))
    assert np.array_equal(ha, hb) and np.array_equal(a.get_pops(), b.get_pops())
