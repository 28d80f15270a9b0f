"""Checkpoint and restore (simlab_cfd_lattice_checkpoint / _restore): a
lattice put back to a checkpoint, or a fresh lattice of the same setup given
it, continues bit for bit as the original did (forces, populations), for
static walls in FP16 and FP32 and for a spinning part; a checkpoint of a
different lattice is refused."""
import sys

import numpy as np
import pytest

import simlab_amd as sa
from simlab_amd import workloads as W

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("storage", [sa.Storage.FP16, sa.Storage.FP32])
def test_restore_continues_bit_for_bit(storage):
    setup, geom, _ = W.sphere_channel((48, 24, 20), 4.0, 14.0, 40.0, 0.05, 0.1, storage, subdiv=3)
    say("case checkpoint static", storage.name)
    lat = sa.Lattice(setup, geom)
    lat.run_still(31)
    ck = lat.checkpoint()
    h1 = lat.run_still(40)
    p1 = lat.get_pops()
    lat.restore(ck)
    assert lat.steps_done == 31
    h2 = lat.run_still(40)
    assert np.array_equal(h1, h2) and np.array_equal(p1, lat.get_pops())
    fresh = sa.Lattice(setup, geom)
    fresh.restore(ck)
    h3 = fresh.run_still(40)
    assert np.array_equal(h1, h3) and np.array_equal(p1, fresh.get_pops())
    other = sa.Lattice(W.sphere_channel((48, 24, 22), 4.0, 14.0, 40.0, 0.05, 0.1, storage, subdiv=3)[0], geom)
    with pytest.raises(sa.SimlabError, match="checkpoint does not fit"):
        other.restore(ck)


def test_restore_spinning_part():
    setup, geom, meta = W.spinning_prop(dims=(48, 48, 40), radius=14.0, tip=0.08, u_axial=0.01, cs=0.16, nu=0.02,
                                        storage=sa.Storage.FP32, center=(23.5, 24.0, 19.5))
    say("case checkpoint spin")
    lat = sa.Lattice(setup, geom)
    drv = W.SpinDriver((0.0, 0.0, 1.0), meta["omega"], ramp_revs=0.0)
    lat.run(drv.next(25))
    ck = lat.checkpoint()
    steps = drv.next(30)
    h1 = lat.run(steps)
    r1 = lat.records()
    lat.restore(ck)
    h2 = lat.run(steps)
    assert np.array_equal(h1, h2) and np.array_equal(r1, lat.records())
