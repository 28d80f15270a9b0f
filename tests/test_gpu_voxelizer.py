"""The GPU static voxelizer (simlab_cfd_voxelize_static_gpu) against the CPU
one (voxel.rs:348-490 restated): the same solid cells, wall cells, link
masks and q values, bit for bit, on the benchmark geometries (the sphere
channels, the drone arm at 512x256x256, the momentum.rs ball box), a
part cut by the domain boundary, thin sheets (the Poiseuille plates, a
disk a fraction of a cell thick: reciprocal links from both sides) and a
static voxelization of the five-inch prop. Times both."""
import time

import numpy as np
import pytest

import simlab_amd as sa
from simlab_amd import workloads as W
from simlab_amd.shapes import PropSpec, cuboid, disk, merge, sphere

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def cases():
    out = []
    nx, ny, nz = 256, 128, 128
    out.append(("sphere channel 256x128x128", merge(sphere([64.0, 63.5, 63.5], 16.0, 5), *W.channel_walls((nx, ny, nz))),
                (nx, ny, nz)))
    out.append(("sphere channel 128x64x64", merge(sphere([32.0, 31.5, 31.5], 8.0, 5), *W.channel_walls((128, 64, 64))),
                (128, 64, 64)))
    out.append(("drone arm 512x256x256", W.drone_arm_mesh(1.0, (0.3 * 512, 128.0, 128.0)), (512, 256, 256)))
    out.append(("ball box 40^3", sphere([20.3, 19.4, 20.17], 6.0, 5), (40, 40, 40)))
    out.append(("sphere cut by the domain", sphere([2.0, 20.0, 18.5], 7.3, 4), (40, 36, 30)))
    out.append(("Poiseuille plates", merge(cuboid([-5.0, -10.0, -5.0], [8.0, 1.37, 8.0]),
                                          cuboid([-5.0, 17.37, -5.0], [8.0, 30.0, 8.0])), (3, 20, 3)))
    out.append(("thin disk", disk([16.2, 15.7, 15.9], 9.0, 0.3), (32, 32, 32)))
    prop = PropSpec.five_inch().mesh()
    pos = np.asarray(prop.positions, np.float64)
    s = 24.0 / (pos[:, 0].max() - pos[:, 0].min())
    prop.positions = [tuple(p) for p in (pos - pos.mean(axis=0)) * s + np.array([30.0, 30.0, 20.0])]
    out.append(("five-inch prop, static", prop, (60, 60, 40)))
    return out


@pytest.mark.parametrize("name,mesh,dims", cases(), ids=[c[0] for c in cases()])
def test_gpu_voxelizer_matches_cpu(name, mesh, dims):
    say("case voxelizer", name, dims)
    t0 = time.perf_counter()
    cpu = sa.voxelize_static(mesh.positions, mesh.triangles, dims)
    t1 = time.perf_counter()
    gpu = sa.voxelize_static(mesh.positions, mesh.triangles, dims, gpu=True)
    t2 = time.perf_counter()
    say(f"  {len(cpu.solid)} solid, {len(cpu.walls)} wall cells; CPU {t1 - t0:.3f} s, GPU {t2 - t1:.3f} s")
    assert cpu.lo == gpu.lo and cpu.hi == gpu.hi
    assert np.array_equal(cpu.solid, gpu.solid)
    assert len(cpu.walls) == len(gpu.walls) and len(cpu.walls) > 0
    assert np.array_equal(cpu.walls["cell"], gpu.walls["cell"])
    assert np.array_equal(cpu.walls["mask"], gpu.walls["mask"])
    assert np.array_equal(cpu.walls["q"].view(np.uint32), gpu.walls["q"].view(np.uint32))
