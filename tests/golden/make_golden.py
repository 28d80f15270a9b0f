"""Writes the golden vectors in this folder from the CPU oracle (C++
restatement of the reference kernels, single-threaded so every sum has one
order) and the library's CPU voxelizer. The reference itself has no golden
vectors for this path (it is Rust + WGSL on Metal and cannot run here), so
these fixtures pin the restatement against accidental change and give the
GPU tests fixed expected outputs.

    python3 tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import simlab_amd as sa  # noqa: E402
from simlab_amd import shapes, workloads as W  # noqa: E402
from oracle import pyoracle as O  # noqa: E402

SEED = 20261203


def traj_case(storage):
    """12x10x8 sphere channel: velocity inlet, open outlet, halfway slab walls,
    sphere r=2.5 (Bouzidi records), Smagorinsky Cs 0.1."""
    return W.sphere_channel((12, 10, 8), 2.5, 5.0, 20.0, 0.05, 0.1, storage, debug_dump=True, subdiv=2)


def kat_scalars():
    rng = np.random.default_rng(SEED)
    xs = np.concatenate([rng.normal(0, 300, 24), [0.0, 1.0, -1.0, 1e-6, 65000.0, -2049.5, 5e-8, 3e-5]]).astype(np.float32)
    keys = rng.integers(0, 2**32, xs.size, dtype=np.uint64)
    seeds = rng.integers(0, 2**32, xs.size, dtype=np.uint64)
    sr = np.array([O.sround_ref(x, k, s) for x, k, s in zip(xs, keys, seeds)], np.float32)
    f16 = np.array([O.f16_rne(v) for v in sr], np.uint16)
    cells = rng.integers(0, 2**31, 8, dtype=np.uint64)
    cseeds = rng.integers(0, 2**32, 8, dtype=np.uint64)
    dith = np.stack([O.dither16(c, s) for c, s in zip(cells, cseeds)])
    hashes = np.array([O.cell_hash(c, s) for c, s in zip(cells, cseeds)], np.uint32)
    vals = (rng.normal(0, 0.02, 32)).astype(np.float32)
    ds = rng.integers(0, 65536, 32, dtype=np.uint64)
    e16 = np.array([O.enc_f16_sr(v, d) for v, d in zip(vals, ds)], np.uint16)
    eb16 = np.array([O.enc_bf16_sr(v, d) for v, d in zip(vals, ds)], np.uint16)
    hq = np.array([0.5, 0.25, 0.75, 1.0, 0.001, 1e-6, 0.3333333, 0.999], np.float32)
    hb = np.array([O.f16_bits_host(q) for q in hq], np.uint16)
    np.savez_compressed(os.path.join(HERE, "kat_rounding.npz"), sround_x=xs, sround_key=keys.astype(np.uint32),
                        sround_seed=seeds.astype(np.uint32), sround_out=sr, sround_f16=f16, cell=cells.astype(np.uint32),
                        cell_seed=cseeds.astype(np.uint32), cell_hash=hashes, dither16=dith, enc_x=vals,
                        enc_d=ds.astype(np.uint16), enc_f16=e16, enc_bf16=eb16, host_q=hq, host_q_bits=hb)


def kat_aa():
    ld0, st0 = O.aa_maps((4, 3, 5), 0)
    ld1, st1 = O.aa_maps((4, 3, 5), 1)
    np.savez_compressed(os.path.join(HERE, "aa_maps_4x3x5.npz"), load0=ld0.astype(np.uint32), store0=st0.astype(np.uint32),
                        load1=ld1.astype(np.uint32), store1=st1.astype(np.uint32))


def kat_collision():
    rng = np.random.default_rng(SEED + 1)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    cases = []
    for tau0, cs, g in ((0.6, 0.0, (0, 0, 0)), (0.5048, 0.1, (0, 0, 0)), (1.0, 0.0, (0, 0, 0)), (0.53, 0.16, (1e-5, 0, -2e-5)),
                        (0.9, 0.2, (0, 3e-6, 0))):
        fin = rng.normal(0, 0.004, 19) + 0.01 * w * rng.normal(0, 1, 3)[0]
        out = O.collide_f64(fin, tau0, cs, g)
        cases.append((fin, tau0, cs, g, out))
    np.savez_compressed(os.path.join(HERE, "kat_collision.npz"), fin=np.array([c[0] for c in cases]),
                        tau0=np.array([c[1] for c in cases]), cs=np.array([c[2] for c in cases]),
                        grav=np.array([c[3] for c in cases], np.float64), fout=np.array([c[4] for c in cases]))


def kat_voxel():
    m = shapes.sphere((6.2, 5.9, 6.1), 4.3, 3)
    verts = np.asarray(m.positions, np.float64)
    tris = np.asarray(m.triangles, np.uint32)
    v = sa.voxelize_static(verts, tris, (13, 12, 12))
    np.savez_compressed(os.path.join(HERE, "voxel_icosphere_r4p3.npz"), verts=verts, tris=tris, solid=v.solid,
                        wall_cell=v.walls["cell"], wall_mask=v.walls["mask"], wall_q=v.walls["q"])


def trajectories():
    out = {}
    for name, st, ost in (("fp64", sa.Storage.FP64, O.F64), ("fp32", sa.Storage.FP32, O.F32),
                          ("fp16", sa.Storage.FP16, O.F16), ("bf16", sa.Storage.BF16, O.BF16)):
        setup, geom, meta = traj_case(st)
        ora = O.OracleLattice(setup, geom, store=ost, threads=1)
        ft = ora.run_still(10)
        out[f"{name}_pops"] = ora.get_pops()
        out[f"{name}_ft"] = ft
        out[f"{name}_macros"] = ora.macros()
        out[f"{name}_records"] = ora.records()
    out["flags"] = ora.flags()
    np.savez_compressed(os.path.join(HERE, "traj_sphere_12x10x8.npz"), **out)


def main():
    kat_scalars()
    kat_aa()
    kat_collision()
    kat_voxel()
    trajectories()
    meta = {"generator": "tests/golden/make_golden.py", "seed": SEED,
            "files": sorted(f for f in os.listdir(HERE) if f.endswith(".npz"))}
    with open(os.path.join(HERE, "README.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(meta)


if __name__ == "__main__":
    main()
