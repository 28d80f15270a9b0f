"""The committed golden vectors (tests/golden, written by make_golden.py from
the oracle and the CPU voxelizer) against the current oracle and voxelizer,
plus the properties each fixture must have regardless of who computed it."""
import os

import numpy as np
import pytest

import simlab_amd as sa
from oracle import pyoracle as O

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(G, name))


C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0], [-1, -1, 0],
              [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1],
              [0, -1, 1]], np.float64)


def test_rounding_kats():
    k = load("kat_rounding.npz")
    for x, key, seed, out, h in zip(k["sround_x"], k["sround_key"], k["sround_seed"], k["sround_out"], k["sround_f16"]):
        got = np.float32(O.sround_ref(x, key, seed))
        assert got.view(np.uint32) == out.view(np.uint32)
        assert O.f16_rne(got) == h
        # sround moves x by less than one half-precision ulp
        e = max(int(np.floor(np.log2(abs(float(x))))) if x != 0 else -14, -14)
        assert abs(float(got) - float(x)) < 2.0 ** (e - 10) or abs(float(x)) >= 65000
    for c, s, hsh, d in zip(k["cell"], k["cell_seed"], k["cell_hash"], k["dither16"]):
        assert O.cell_hash(c, s) == hsh
        assert np.array_equal(O.dither16(c, s), d)
        # dither words are the hash, then xorshift32 steps (2 x 16 bits each)
        assert d[0] == (hsh & 0xFFFF) and d[1] == (hsh >> 16)
    for x, d, e16, eb in zip(k["enc_x"], k["enc_d"], k["enc_f16"], k["enc_bf16"]):
        assert O.enc_f16_sr(x, d) == e16 and O.enc_bf16_sr(x, d) == eb
        # the stored value brackets x: truncation toward zero of x plus a dither
        v16 = float(np.uint16(e16).view(np.float16)) / 32768.0
        assert abs(v16 - float(x)) <= abs(float(x)) * 2.0 ** -10 + 2.0 ** -39
    for q, b in zip(k["host_q"], k["host_q_bits"]):
        assert O.f16_bits_host(q) == b
    assert k["host_q_bits"][0] == 0x3800    # q = 1/2 is the halfway marker


def test_aa_maps_fixture():
    k = load("aa_maps_4x3x5.npz")
    for p in (0, 1):
        ld, st = O.aa_maps((4, 3, 5), p)
        assert np.array_equal(ld.astype(np.uint32), k[f"load{p}"])
        assert np.array_equal(st.astype(np.uint32), k[f"store{p}"])
    # population i that cell n stores in one step is what cell n + c_i loads
    # as its input i in the next step (the other parity), periodically
    nx, ny, nz = 4, 3, 5
    n = np.arange(nx * ny * nz)
    x, y, z = n % nx, (n // nx) % ny, n // (nx * ny)
    for p in (0, 1):
        for i in range(19):
            t = (x + int(C[i, 0])) % nx + nx * ((y + int(C[i, 1])) % ny + ny * ((z + int(C[i, 2])) % nz))
            assert np.array_equal(k[f"store{p}"][n, i], k[f"load{1 - p}"][t, i])


def test_collision_kats():
    k = load("kat_collision.npz")
    for fin, tau0, cs, g, fout in zip(k["fin"], k["tau0"], k["cs"], k["grav"], k["fout"]):
        got = O.collide_f64(fin, tau0, cs, g)
        assert np.array_equal(got, fout)
        # mass conserved; momentum changes by the Guo force rho * g
        rho = 1.0 + fin.sum()
        assert abs(fout.sum() - fin.sum()) < 1e-15
        assert np.allclose(C.T @ fout - C.T @ fin, rho * np.asarray(g), atol=1e-15)


def test_voxelizer_fixture():
    k = load("voxel_icosphere_r4p3.npz")
    v = sa.voxelize_static(k["verts"], k["tris"], (13, 12, 12))
    assert np.array_equal(v.solid, k["solid"])
    assert np.array_equal(v.walls["cell"], k["wall_cell"])
    assert np.array_equal(v.walls["mask"], k["wall_mask"])
    assert np.array_equal(v.walls["q"], k["wall_q"])
    # solid volume close to 4/3 pi r^3 (voxel_stats, the reference's 3 % bound)
    assert abs(len(k["solid"]) / (4.0 / 3.0 * np.pi * 4.3 ** 3) - 1.0) < 0.08


@pytest.mark.parametrize("name,store", [("fp64", O.F64), ("fp32", O.F32), ("fp16", O.F16), ("bf16", O.BF16)])
def test_trajectory_fixture(name, store):
    from golden.make_golden import traj_case
    st = {"fp64": sa.Storage.FP64, "fp32": sa.Storage.FP32, "fp16": sa.Storage.FP16, "bf16": sa.Storage.BF16}[name]
    setup, geom, _ = traj_case(st)
    k = load("traj_sphere_12x10x8.npz")
    ora = O.OracleLattice(setup, geom, store=store, threads=1)
    ft = ora.run_still(10)
    assert np.array_equal(ora.flags(), k["flags"])
    assert np.array_equal(ora.get_pops(), k[f"{name}_pops"])
    assert np.array_equal(ft, k[f"{name}_ft"])
    assert np.array_equal(ora.records(), k[f"{name}_records"])
