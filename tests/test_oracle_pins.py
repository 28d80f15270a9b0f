"""Pins the CPU oracle (oracle/lbm_oracle.cpp) before it is trusted as the
checker: the reference's own lattice tests restated (crates/cfd/tests/
channel.rs, momentum.rs), closed forms and invariants."""
import math

import numpy as np
import pytest

import simlab_amd as sa
from simlab_amd import workloads as W
from oracle import pyoracle as O

W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0], [-1, -1, 0],
              [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0], [-1, 1, 0], [1, 0, -1], [-1, 0, 1], [0, 1, -1],
              [0, -1, 1]])
OPP = [0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17]


def feq_shift(rho, u):
    cu = C @ np.asarray(u, float)
    usq = 1.5 * float(np.dot(u, u))
    return W19 * ((rho - 1.0) + rho * (3 * cu + 4.5 * cu * cu - usq))


def poiseuille(h, store, umax=0.02):
    setup, geom, meta = W.poiseuille(h, umax=umax)
    lat = O.OracleLattice(setup, geom, store=store)
    hist = lat.run_still(meta["steps"])
    return W.poiseuille_error(meta, lat.macros(), hist[-1])


def test_poiseuille_second_order():
    # channel.rs:75-89
    errs = []
    for h in (8.0, 16.0, 32.0):
        e, bal = poiseuille(h, O.F32)
        assert abs(bal) < 2e-3
        errs.append(e)
    o2 = math.log2(errs[1] / errs[2])
    assert errs[2] < 2e-3
    assert o2 > 1.6 or errs[2] < 1e-4


@pytest.mark.parametrize("store", [O.F16, O.F16REF, O.BF16])
def test_poiseuille_stochastic_rounding(store):
    # channel.rs:91-101 (FP16 < 3e-3); BF16 has 3 fewer mantissa bits than
    # FP16 at the 2^15 scale, its bar is stated separately (1e-2)
    bar = 1e-2 if store == O.BF16 else 3e-3
    for umax in (0.02, 0.06):
        e, bal = poiseuille(16.0, store, umax)
        assert e < bar, (store, umax, e)


def test_fp64_oracle_poiseuille_matches_fp32():
    e32, _ = poiseuille(16.0, O.F32)
    e64, b64 = poiseuille(16.0, O.F64)
    assert abs(b64) < 1e-6
    assert abs(e64 - e32) < 2e-4


def test_momentum_lost_by_fluid_equals_force_on_body():
    # momentum.rs:39-51
    n = 40
    setup, geom = W.ball_box(n, 6.0, 0.05, 0.05, 0.0, [sa.FaceBc.PERIODIC] * 6)
    lat = O.OracleLattice(setup, geom, store=O.F32)

    def pmom():
        m = lat.macros()
        ok = ~np.isnan(m[:, 0])
        return float(np.sum((m[ok, 0] * m[ok, 1]).astype(np.float64)))

    p0 = pmom()
    h = lat.run_still(600)
    p1 = pmom()
    impulse = float(h[:, 0].sum())
    err = (p0 - p1 - impulse) / impulse
    assert impulse > 0
    assert abs(err) < 5e-3


def test_uniform_states_feel_no_force():
    # momentum.rs:73-85
    setup, geom = W.ball_box(32, 5.0, 0.05, 0.0, 0.0, [sa.FaceBc.OPEN] * 6)
    lat = O.OracleLattice(setup, geom, store=O.F32)
    h = lat.run_still(50)
    assert np.all(np.abs(h) < 1e-4)


def test_tau_one_gives_equilibrium():
    # closed form: with tau = 1 (nu = 1/6) and Cs = 0 the regularized
    # collision returns feq(rho, u) exactly
    rng = np.random.default_rng(1)
    u = np.array([0.03, -0.02, 0.01])
    fin = feq_shift(1.02, u) + 1e-3 * rng.standard_normal(19)
    out = O.collide_f64(fin, 1.0)
    rho1 = fin.sum()
    j = C.T @ fin
    uu = j / (1 + rho1)
    assert np.allclose(out, feq_shift(1 + rho1, uu), atol=1e-15)


def test_collision_conserves_mass_and_momentum_and_relaxes_stress():
    rng = np.random.default_rng(2)
    u = np.array([0.05, 0.01, -0.02])
    fin = feq_shift(0.99, u) + 2e-3 * rng.standard_normal(19)
    for tau0, cs in ((0.52, 0.0), (0.6, 0.1), (0.9, 0.16)):
        out = O.collide_f64(fin, tau0, cs)
        assert abs(out.sum() - fin.sum()) < 1e-15
        assert np.allclose(C.T @ out, C.T @ fin, atol=1e-15)
        # non-equilibrium stress scales by (1 - 1/tau) with tau >= tau0
        rho = 1 + fin.sum()
        uu = (C.T @ fin) / rho

        def pneq(f):
            s = np.einsum("i,ia,ib->ab", f, C, C)
            return s - rho * np.outer(uu, uu) - fin.sum() / 3 * np.eye(3)

        r = np.linalg.norm(pneq(out)) / np.linalg.norm(pneq(fin))
        assert r <= abs(1 - 1 / tau0) + 1e-12
        if cs == 0:
            assert np.allclose(pneq(out), (1 - 1 / tau0) * pneq(fin), rtol=1e-9, atol=1e-17)


def test_guo_forcing_adds_the_body_force_momentum():
    # with Guo forcing the post-collision momentum gains exactly G = rho g
    fin = feq_shift(1.0, [0.02, 0.0, 0.0])
    g = np.array([1e-5, 2e-6, -3e-6])
    out = O.collide_f64(fin, 0.8, 0.0, g)
    rho = 1 + fin.sum()
    assert np.allclose(C.T @ out - C.T @ fin, rho * g, rtol=1e-9, atol=1e-18)
    assert abs(out.sum() - fin.sum()) < 1e-15


def test_aa_maps_stream_each_population_once():
    # a parity-p step stores f_i of cell n where the next (1-p) step loads
    # f_i of cell n + c_i (periodic): AA streaming is a permutation
    dims = (4, 3, 5)
    nx, ny, nz = dims
    n = nx * ny * nz
    for p in (0, 1):
        _, st = O.aa_maps(dims, p)
        ld, _ = O.aa_maps(dims, 1 - p)
        assert len(set(st.ravel().tolist())) == 19 * n
        for cell in range(n):
            x, y, z = cell % nx, (cell // nx) % ny, cell // (nx * ny)
            for i in range(19):
                t = ((x + C[i][0]) % nx) + nx * (((y + C[i][1]) % ny) + ny * ((z + C[i][2]) % nz))
                assert st[cell, i] == ld[t, i]


def test_f16_host_rounding_and_conversions():
    # lattice.rs:172-189: q rounding adds bit 12, subnormals flush to zero
    assert O.f16_bits_host(0.5) == 0x3800
    assert O.f16_bits_host(1.0) == 0x3C00
    assert O.f16_bits_host(0.001) == 0x1419
    assert O.f16_bits_host(1e-6) == 0
    assert O.f16_bits_host(-0.25) == 0xB400
    for x in (0.1, 0.37, 0.999, 1e-3, 12.5, 65000.0):
        h = O.f16_rne(x)
        assert O.half_to_float(h) == pytest.approx(np.float16(x), rel=0, abs=0)
        assert O.half_to_float(O.f16_rtz(x)) <= x
    assert O.f16_rne(2049.0) == np.float16(2049.0).view(np.uint16)
    # round toward zero saturates finite overflow at the largest half (the
    # behaviour of v_cvt_pkrtz_f16_f32), keeps infinities and NaN
    assert O.f16_rtz(65519.0) == 0x7BFF and O.f16_rtz(1e30) == 0x7BFF and O.f16_rtz(-7e4) == 0xFBFF
    assert O.f16_rtz(float("inf")) == 0x7C00 and (O.f16_rtz(float("nan")) & 0x7FFF) == 0x7E00
    assert O.f16_rtz(65503.9) == 0x7BFE and O.f16_rtz(6e-8) == 0x0001
    # the storage encoder: scaled by 2^15, dithered, saturating
    assert O.enc_f16_sr(3.0, 0xFFFF) == 0x7BFF and O.enc_f16_sr(-3.0, 0) == 0xFBFF
    assert O.enc_f16_sr(0.5, 0) == np.float16(16384.0).view(np.uint16)


def test_sround_is_unbiased_and_within_one_ulp():
    # shader.rs:510-517: E[sround(x)] = x; |sround(x) - x| < 1 ulp
    x = 1234.5678
    ulp = 2.0 ** (math.floor(math.log2(x)) - 10)
    vals = np.array([O.half_to_float(O.f16_rne(O.sround_ref(x, k, 77))) for k in range(4000)])
    assert np.all(np.abs(vals - x) <= ulp)
    assert abs(vals.mean() - x) < 0.05 * ulp
    d = np.array([O.half_to_float(O.enc_f16_sr(x / 32768.0, int(k))) for k in range(0, 65536, 13)])
    assert np.all(np.abs(d - x) <= ulp)
    assert abs(d.mean() - x) < 0.01 * ulp


def test_face_update_one_step_early_is_the_same_schedule():
    # The GPU library rebuilds the face cells' inputs for step k+1 right
    # after step k (from the source cells' fresh outputs) instead of at the
    # start of step k+1, and at the start of each run does this step's face
    # update before the next step's. Run in the oracle, that schedule gives
    # bit-identical results to the reference's order.
    setup, geom, meta = W.sphere_channel((40, 20, 16), 3.5, 12.0, 40.0, 0.05, 0.0, sa.Storage.FP64, debug_dump=True,
                                         subdiv=3)
    a = O.OracleLattice(setup, geom, store=O.F64)
    b = O.OracleLattice(setup, geom, store=O.F64)
    b.set_fused(True)
    ha = a.run_still(60)
    hb = np.concatenate([b.run_still(25), b.run_still(35)])
    assert np.array_equal(ha, hb)
    assert np.array_equal(a.macros_f64(), b.macros_f64(), equal_nan=True)
