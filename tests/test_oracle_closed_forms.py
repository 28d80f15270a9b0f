"""Closed-form pins of the oracle's Smagorinsky closure and face rules
(crates/cfd/src/shader.rs:815-837, 630-687), stated as the physics they
implement rather than as the code's formula:

- Smagorinsky: the relaxation time a collision applies to the
  non-equilibrium stress holds an eddy viscosity nu_t = (tau - tau0) / 3
  equal to C_s^2 |S|, with the strain rate read from the stress
  (S = -Pi / (2 rho cs^2 tau));
- faces: after the face update, a face cell's 19 populations carry the
  density, velocity and non-equilibrium stress the boundary prescribes:
  open: rho = 1, u and Pi from the inner neighbour; velocity: u = u_inf,
  rho and Pi from the inner neighbour; total-pressure inlet: u and Pi from
  the inner neighbour and rho (cs^2 + u^2/2) = cs^2 + u_inf^2/2."""
import numpy as np
import pytest

import simlab_amd as sa
from oracle import pyoracle as O

import closed_forms as CF


@pytest.mark.parametrize("tau0,cs", [(0.5048, 0.1), (0.524, 0.16), (0.6, 0.3)])
def test_smagorinsky_tau_is_cs2_times_strain_rate(tau0, cs):
    rng = np.random.default_rng(11)
    fin = CF.random_states(64, rng, pi_amp=5e-3)
    fout = np.array([O.collide_f64(f, tau0, cs) for f in fin])
    rho, u, pi_in = CF.moments(fin)
    rho2, u2, pi_out = CF.moments(fout)
    assert np.allclose(rho2, rho, rtol=0, atol=1e-14) and np.allclose(u2, u, rtol=0, atol=1e-14)
    tau, spread = CF.relaxation_time(pi_in, pi_out)
    assert spread.max() < 1e-9
    assert np.all(tau > tau0)
    res = CF.smagorinsky_tau(tau, tau0, cs, rho, pi_in)
    assert np.abs(res).max() < 1e-9, res


def test_smagorinsky_off_relaxes_with_tau0():
    rng = np.random.default_rng(12)
    fin = CF.random_states(16, rng, pi_amp=5e-3)
    fout = np.array([O.collide_f64(f, 0.55, 0.0) for f in fin])
    tau, _ = CF.relaxation_time(CF.moments(fin)[2], CF.moments(fout)[2])
    assert np.allclose(tau, 0.55, rtol=1e-10)


def face_case(kind, periodic_yz, store):
    dims = (10, 6, 5)
    P, V, Op = sa.FaceBc.PERIODIC, sa.FaceBc.VELOCITY, sa.FaceBc.OPEN
    side = P if periodic_yz else Op
    faces = (V, Op, side, side, side, side)
    setup = sa.LatticeSetup(dims=dims, faces=faces, nu=0.02, u_inf=(0.05, 0.012, -0.02), u_init=(0.05, 0.0, 0.0),
                            total_pressure_inlet=(kind == "total_pressure"), storage=sa.Storage.FP64)
    return setup


def face_cells(setup):
    """Face cells with their inner neighbour (clamped on the non-periodic
    axes) and their type: velocity if any face they lie on is a velocity
    face (lattice.rs:295-302), open otherwise."""
    nx, ny, nz = setup.dims
    per = setup.periodic()
    out = []
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                c = (x, y, z)
                on = [(not per[a]) and (c[a] == 0 or c[a] == setup.dims[a] - 1) for a in range(3)]
                if not any(on):
                    continue
                vel = False
                for a in range(3):
                    if on[a]:
                        f = setup.faces[2 * a + (0 if c[a] == 0 else 1)]
                        vel = vel or f == sa.FaceBc.VELOCITY
                m = [min(max(c[a], 1), setup.dims[a] - 2) if not per[a] else c[a] for a in range(3)]
                out.append(((z * ny + y) * nx + x, (m[2] * ny + m[1]) * nx + m[0], vel))
    return out


@pytest.mark.parametrize("kind", ["velocity", "total_pressure"])
@pytest.mark.parametrize("periodic_yz", [True, False])
def test_face_cells_carry_the_prescribed_moments(kind, periodic_yz):
    setup = face_case(kind, periodic_yz, O.F64)
    n = setup.cells()
    lat = O.OracleLattice(setup, None, store=O.F64)
    rng = np.random.default_rng(5)
    st = CF.random_states(n, rng, pi_amp=3e-3)
    lat.set_pops(st.T.ravel())          # slot i of cell n at i * cells + n (parity-0 inputs)
    lat.run_still(1)                     # a parity-0 step: outputs of cell n at (n, opp i)
    lat.face_update()                    # the next step's face inputs: face cell b at (b, opp i)
    pops = lat.get_pops().reshape(19, n)
    post = pops[CF.OPP19, :].T           # f_i of every cell after the step / face update
    fc = face_cells(setup)
    b = np.array([t[0] for t in fc])
    m = np.array([t[1] for t in fc])
    vel = np.array([t[2] for t in fc])
    rho_b, u_b, pi_b = CF.moments(post[b])
    rho_m, u_m, pi_m = CF.moments(post[m])
    for sel, k in ((vel, kind), (~vel, "open")):
        if not sel.any():
            continue
        er, eu = CF.face_expected(k, rho_m[sel], u_m[sel], setup.u_inf)
        assert np.abs(rho_b[sel] - er).max() < 1e-12, k
        assert np.abs(u_b[sel] - eu).max() < 1e-12, k
        assert np.abs(pi_b[sel] - pi_m[sel]).max() < 1e-12, k
        # the inner states are not trivial: the rule is what makes them agree
        assert np.abs(pi_m[sel]).max() > 1e-4 and np.abs(rho_m[sel] - 1).max() > 1e-3
    if kind == "total_pressure":
        # the far field's total pressure at the inlet
        tp = rho_b[vel] * (CF.CS2 + 0.5 * (u_b[vel] ** 2).sum(-1))
        assert np.allclose(tp, CF.CS2 + 0.5 * np.dot(setup.u_inf, setup.u_inf), rtol=0, atol=1e-13)
