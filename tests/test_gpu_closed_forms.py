"""Closed-form pins of the HIP path's Smagorinsky closure and face rules
(crates/cfd/src/shader.rs:815-837, 630-687), through the C ABI, stated as
the physics they implement (tests/closed_forms.py):

- one step of a periodic lattice from random states: every cell's collision
  relaxes its non-equilibrium stress with a tau whose eddy viscosity
  (tau - tau0) / 3 is C_s^2 |S| with S = -Pi / (2 rho cs^2 tau);
- after one step, the populations the library has written into the face
  cells for the next step are those of the regularized state with the
  prescribed density and velocity (open: rho = 1 and u from the inner
  neighbour; velocity: u = u_inf, rho from inside; total-pressure inlet:
  rho (cs^2 + u^2/2) = cs^2 + u_inf^2/2, u from inside) and the inner
  neighbour's post-collision non-equilibrium stress. Only the directions a
  colliding cell reads are written (the others are read by nobody)."""
import numpy as np
import pytest

import simlab_amd as sa

import closed_forms as CF

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def encode(v, storage):
    v = np.asarray(v, np.float64)
    if storage == sa.Storage.FP64:
        return v
    if storage == sa.Storage.FP32:
        return v.astype(np.float32)
    if storage == sa.Storage.FP16:
        return (v * 32768.0).astype(np.float16).view(np.uint16)
    raise ValueError(storage)


def decode(p, storage):
    if storage == sa.Storage.FP16:
        return p.view(np.float16).astype(np.float64) / 32768.0
    return p.astype(np.float64)


@pytest.mark.parametrize("storage,cs,tol", [(sa.Storage.FP64, 0.16, 1e-8), (sa.Storage.FP32, 0.4, 1e-3)])
def test_smagorinsky_tau_is_cs2_times_strain_rate(storage, cs, tol):
    dims = (24, 8, 6)
    setup = sa.LatticeSetup(dims=dims, nu=0.0016, smagorinsky=cs, storage=storage)
    say("case smagorinsky", storage.name, cs)
    lat = sa.Lattice(setup)
    n = setup.cells()
    rng = np.random.default_rng(3)
    st = CF.random_states(n, rng, pi_amp=5e-3)
    lat.set_pops(encode(st.T.ravel(), storage))
    fin = decode(lat.get_pops(), storage).reshape(19, n).T      # parity-0 inputs: slot (i, n)
    lat.run_still(1)
    fout = decode(lat.get_pops(), storage).reshape(19, n)[CF.OPP19, :].T   # outputs at (opp i, n)
    rho, u, pi_in = CF.moments(fin)
    rho2, u2, pi_out = CF.moments(fout)
    tau0 = 0.5 + 3 * setup.nu
    tau, spread = CF.relaxation_time(pi_in, pi_out)
    res = CF.smagorinsky_tau(tau, tau0, cs, rho, pi_in)
    say(f"  tau - tau0 {np.median(tau - tau0):.3e}, max |residual| {np.abs(res).max():.2e}, spread {spread.max():.1e}")
    assert np.abs(rho2 - rho).max() < (1e-13 if storage == sa.Storage.FP64 else 2e-6)
    assert np.all(tau > tau0)
    assert np.abs(res).max() < tol


def face_setup(kind, periodic_yz, storage):
    P, V, Op = sa.FaceBc.PERIODIC, sa.FaceBc.VELOCITY, sa.FaceBc.OPEN
    side = P if periodic_yz else Op
    return sa.LatticeSetup(dims=(40, 12, 10), faces=(V, Op, side, side, side, side), nu=0.02,
                           u_inf=(0.05, 0.012, -0.02), u_init=(0.05, 0.0, 0.0),
                           total_pressure_inlet=(kind == "total_pressure"), storage=storage)


def face_cells(setup):
    """(face cell, clamped inner neighbour, velocity type, live directions)
    for every face cell (lattice.rs:283-324)."""
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
                vel = any(on[a] and setup.faces[2 * a + (0 if c[a] == 0 else 1)] == sa.FaceBc.VELOCITY
                          for a in range(3))
                m = [min(max(c[a], 1), setup.dims[a] - 2) if not per[a] else c[a] for a in range(3)]
                t = np.asarray(c)[None, :] + CF.C19.astype(int)
                live = np.ones(19, bool)
                for a in range(3):
                    if not per[a]:
                        live &= (t[:, a] >= 1) & (t[:, a] <= setup.dims[a] - 2)
                out.append(((z * ny + y) * nx + x, (m[2] * ny + m[1]) * nx + m[0], vel, live))
    return out


@pytest.mark.parametrize("storage,tol", [(sa.Storage.FP64, 1e-12), (sa.Storage.FP32, 1e-6), (sa.Storage.FP16, 4e-4)])
@pytest.mark.parametrize("kind", ["velocity", "total_pressure"])
@pytest.mark.parametrize("periodic_yz", [True, False])
def test_face_cells_get_the_prescribed_state(kind, periodic_yz, storage, tol):
    setup = face_setup(kind, periodic_yz, storage)
    say("case faces", kind, "periodic y/z" if periodic_yz else "open y/z", storage.name)
    lat = sa.Lattice(setup)
    n = setup.cells()
    rng = np.random.default_rng(8)
    lat.set_pops(encode(CF.random_states(n, rng, pi_amp=3e-3).T.ravel(), storage))
    lat.run_still(1)
    post = decode(lat.get_pops(), storage).reshape(19, n)[CF.OPP19, :].T
    fc = face_cells(setup)
    b = np.array([t[0] for t in fc])
    m = np.array([t[1] for t in fc])
    vel = np.array([t[2] for t in fc])
    live = np.array([t[3] for t in fc])
    rho_m, u_m, pi_m = CF.moments(post[m])
    worst = 0.0
    for sel, k in ((vel, kind), (~vel, "open")):
        if not sel.any():
            continue
        er, eu = CF.face_expected(k, rho_m[sel], u_m[sel], setup.u_inf)
        want = CF.regularized(er, eu, pi_m[sel])
        got = post[b[sel]]
        d = np.abs(got - want)[live[sel]]
        assert live[sel].sum() > 0
        worst = max(worst, float(d.max()))
        assert d.max() < tol, (k, float(d.max()))
    say(f"  {len(b)} face cells, max |f - f_rule| {worst:.2e}")
