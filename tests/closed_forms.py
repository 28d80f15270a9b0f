"""Closed forms that pin the collision closure and the face rules without
restating the code under test: moments of 19 stored deviations, the
Smagorinsky eddy viscosity from the strain rate, and the three face rules
stated as what the face cell's state must hold (crates/cfd/src/shader.rs:
630-687, 815-837). Shared by the oracle tests and the GPU tests."""
import numpy as np

W19 = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
C19 = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [1, 1, 0],
                [-1, -1, 0], [1, 0, 1], [-1, 0, -1], [0, 1, 1], [0, -1, -1], [1, -1, 0], [-1, 1, 0], [1, 0, -1],
                [-1, 0, 1], [0, 1, -1], [0, -1, 1]], dtype=np.float64)
OPP19 = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17])
CS2 = 1.0 / 3.0   # lattice speed of sound squared


def moments(ft):
    """Density, velocity and non-equilibrium stress of populations given as
    deviations f~ = f - w (last axis: the 19 directions):
    rho = sum f, rho u = sum c f, Pi = sum c c f - rho u u - rho cs^2 I."""
    ft = np.asarray(ft, np.float64)
    f = ft + W19
    rho = f.sum(-1)
    u = (f @ C19) / rho[..., None]
    s = np.einsum("...i,ia,ib->...ab", f, C19, C19)
    pi = s - rho[..., None, None] * (u[..., :, None] * u[..., None, :]) - (rho * CS2)[..., None, None] * np.eye(3)
    return rho, u, pi


def random_states(n, rng, u_amp=0.06, rho_amp=0.02, pi_amp=2e-3):
    """n near-equilibrium populations (deviations) with random density,
    velocity and a random symmetric non-equilibrium part."""
    rho = 1.0 + rho_amp * rng.uniform(-1, 1, n)
    u = u_amp * rng.uniform(-1, 1, (n, 3))
    cu = u @ C19.T
    usq = 1.5 * (u * u).sum(1)
    feq = W19 * rho[:, None] * (1 + 3 * cu + 4.5 * cu * cu - usq[:, None])
    return feq - W19 + pi_amp * rng.standard_normal((n, 19)) * W19 * 18


def smagorinsky_tau(tau, tau0, cs, rho, pi_in):
    """Residual of the Smagorinsky closure for a collision that relaxed the
    non-equilibrium stress pi_in with the time `tau`: the eddy viscosity
    nu_t = (tau - tau0) cs_lattice^2 must equal (C_s Delta)^2 |S| with
    |S| = sqrt(2 S:S) and the strain rate read from the stress,
    S = -Pi / (2 rho cs_lattice^2 tau) (Hou et al. 1996, Delta = dt = 1).
    Returns nu_t / ((C_s)^2 |S|) - 1."""
    s = -pi_in / (2.0 * rho[..., None, None] * CS2 * tau[..., None, None])
    smag = np.sqrt(2.0 * np.einsum("...ab,...ab->...", s, s))
    nu_t = (tau - tau0) * CS2
    return nu_t / (cs * cs * smag) - 1.0


def relaxation_time(pi_in, pi_out):
    """tau from a regularized BGK collision, Pi_out = (1 - 1/tau) Pi_in, and
    the largest spread of the per-component ratio (0 for a collision that
    relaxes every component alike)."""
    a = pi_in.reshape(pi_in.shape[:-2] + (9,))
    b = pi_out.reshape(pi_out.shape[:-2] + (9,))
    r = (a * b).sum(-1) / (a * a).sum(-1)
    spread = np.abs(b - r[..., None] * a).max(-1) / np.abs(a).max(-1)
    return 1.0 / (1.0 - r), spread


def face_expected(kind, rho_in, u_in, u_inf):
    """Density and velocity a face cell takes from its inner neighbour's
    (rho_in, u_in) (shader.rs:670-681): an open face holds the far-field
    pressure (rho = 1) with the velocity from inside; a velocity face holds
    u_inf with the density from inside; a total-pressure inlet keeps the
    velocity from inside and the far field's total pressure,
    rho (cs^2 + |u|^2 / 2) = cs^2 + |u_inf|^2 / 2."""
    u_inf = np.asarray(u_inf, np.float64)
    if kind == "open":
        return np.ones_like(rho_in), u_in
    if kind == "velocity":
        return rho_in, np.broadcast_to(u_inf, u_in.shape)
    if kind == "total_pressure":
        return (CS2 + 0.5 * u_inf @ u_inf) / (CS2 + 0.5 * (u_in * u_in).sum(-1)), u_in
    raise ValueError(kind)


def regularized(rho, u, pi):
    """The regularized state with these moments: f~_i = w_i rho (1 + c.u /
    cs^2 + (c.u)^2 / (2 cs^4) - u^2 / (2 cs^2)) - w_i
    + w_i / (2 cs^4) (c_a c_b - cs^2 delta_ab) Pi_ab; its zeroth, first and
    second moments are rho, rho u and rho cs^2 I + rho u u + Pi on D3Q19."""
    cu = u @ C19.T
    usq = (u * u).sum(-1)
    feq = W19 * rho[..., None] * (1 + cu / CS2 + cu * cu / (2 * CS2 * CS2) - usq[..., None] / (2 * CS2))
    q = np.einsum("ia,ib->iab", C19, C19) - CS2 * np.eye(3)
    fneq = W19 / (2 * CS2 * CS2) * np.einsum("iab,...ab->...i", q, pi)
    return feq - W19 + fneq
